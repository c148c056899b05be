// dea_kernels.h -- gfx950 kernels of the Direct Electrifying conductance solve.
//
// Matrix layout ("blocked SSS").  The reference stores the strictly-lower triangle of the
// symmetric conductance matrix plus its diagonal (Direct_Electrifying.cpp:1329-1394) and
// applies it with a sequential scatter (spM_V_SSS, :1678-1695).  Here the rows are cut into
// blocks of RB consecutive rows, one work-group per block:
//   * a lower entry (r, c) with c in r's block is read once and applied twice: the row
//     side gathers p[c] from LDS, the transpose side adds a*p[r] into an LDS accumulator
//     (ds_add_f64 / ds_add_f32);
//   * a lower entry whose column lies in an earlier block is applied on the row side by a
//     gather, and its transpose is stored once more in a small CSR keyed by the column
//     ("remote" list urp/ucol/uval), so the owner of row c gathers it.  No global atomics,
//     every output row is written exactly once by its owner.
// Tube networks are numbered tube by tube, so nearly every entry is the in-block chain
// entry (r, r-1); the remote list holds junctions and block-crossing chain links (a few
// percent of L).
//
// The PCG iteration is two launches:
//   k_apply<PCG>:  p = M^-1 r + beta p_old (own rows into LDS and HBM; gathered columns
//                  recompute it from r, diag, p_old), q = A p, partial p.q; the last
//                  work-group to finish reduces the partials in a fixed order and sets
//                  alpha = rz / p.q.
//   k_update:      x += alpha p, r -= alpha q, z = M^-1 r, partial r.z and r.r; the last
//                  work-group sets beta, the iteration count and the convergence flag.
// Both return at once when the flag is set, so chunks of iterations are captured in a
// hipGraph and launched without a host round trip per iteration.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cicy {

constexpr int kThreads = 256;
constexpr int kWave = 64;

struct PcgState {
    double alpha, beta, rz, rr, pq;
    double rr_init;    // ||r_0||^2
    double stop2;      // squared stopping threshold on r.r
    double pad0;
    long long iters, max_iter;
    int done, status;
    unsigned int counter;
    int pad1;
};

template <typename T>
struct SssDev {
    int n, nblocks;
    const int *__restrict__ rp;
    const int *__restrict__ col;
    const T *__restrict__ val;
    const T *__restrict__ diag;   // nullptr: unit diagonal (Jacobi-scaled system)
    const int *__restrict__ urp;
    const int *__restrict__ ucol;
    const T *__restrict__ uval;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_down(v, o, kWave);
    return v;
}

// Sum over a 256-thread work-group; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds4) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) lds4[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) s = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
    return s;
}

// Two simultaneous sums; results valid in thread 0.
__device__ __forceinline__ void block_sum2(double &a, double &b, double *lds8) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
        lds8[w] = a;
        lds8[4 + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = (lds8[0] + lds8[1]) + (lds8[2] + lds8[3]);
        b = (lds8[4] + lds8[5]) + (lds8[6] + lds8[7]);
    }
}

// Publishes this work-group's partial(s) and returns true in every thread of the last
// work-group to arrive, after which all partials are visible to it.
__device__ __forceinline__ bool arrive_last(unsigned int *counter, int *lds_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned int v = atomicAdd(counter, 1u);
        *lds_flag = (v == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    bool last = *lds_flag != 0;
    if (last) __threadfence();
    return last;
}

// Fixed-order sum of nb partials by one work-group (deterministic); valid in thread 0.
__device__ __forceinline__ double reduce_partials(const double *part, int nb, int stride,
                                                  double *lds4) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += kThreads) s += part[(size_t)i * stride];
    return block_sum(s, lds4);
}

// Jacobi-preconditioned search direction p = (1/d) r + beta p_old.  One definition so
// that the owner of a row and every gather of it compute the same bits.
template <typename T>
__device__ __forceinline__ T pdir(T r, T d, T beta, T pold) {
    T z = ((T)1 / d) * r;
    return fma(beta, pold, z);
}

template <typename T>
__device__ __forceinline__ T pdir_unit(T r, T beta, T pold) {
    return fma(beta, pold, r);
}

#ifdef __HIPCC__
// q = A p over one row block; the guard, when given, skips the launch unless *guard != 0.
// PCG: p = M^-1 r + beta p_old is formed on the fly (own rows are also written to pnew)
// and the block's p.q partial is reduced into alpha by the last work-group.
// SPMV: p = src.  UNIT: unit diagonal (Jacobi-scaled system), M^-1 = I.
template <typename T, int RB, bool PCG, bool UNIT>
__global__ __launch_bounds__(kThreads) void k_apply(SssDev<T> s, const T *__restrict__ src,
                                                    const T *__restrict__ rvec,
                                                    T *__restrict__ pnew, T *__restrict__ q,
                                                    PcgState *__restrict__ st,
                                                    double *__restrict__ partials,
                                                    const int *__restrict__ guard) {
    if (PCG && st->done) return;
    if (guard && !*guard) return;
    __shared__ T sp[RB];
    __shared__ T sy[RB];
    __shared__ double red[8];
    __shared__ int flag;
    const int r0 = blockIdx.x * RB;
    const int nrows = min(RB, s.n - r0);
    const T beta = PCG ? (T)st->beta : (T)0;
    auto pval = [&](int i) -> T {
        if (!PCG) return src[i];
        if (UNIT) return pdir_unit<T>(rvec[i], beta, src[i]);
        return pdir<T>(rvec[i], s.diag[i], beta, src[i]);
    };
    for (int t = threadIdx.x; t < nrows; t += kThreads) {
        const int row = r0 + t;
        const T p = pval(row);
        sp[t] = p;
        if (PCG) pnew[row] = p;
        sy[t] = UNIT ? p : s.diag[row] * p;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nrows; t += kThreads) {
        const int row = r0 + t;
        const T pr = sp[t];
        T acc = (T)0;
        const int j1 = s.rp[row + 1];
        for (int j = s.rp[row]; j < j1; j++) {
            const int c = s.col[j];
            const T a = s.val[j];
            if (c >= r0) {
                acc = fma(a, sp[c - r0], acc);
                atomicAdd(&sy[c - r0], a * pr);
            } else {
                acc = fma(a, pval(c), acc);
            }
        }
        const int u1 = s.urp[row + 1];
        for (int j = s.urp[row]; j < u1; j++) acc = fma(s.uval[j], pval(s.ucol[j]), acc);
        atomicAdd(&sy[t], acc);
    }
    __syncthreads();
    double part = 0.0;
    for (int t = threadIdx.x; t < nrows; t += kThreads) {
        const T qv = sy[t];
        q[r0 + t] = qv;
        if (PCG) part += (double)sp[t] * (double)qv;
    }
    if (!PCG) return;
    part = block_sum(part, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = part;
    if (arrive_last(&st->counter, &flag)) {
        double pq = reduce_partials(partials, gridDim.x, 1, red);
        if (threadIdx.x == 0) {
            double alpha = st->rz / pq;
            st->pq = pq;
            st->alpha = alpha;
            st->counter = 0;
            if (!(pq > 0.0) || !isfinite(alpha)) {
                st->status = CICY_BREAKDOWN;
                st->done = 1;
            }
        }
    }
}

#endif

}  // namespace cicy
