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
//                  work-group of the system to finish reduces the partials in a fixed
//                  order and sets alpha = rz / p.q.
//   k_update:      x += alpha p, r -= alpha q, z = M^-1 r, partial r.z and r.r; the last
//                  work-group sets beta, the iteration count and the convergence flag.
// Both return at once when the flag is set, so chunks of iterations are captured in a
// hipGraph and launched without a host round trip per iteration.
//
// Batches (BATCH=true): many independent systems are stored one after another (block-
// diagonal), every row block belongs to one system, and a block table maps a work-group
// to (system, first row).  Each system has its own PcgState (alpha, beta, flags) and
// reduction counter, so systems of different size converge independently inside one
// launch; an optional `active` list restricts a launch to the blocks of unfinished systems.
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

// Mixed-precision state (one per system).
struct MixState {
    double rr_prev;     // scaled recursive r.r before the last update
    double m2;          // scaled r.r right after the last reliable update
    double tgt2;        // scaled r.r at which the stopping rule is expected to hold
    double rtrue2;      // ||b - A x||^2 at the last reliable update
    double delta2;      // update when r.r <= delta2 * m2
    int need_update;
    int refinements;
    unsigned int counter;
    int stall;
    int restarts;
    int run;            // FP32 iterations since the last reliable update
};

// FP32 iterations after which a reliable update is taken even if the recursion has not
// reached its trigger (a recursion stuck at FP32's resolution then restarts from the true
// residual, or is found stalled, instead of running on into overflow)
constexpr int kMixedMaxRun = 4096;

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

// "Band + remainder" layout of the same SSS matrix: band[r] = a(r, r-1) (0 when row r
// has no entry in column r-1; band[0] = band[n] = 0) plus every other stored entry as two
// half-entries (nbr, val), one in each of its rows, sorted by row and indexed per row by
// xrp (per row block by xblk).
template <typename T>
struct BandDev {
    int n, nblocks;
    const T *__restrict__ band;   // [n+1]
    const T *__restrict__ diag;   // nullptr: unit diagonal
    const int *__restrict__ xblk; // [nblocks+1] first half-entry of each row block
    const int *__restrict__ xrp;  // [n+1] first half-entry of each row
    const int *__restrict__ xnbr; // [m]
    const T *__restrict__ xval;   // [m]
    // [nblocks] row block of each work-group (heaviest blocks first); nullptr: row order
    const int *__restrict__ order = nullptr;
};

struct BatchDev {
    const int *active;    // [grid] block ids of this launch (nullptr: blockIdx.x)
    const int *blk_sys;   // [blocks] system of each block
    const int *blk_row0;  // [blocks] first (global) row of each block
    const int *sys_row0;  // [nsys+1] row offsets
    const int *sys_blk0;  // [nsys+1] block offsets
};

struct Blk {
    int b, sys, r0, nrows, first, count;
};

template <int RB, bool BATCH>
__device__ __forceinline__ Blk block_info(const BatchDev &bd, int n) {
    Blk k;
    if (!BATCH) {
        k.b = blockIdx.x;
        k.sys = 0;
        k.r0 = k.b * RB;
        k.nrows = min(RB, n - k.r0);
        k.first = 0;
        k.count = gridDim.x;
    } else {
        k.b = bd.active ? bd.active[blockIdx.x] : (int)blockIdx.x;
        k.sys = bd.blk_sys[k.b];
        k.r0 = bd.blk_row0[k.b];
        k.nrows = min(RB, bd.sys_row0[k.sys + 1] - k.r0);
        k.first = bd.sys_blk0[k.sys];
        k.count = bd.sys_blk0[k.sys + 1] - k.first;
    }
    return k;
}

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

// Reduction scratch: one slot of NV partial sums per work-group (gpart), one per group of
// kGroup consecutive work-groups of a system (part, at the slot of the group's first
// work-group), and the group arrival counters (gcount, same indexing).
constexpr int kGroup = 64;
struct RedBuf {
    double *part;            // [2 * blocks] group sums
    double *gpart;           // [2 * blocks] work-group sums
    unsigned int *gcount;    // [blocks]     group arrival counters
};

__device__ __forceinline__ double ld_dev(const double *p) {
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduction of NV values over the work-groups of one system, in a fixed order and without
// cache-wide fences.  Thread 0 stores the work-group's sums into its slot with
// device-coherent (agent-scope) stores, waits for them to complete, and arrives on its
// group's counter.  The last arriver of a group sums the group's slots in work-group order
// (lane j of wave 0 reads slot j, then a fixed shuffle tree), stores the group sum, and
// arrives on `top`; the last group sums the group sums in group order.  Each partial is
// summed in the same order whichever work-group arrives last, so the totals are bit-identical
// from run to run.  Stores and loads of shared slots are performed at the point of
// coherence (the eight XCDs' L2 caches are not coherent with each other), so no L2
// write-back/invalidate fence is needed.  Returns true in the final work-group with out[]
// valid in thread 0.
template <int NV>
__device__ __forceinline__ bool reduce_sys(double (&v)[NV], const Blk &k, const RedBuf &rb,
                                           unsigned int *top, double *red, int *flag,
                                           double (&out)[NV]) {
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) s[i] = block_sum(v[i], red);
    const int g0 = k.first + ((k.b - k.first) / kGroup) * kGroup;
    const int gsz = min(kGroup, k.first + k.count - g0);
    const int ng = (k.count + kGroup - 1) / kGroup;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) st_dev(&rb.gpart[(size_t)k.b * NV + i], s[i]);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        *flag = atomicAdd(&rb.gcount[g0], 1u) == (unsigned)(gsz - 1) ? 1 : 0;
    }
    __syncthreads();
    const int group_last = *flag;
    __syncthreads();  // every thread has read the flag before it is rewritten below
    if (!group_last) return false;
    if (threadIdx.x < kWave) {
        const int lane = threadIdx.x;
        double g[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) {
            g[i] = lane < gsz ? ld_dev(&rb.gpart[(size_t)(g0 + lane) * NV + i]) : 0.0;
            g[i] = wave_sum(g[i]);
        }
        int last = 0;
        if (lane == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) st_dev(&rb.part[(size_t)g0 * NV + i], g[i]);
            __hip_atomic_store(&rb.gcount[g0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            last = atomicAdd(top, 1u) == (unsigned)(ng - 1) ? 1 : 0;
        }
        last = __shfl(last, 0, kWave);
        if (last) {
            double t[NV];
#pragma unroll
            for (int i = 0; i < NV; i++) t[i] = 0.0;
            for (int gi = lane; gi < ng; gi += kWave) {
#pragma unroll
                for (int i = 0; i < NV; i++)
                    t[i] += ld_dev(&rb.part[(size_t)(k.first + gi * kGroup) * NV + i]);
            }
#pragma unroll
            for (int i = 0; i < NV; i++) t[i] = wave_sum(t[i]);
            if (lane == 0) {
#pragma unroll
                for (int i = 0; i < NV; i++) out[i] = t[i];
                __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) *flag = last;
    }
    __syncthreads();
    return *flag != 0;
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

// q = A p over one row block.
// PCG: p = M^-1 r + beta p_old is formed on the fly (own rows are also written to pnew)
// and the block's p.q partial is reduced into alpha by the system's last work-group.
// SPMV: p = src.  UNIT: unit diagonal (Jacobi-scaled system), M^-1 = I.
// guard: when given, the launch does nothing for systems whose need_update flag is 0.
template <typename T, int RB, bool PCG, bool UNIT, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_apply(SssDev<T> s, BatchDev bd,
                                                    const T *__restrict__ src,
                                                    const T *__restrict__ rvec,
                                                    T *__restrict__ pnew, T *__restrict__ q,
                                                    PcgState *__restrict__ st_base,
                                                    RedBuf rbuf,
                                                    const MixState *__restrict__ guard) {
    const Blk k = block_info<RB, BATCH>(bd, s.n);
    PcgState *st = st_base ? st_base + k.sys : nullptr;
    if (PCG && st->done) return;
    if (guard && !guard[k.sys].need_update) return;
    __shared__ T sp[RB];
    __shared__ T sy[RB];
    __shared__ double red[8];
    __shared__ int flag;
    const int r0 = k.r0;
    const int nrows = k.nrows;
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
    double v[1] = {part}, out[1];
    if (reduce_sys<1>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double pq = out[0];
        if (threadIdx.x == 0) {
            double alpha = st->rz / pq;
            st->pq = pq;
            st->alpha = alpha;
            if (!(pq > 0.0) || !isfinite(alpha)) {
                st->status = 2;  // CICY_BREAKDOWN
                st->done = 1;
            }
        }
    }
}

// q = A p over one row block in the band + remainder layout.  The band part is a
// three-point stencil on LDS (p[r-1], p[r], p[r+1] with a halo row on each side, band[r],
// band[r+1]); each row's remainder half-entries (xrp[r]..xrp[r+1]) are then added by the
// thread that owns the row, in stored order, four loads in flight.  Every sum has a fixed
// order, so the product is bit-identical from run to run.  Same PCG/SPMV/UNIT/BATCH/guard
// semantics as k_apply.
template <typename T, int RB, bool PCG, bool UNIT, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_apply_band(BandDev<T> s, BatchDev bd,
                                                         const T *__restrict__ src,
                                                         const T *__restrict__ rvec,
                                                         T *__restrict__ pnew,
                                                         T *__restrict__ q,
                                                         PcgState *__restrict__ st_base,
                                                         RedBuf rbuf,
                                                         const MixState *__restrict__ guard) {
    Blk k = block_info<RB, BATCH>(bd, s.n);
    if (!BATCH && s.order) {
        // blocks with many off-band entries (plate rows) go first, so that they do not
        // run last on a part of the GPU while the rest of it idles
        k.b = s.order[blockIdx.x];
        k.r0 = k.b * RB;
        k.nrows = min(RB, s.n - k.r0);
    }
    PcgState *st = st_base ? st_base + k.sys : nullptr;
    if (PCG && st->done) return;
    if (guard && !guard[k.sys].need_update) return;
    __shared__ T sp[RB + 2];
    __shared__ T sb[RB + 1];
    __shared__ double red[8];
    __shared__ int flag;
    const int r0 = k.r0;
    const int nrows = k.nrows;
    const int lo = BATCH ? bd.sys_row0[k.sys] : 0;
    const int hi = BATCH ? bd.sys_row0[k.sys + 1] : s.n;
    const T beta = PCG ? (T)st->beta : (T)0;
    auto pval = [&](int i) -> T {
        if (!PCG) return src[i];
        if (UNIT) return pdir_unit<T>(rvec[i], beta, src[i]);
        return pdir<T>(rvec[i], s.diag[i], beta, src[i]);
    };
    for (int t = threadIdx.x; t < nrows + 2; t += kThreads) {
        const int i = r0 - 1 + t;
        T p = (T)0;
        if (i >= lo && i < hi) p = pval(i);
        sp[t] = p;
        if (PCG && t >= 1 && t <= nrows) pnew[i] = p;
    }
    for (int t = threadIdx.x; t < nrows + 1; t += kThreads) sb[t] = s.band[r0 + t];
    // the diagonal and remainder range of this thread's rows, loaded with the staging above
    constexpr int kRows = (RB + kThreads - 1) / kThreads;
    T dg[kRows];
    int xb[kRows], xe[kRows];
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const int t = (int)threadIdx.x + j * kThreads;
        dg[j] = (!UNIT && t < nrows) ? s.diag[r0 + t] : (T)1;
        xb[j] = t < nrows ? s.xrp[r0 + t] : 0;
        xe[j] = t < nrows ? s.xrp[r0 + t + 1] : 0;
    }
    __syncthreads();
    double part = 0.0;
#pragma unroll
    for (int j = 0; j < kRows; j++) {
        const int t = (int)threadIdx.x + j * kThreads;
        if (t < nrows) {
            const T pr = sp[t + 1];
            T y = UNIT ? pr : dg[j] * pr;
            y = fma(sb[t], sp[t], y);
            y = fma(sb[t + 1], sp[t + 2], y);
            for (int e = xb[j]; e < xe[j]; e += 4) {
                int c[4];
                T xv[4], pc[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const bool in = e + u < xe[j];
                    c[u] = in ? s.xnbr[e + u] : r0;
                    xv[u] = in ? s.xval[e + u] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    pc[u] = (c[u] >= r0 && c[u] < r0 + nrows) ? sp[c[u] - r0 + 1] : pval(c[u]);
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (e + u < xe[j]) y = fma(xv[u], pc[u], y);
            }
            q[r0 + t] = y;
            if (PCG) part += (double)pr * (double)y;
        }
    }
    if (!PCG) return;
    double v[1] = {part}, out[1];
    if (reduce_sys<1>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double pq = out[0];
        if (threadIdx.x == 0) {
            double alpha = st->rz / pq;
            st->pq = pq;
            st->alpha = alpha;
            if (!(pq > 0.0) || !isfinite(alpha)) {
                st->status = 2;
                st->done = 1;
            }
        }
    }
}

// x += alpha p; r -= alpha q; partial r.z, r.r; the last work-group closes the iteration:
// convergence test sqrt(r.r) <= tol*||r_0|| (Direct_Electrifying.cpp:1619-1620), then
// beta = r.z / rz_old (:1627-1632) and the iteration cap (:1559).
template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_update_f64(int n, BatchDev bd,
                                                         const double *__restrict__ diag,
                                                         double *__restrict__ x,
                                                         double *__restrict__ r,
                                                         const double *__restrict__ p,
                                                         const double *__restrict__ q,
                                                         PcgState *__restrict__ st_base,
                                                         RedBuf rbuf) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    PcgState *st = st_base + k.sys;
    if (st->done) return;
    __shared__ double red[8];
    __shared__ int flag;
    const double alpha = st->alpha;
    double prz = 0.0, prr = 0.0;
    for (int t = threadIdx.x; t < k.nrows; t += kThreads) {
        const int i = k.r0 + t;
        x[i] = fma(alpha, p[i], x[i]);
        const double rn = fma(-alpha, q[i], r[i]);
        r[i] = rn;
        const double z = (1.0 / diag[i]) * rn;
        prz += z * rn;
        prr += rn * rn;
    }
    double v[2] = {prz, prr}, out[2];
    if (reduce_sys<2>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double rz = out[0], rr = out[1];
        if (threadIdx.x == 0) {
            st->iters += 1;
            st->rr = rr;
            if (rr <= st->stop2) {
                st->status = 0;  // CICY_OK
                st->done = 1;
            } else if (!isfinite(rr)) {
                st->status = 2;
                st->done = 1;
            } else {
                st->beta = rz / st->rz;
                st->rz = rz;
                if (st->iters >= st->max_iter) {
                    st->status = 1;  // CICY_MAXITER
                    st->done = 1;
                }
            }
        }
    }
}

// Start of a solve: rz = r.(M^-1 r), rr = r.r, stopping threshold.  The reference norm
// is stop_ref2[sys] (or stop_ref2_all when stop_ref2 is null); a negative value means
// "relative to ||r_0||" (the reference's rule), otherwise it is ||b||^2.  Also zeroes both p buffers so that the first direction is p = M^-1 r.
template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_init_f64(int n, BatchDev bd,
                                                       const double *__restrict__ diag,
                                                       const double *__restrict__ r,
                                                       double *__restrict__ p0,
                                                       double *__restrict__ p1,
                                                       PcgState *__restrict__ st_base,
                                                       RedBuf rbuf,
                                                       double tol,
                                                       const double *__restrict__ stop_ref2,
                                                       double stop_ref2_all,
                                                       long long max_iter) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    PcgState *st = st_base + k.sys;
    __shared__ double red[8];
    __shared__ int flag;
    double prz = 0.0, prr = 0.0;
    for (int t = threadIdx.x; t < k.nrows; t += kThreads) {
        const int i = k.r0 + t;
        const double ri = r[i];
        const double z = (1.0 / diag[i]) * ri;
        p0[i] = 0.0;
        p1[i] = 0.0;
        prz += z * ri;
        prr += ri * ri;
    }
    double v[2] = {prz, prr}, out[2];
    if (reduce_sys<2>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double rz = out[0], rr = out[1];
        if (threadIdx.x == 0) {
            st->rz = rz;
            st->rr = rr;
            st->rr_init = rr;
            const double sr = stop_ref2 ? stop_ref2[k.sys] : stop_ref2_all;
            double ref2 = sr >= 0.0 ? sr : rr;
            double thr = tol * sqrt(ref2);
            st->stop2 = thr * thr;
            st->alpha = 0.0;
            st->beta = 0.0;
            st->pq = 0.0;
            st->iters = 0;
            st->max_iter = max_iter;
            st->status = 1;
            st->done = 0;
            if (rr == 0.0) {  // exact initial guess
                st->status = 0;
                st->done = 1;
            }
        }
    }
}

// ---- mixed precision (FP32 inner CG on the Jacobi-scaled system, FP64 reliable updates)

template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_mixed_init(int n, BatchDev bd,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ r0v,
                                                         float *__restrict__ rs,
                                                         float *__restrict__ xs,
                                                         float *__restrict__ ps0,
                                                         float *__restrict__ ps1,
                                                         PcgState *__restrict__ st_base,
                                                         MixState *__restrict__ mx_base,
                                                         RedBuf rbuf,
                                                         double tol,
                                                         const double *__restrict__ stop_ref2,
                                                         double stop_ref2_all,
                                                         long long max_iter, double delta2) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    PcgState *st = st_base + k.sys;
    MixState *mx = mx_base + k.sys;
    __shared__ double red[8];
    __shared__ int flag;
    double a = 0.0, t = 0.0;
    for (int j = threadIdx.x; j < k.nrows; j += kThreads) {
        const int i = k.r0 + j;
        const double ri = r0v[i];
        const float v = (float)(ri * (1.0 / sqrt(diag[i])));
        rs[i] = v;
        xs[i] = 0.f;
        ps0[i] = 0.f;
        ps1[i] = 0.f;
        a += (double)v * (double)v;
        t += ri * ri;
    }
    double v[2] = {a, t}, out[2];
    if (reduce_sys<2>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double rs2 = out[0], r02 = out[1];
        if (threadIdx.x == 0) {
            st->rz = rs2;
            st->rr = rs2;
            st->rr_init = r02;
            const double sr = stop_ref2 ? stop_ref2[k.sys] : stop_ref2_all;
            double ref2 = sr >= 0.0 ? sr : r02;
            double thr = tol * sqrt(ref2);
            st->stop2 = thr * thr;
            st->alpha = st->beta = st->pq = 0.0;
            st->iters = 0;
            st->max_iter = max_iter;
            st->status = 1;
            st->done = 0;
            mx->rr_prev = rs2;
            mx->m2 = rs2;
            mx->rtrue2 = r02;
            mx->tgt2 = r02 > 0.0 ? rs2 * (st->stop2 / r02) : 0.0;
            mx->delta2 = delta2;
            mx->need_update = 0;
            mx->refinements = 0;
            mx->counter = 0;
            mx->stall = 0;
            mx->restarts = 0;
            mx->run = 0;
            if (r02 <= st->stop2) {
                st->status = 0;
                st->done = 1;
            }
        }
    }
}

// x^ += alpha p; r^ -= alpha q (FP32); partial r^.r^ (FP64).  The last work-group sets
// beta, counts the iteration and raises the reliable-update flag.
template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_update_mixed(int n, BatchDev bd,
                                                           float *__restrict__ xs,
                                                           float *__restrict__ rs,
                                                           const float *__restrict__ p,
                                                           const float *__restrict__ q,
                                                           PcgState *__restrict__ st_base,
                                                           MixState *__restrict__ mx_base,
                                                           RedBuf rbuf) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    PcgState *st = st_base + k.sys;
    if (st->done) return;
    MixState *mx = mx_base + k.sys;
    __shared__ double red[8];
    __shared__ int flag;
    const float alpha = (float)st->alpha;
    double prr = 0.0;
    for (int j = threadIdx.x; j < k.nrows; j += kThreads) {
        const int i = k.r0 + j;
        xs[i] = fmaf(alpha, p[i], xs[i]);
        const float rn = fmaf(-alpha, q[i], rs[i]);
        rs[i] = rn;
        prr += (double)rn * (double)rn;
    }
    double v[1] = {prr}, out[1];
    if (reduce_sys<1>(v, k, rbuf, &st->counter, red, &flag, out)) {
        const double rr = out[0];
        if (threadIdx.x == 0) {
            st->iters += 1;
            if (!isfinite(rr)) {
                st->status = 2;
                st->done = 1;
            } else {
                st->beta = rr / st->rz;
                mx->rr_prev = st->rz;
                st->rz = rr;
                st->rr = rr;
                mx->run += 1;
                if (rr <= mx->delta2 * mx->m2 || rr <= mx->tgt2 || mx->run >= kMixedMaxRun)
                    mx->need_update = 1;
                if (st->iters >= st->max_iter) {
                    // fold what has been accumulated and stop
                    mx->need_update = 1;
                    st->done = 1;
                    st->status = 1;
                }
            }
        }
    }
}

// x += D^-1/2 x^; x^ = 0 (only for systems with the update flag unless `always`).
template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_fold(int n, BatchDev bd,
                                                   const double *__restrict__ diag,
                                                   double *__restrict__ x, float *__restrict__ xs,
                                                   const MixState *__restrict__ mx_base,
                                                   int always) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    if (!always && !mx_base[k.sys].need_update) return;
    for (int j = threadIdx.x; j < k.nrows; j += kThreads) {
        const int i = k.r0 + j;
        const float c = xs[i];
        // a correction made non-finite by an FP32 breakdown is dropped: x stays the last
        // finite iterate, from which the FP64 continuation starts
        if (c != 0.f && isfinite(c)) x[i] = fma((double)c, 1.0 / sqrt(diag[i]), x[i]);
        xs[i] = 0.f;
    }
}

// r = b - A x (FP64, ax = A x already formed), r^ = D^-1/2 r in FP32; the last work-group
// tests the stopping rule on ||r|| and restarts the recursion from r^ (same direction p).
template <int RB, bool BATCH>
__global__ __launch_bounds__(kThreads) void k_reliable_resid(int n, BatchDev bd,
                                                             const double *__restrict__ diag,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ ax,
                                                             double *__restrict__ r,
                                                             float *__restrict__ rs,
                                                             PcgState *__restrict__ st_base,
                                                             MixState *__restrict__ mx_base,
                                                             RedBuf rbuf,
                                                             int stall_limit) {
    const Blk k = block_info<RB, BATCH>(bd, n);
    MixState *mx = mx_base + k.sys;
    if (!mx->need_update) return;
    PcgState *st = st_base + k.sys;
    __shared__ double red[8];
    __shared__ int flag;
    double a = 0.0, t = 0.0;
    for (int j = threadIdx.x; j < k.nrows; j += kThreads) {
        const int i = k.r0 + j;
        const double ri = b[i] - ax[i];
        r[i] = ri;
        const float v = (float)(ri * (1.0 / sqrt(diag[i])));
        rs[i] = v;
        a += (double)v * (double)v;
        t += ri * ri;
    }
    double v[2] = {a, t}, out[2];
    if (reduce_sys<2>(v, k, rbuf, &mx->counter, red, &flag, out)) {
        const double rs2 = out[0], rt2 = out[1];
        if (threadIdx.x == 0) {
            mx->refinements += 1;
            mx->need_update = 0;
            mx->run = 0;
            if (rt2 > 0.25 * mx->rtrue2) mx->stall += 1;
            else mx->stall = 0;
            mx->rtrue2 = rt2;
            st->rr = rt2;
            if (rt2 <= st->stop2) {
                st->status = 0;
                st->done = 1;
            } else if (!isfinite(rt2)) {
                st->status = 2;
                st->done = 1;
            } else if (!st->done) {
                // when the FP32 recursion has drifted from the true residual (r.r off by
                // more than 2x) its search direction is stale: restart from r^ (beta = 0)
                if (rs2 > 2.0 * st->rz) {
                    st->beta = 0.0;
                    mx->restarts += 1;
                } else {
                    st->beta = mx->rr_prev > 0.0 ? rs2 / mx->rr_prev : 0.0;
                }
                st->rz = rs2;
                mx->m2 = rs2;
                mx->tgt2 = rs2 * (st->stop2 / rt2);
                if (mx->stall >= stall_limit) {
                    st->status = 1;
                    st->done = 1;
                }
            }
        }
    }
}

#endif  // __HIPCC__

}  // namespace cicy
