// dea_resident.hip -- register-resident FP64 CG for band-dominated DEA systems.
//
// Same algorithm as the reference's solve (Jacobi-PCG in FP64,
// Direct_Electrifying.cpp:1538-1659) run on the Jacobi-scaled matrix
// A^ = D^-1/2 A D^-1/2 (unit diagonal; Jacobi-PCG on A is plain CG on A^), with the stopping
// test on the unscaled residual ||r|| = ||D^1/2 r^||.  What differs is where the vectors live:
//
//   * one cooperative launch runs a chunk of iterations with every work-group resident
//     (512 threads, one work-group per CU); work-group g owns rows [g P, (g+1) P),
//     P = 512 R, thread t holding rows g P + j 512 + t (j < R) in registers: r^ and p
//     never leave the CU inside a chunk;
//   * per iteration only the sub-diagonal band values (twice), x^ (read + write) and, every
//     kResCheck iterations, the diagonal are streamed: about 32 bytes per row instead of
//     ~130 for the two-kernel iteration, and the streamed set fits the Infinity Cache for
//     the 5.6M-row system;
//   * neighbours inside a wave come from lane shuffles, across waves of a work-group from
//     LDS; rows that another work-group (or a remainder entry) needs -- junction rows and
//     the first/last row of each work-group -- are "published" to small slot arrays with
//     device-coherent (agent-scope, sc1) stores and read with agent-scope loads, so no
//     cache-wide L2 write-back is needed on this multi-XCD part;
//   * standard CG has two global reductions per iteration (p.Ap, then r.r); each is a grid
//     barrier built from device-scope atomics on a rotating accumulator slot.  A remote
//     p_k is recomputed from the published r_k and p_{k-1} (p_k = r_k + beta p_{k-1}), so
//     the p.Ap phase needs no extra barrier after the p update.
// Sums are accumulated in arrival order (last bits vary from run to run).
#include "dea_grid.h"
#include "dea_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

using namespace cicy;

namespace cicy {

constexpr int kResThreads = 512;
constexpr int kResWaves = kResThreads / kWave;
constexpr int kResMaxSlots = 6144;          // published rows per work-group (LDS accumulator)
constexpr int kResCheck = 8;                // unscaled-residual test every kResCheck iterations
constexpr int kResChunk = 2048;             // iterations per cooperative launch (at most)
// rows per thread whose loads are issued together before they are used (R is a multiple)
constexpr int kResGroup = 4;
constexpr int kResRows[] = {4, 8, 16, 24, 32, 40, 44, 48};

// Grid-barrier control block (dea_grid.h); cleared by the host before every launch.
using ResCtl = GridCtl;

struct ResState {
    long long k;        // iterations done in this solve
    long long max_iter;
    double rz;          // r^.r^
    double beta;
    double dr2;         // last ||D^1/2 r^||^2 = ||r||^2
    double stop2;       // (tol * ||ref||)^2
    int done;           // 0 running, 1 converged, 2 max_iter, 3 breakdown, 4 barrier timeout
    int started;        // r^ published
};

struct ResDev {
    int n, G, S;
    const double *band;          // [G P + 1] a^(i, i-1), zero where absent and past n
    const double *dvec;          // [G P] diagonal of A (zero past n)
    double *x;                   // [G P] x^ = D^1/2 x
    double *rs, *ps;             // [G P] r^ and p between launches
    double *rpub, *ppub;         // [S], [2 S] published r^ and p (p double-buffered by parity)
    const int *slot_base;        // [G + 1] first slot of each work-group
    const unsigned long long *jbits;  // [G T] bit j: row (g, j, t) is published
    const int *ent_off;          // [G + 1] remainder half-entries of each work-group
    const int *ent_slot;         // local slot of the entry's row
    const int *ent_nbr;          // global slot of the entry's column
    const double *ent_val;       // a^(row, col)
    ResCtl *ctl;
    double *part;                // [2][G][2] per-work-group partial sums, by barrier parity
};

}  // namespace cicy

struct cicy_res {
    int R = 0, G = 0, S = 0, P = 0, nent = 0, max_slots = 0;
    size_t npad = 0;
    double *band = nullptr, *dvec = nullptr, *x = nullptr, *rs = nullptr, *ps = nullptr;
    double *rpub = nullptr, *ppub = nullptr, *ent_val = nullptr, *part = nullptr;
    int *slot_base = nullptr, *ent_off = nullptr, *ent_slot = nullptr, *ent_nbr = nullptr;
    unsigned long long *jbits = nullptr;
    cicy::ResCtl *ctl = nullptr;
    cicy::ResState *st = nullptr;
    cicy::ResState *h_st = nullptr;   // pinned
    size_t bytes = 0;
};

namespace cicy {
namespace {

// ------------------------------------------------------------------------------------
// host layout
// ------------------------------------------------------------------------------------

struct ResHost {
    int R = 0, G = 0, P = 0, S = 0, max_slots = 0;
    size_t npad = 0;
    std::vector<double> band, dvec, ent_val;
    std::vector<int> slot_base, ent_off, ent_slot, ent_nbr;
    std::vector<unsigned long long> jbits;
};

// Resident layout of a validated system for R rows per thread (G = ceil(n / (T R))).
// Returns false (with a reason) when a work-group would publish more than kResMaxSlots rows.
bool build_resident_host(const BlockedHost &h, const BandHost &bh, int R, ResHost &L,
                         const char **why) {
    const int n = h.n, T = kResThreads;
    L.R = R;
    L.P = R * T;
    L.G = (n + L.P - 1) / L.P;
    L.npad = (size_t)L.G * L.P;
    std::vector<double> sd((size_t)n);
    for (int i = 0; i < n; i++) sd[i] = std::sqrt(h.diag[i]);
    L.band.assign(L.npad + 1, 0.0);
    for (int i = 1; i < n; i++)
        if (bh.band[i] != 0.0) L.band[i] = bh.band[i] / (sd[i] * sd[i - 1]);
    L.dvec.assign(L.npad, 0.0);
    for (int i = 0; i < n; i++) L.dvec[i] = h.diag[i];
    // published rows: rows with remainder entries, the first row of every work-group but
    // the first and the last row of every work-group but the last
    std::vector<char> pub((size_t)n, 0);
    for (int k = 0; k < bh.m; k++) pub[bh.xrow[k]] = 1;
    for (int g = 0; g < L.G; g++) {
        if (g > 0) pub[(size_t)g * L.P] = 1;
        if (g + 1 < L.G) pub[(size_t)(g + 1) * L.P - 1] = 1;
    }
    std::vector<int> slot((size_t)n, -1);
    L.slot_base.assign((size_t)L.G + 1, 0);
    L.jbits.assign((size_t)L.G * T, 0u);
    int S = 0;
    L.max_slots = 0;
    for (int g = 0; g < L.G; g++) {
        L.slot_base[g] = S;
        const int lo = g * L.P, hi = std::min(n, lo + L.P);
        for (int i = lo; i < hi; i++)
            if (pub[i]) {
                slot[i] = S++;
                const int off = i - lo;
                L.jbits[(size_t)g * T + (off % T)] |= 1ull << (off / T);
            }
        L.max_slots = std::max(L.max_slots, S - L.slot_base[g]);
    }
    L.slot_base[L.G] = S;
    L.S = S;
    if (L.max_slots > kResMaxSlots) {
        *why = "too many junction rows per work-group for the LDS accumulator";
        return false;
    }
    // remainder half-entries, grouped by work-group (bh is sorted by row, hence by slot)
    L.ent_off.assign((size_t)L.G + 1, 0);
    L.ent_slot.resize((size_t)bh.m);
    L.ent_nbr.resize((size_t)bh.m);
    L.ent_val.resize((size_t)bh.m);
    for (int k = 0; k < bh.m; k++) {
        const int i = bh.xrow[k], j = bh.xnbr[k], g = i / L.P;
        L.ent_off[g + 1]++;
        L.ent_slot[k] = slot[i] - L.slot_base[g];
        L.ent_nbr[k] = slot[j];
        const int a = std::max(i, j), b = std::min(i, j);
        L.ent_val[k] = bh.xval[k] / (sd[a] * sd[b]);
    }
    for (int g = 0; g < L.G; g++) L.ent_off[g + 1] += L.ent_off[g];
    return true;
}

// ------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------

// Element idx of a row-indexed array, addressed as a 32-bit byte offset from a uniform base
// (the arrays are below 4 GiB), so loads and stores use the scalar-base + vector-offset form
// and need one offset register instead of a 64-bit address each.
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned idx) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T &at(const T *base, unsigned idx) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        idx * (unsigned)sizeof(T));
}

// rank of this lane among the lanes of its wave whose predicate mask bit is set
__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Grid-wide sum and barrier m of this launch (grid_sum, dea_grid.h).
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], const ResDev &a, unsigned m,
                                            double *lds, int *flag) {
    return grid_sum<NV, kResWaves>(v, a.G, a.part, a.ctl, m, lds, flag);
}

// ------------------------------------------------------------------------------------
// the resident CG kernel
// ------------------------------------------------------------------------------------

// Hides a value from the optimiser: per-row offsets and slot ranks derived from it are
// recomputed where used instead of being hoisted out of the iteration loop, where they
// would hold a register per row on top of r^ and p.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ unsigned long long opaque64(unsigned long long v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int R>
__global__ __launch_bounds__(kResThreads) void k_cg_resident(ResDev a, ResState *sp, int chunk) {
    constexpr int T = kResThreads, W = kResWaves, P = R * T, NSEG = R * W;
    __shared__ double acc[kResMaxSlots];
    // wave segments in row order (segment j W + w = rows base + j T + w 64 .. + 63):
    // lastv[1 + q] = p of the last row of segment q, lastv[0] = p of row base - 1;
    // firstv[q] = p of the first row of segment q, firstv[NSEG] = p of row base + P
    __shared__ double lastv[NSEG + 1], firstv[NSEG + 1];
    __shared__ int woff[NSEG];
    __shared__ double red[2 * W];
    __shared__ int flag, nsl_s;
    const int g = blockIdx.x;
    const unsigned base = (unsigned)g * P;
    ResState s = *sp;
    if (s.done) return;

    double r[R], p[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        r[j] = at(a.rs, base + j * T + threadIdx.x);
        p[j] = at(a.ps, base + j * T + threadIdx.x);
    }
    const unsigned long long bits0 = a.jbits[(size_t)g * T + threadIdx.x];
#pragma unroll
    for (int j = 0; j < R; j++) {
        const unsigned long long bm = __ballot((bits0 >> j) & 1ull);
        if ((threadIdx.x & (kWave - 1)) == 0) woff[j * W + threadIdx.x / kWave] = __popcll(bm);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int o = 0;
        for (int q = 0; q < NSEG; q++) {
            const int c = woff[q];
            woff[q] = o;
            o += c;
        }
        nsl_s = o;
    }
    __syncthreads();
    const int nsl = nsl_s;
    const int S = a.S;
    const unsigned sbase = (unsigned)a.slot_base[g];
    const int ebeg = a.ent_off[g], eend = a.ent_off[g + 1];
    const int sprev = g > 0 ? (int)sbase - 1 : -1;
    const int snext = g + 1 < a.G ? a.slot_base[g + 1] : -1;
    unsigned m = 0;

    // Re-read through opaque() at the top of every phase, so that the compiler derives
    // row offsets, LDS addresses and slot ranks from one register each time instead of
    // hoisting R copies of them out of the iteration loop.
    unsigned t = threadIdx.x;
    unsigned long long bits = bits0;
    unsigned lane = t & (kWave - 1), w = t / kWave, bt = base + t;
    auto refresh = [&]() {
        t = (unsigned)opaque((int)t);
        bits = opaque64(bits);
        lane = t & (kWave - 1);
        w = t / kWave;
        bt = base + t;
    };
    // local slot of row j (meaningful where bit j is set); every lane must call it
    auto lslot = [&](int j) -> unsigned {
        const unsigned long long bm = __ballot((bits >> j) & 1ull);
        return (unsigned)woff[j * W + w] + (unsigned)lane_rank(bm);
    };
    // acc[slot] += a^(row, col) p_col over the work-group's remainder entries; pval(nb)
    // returns p of published row nb.  Four entries per thread are in flight together
    // (indices, values and published rows loaded before any is used); lanes past the end
    // add 0 to slot 0.
    auto entries = [&](unsigned tt, auto pval) {
        for (int e0 = ebeg + (int)tt; e0 < eend; e0 += 4 * T) {
            int nb[4], sl[4];
            double va[4], pv[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int e = e0 + u * T;
                const bool in = e < eend;
                nb[u] = in ? a.ent_nbr[e] : 0;
                sl[u] = in ? a.ent_slot[e] : 0;
                va[u] = in ? a.ent_val[e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) pv[u] = pval(nb[u]);
#pragma unroll
            for (int u = 0; u < 4; u++) atomicAdd(&acc[sl[u]], va[u] * pv[u]);
        }
    };
    // row j of A^ p from its band values: unit diagonal, band neighbours from the lanes or
    // the segment edges, remainder entries from acc.  Branch-free (every lane reads the
    // broadcast edge values and an in-range acc entry), so that the band loads of a whole
    // group of rows are in flight together.
    auto qrow = [&](int j, double blo, double bhi, unsigned long long bg) -> double {
        const double pj = p[j];
        const double up = __shfl_up(pj, 1, kWave);
        const double dn = __shfl_down(pj, 1, kWave);
        const double em = lastv[j * W + w];
        const double ep = firstv[j * W + w + 1];
        const double pm = lane == 0 ? em : up;
        const double pp = lane == kWave - 1 ? ep : dn;
        double q = fma(blo, pm, pj);
        q = fma(bhi, pp, q);
        const unsigned long long bm = __ballot((bg >> j) & 1ull);
        const unsigned ls =
            min((unsigned)woff[j * W + w] + (unsigned)lane_rank(bm), (unsigned)(kResMaxSlots - 1));
        const double av = acc[ls];
        return ((bg >> j) & 1ull) ? q + av : q;
    };
    // p.q over one group of rows (band loads of the group issued together)
    auto pq_group = [&](int j0, double &sum) {
        const unsigned b0 = (unsigned)opaque((int)(bt + j0 * T));
        const unsigned long long bg = opaque64(bits);
        double blo[kResGroup], bhi[kResGroup];
#pragma unroll
        for (int u = 0; u < kResGroup; u++) {
            blo[u] = at(a.band, b0 + u * T);
            bhi[u] = at(a.band, b0 + u * T + 1);
        }
        // the group's loads are issued above; rows are then finished one at a time so that
        // only one row's temporaries are live next to the loaded values
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kResGroup; u++) {
            sum = fma(p[j0 + u], qrow(j0 + u, blo[u], bhi[u], bg), sum);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // phase B body for one group of rows; CHECK adds the ||D^1/2 r^||^2 term
    auto update_group = [&](auto check_tag, int j0, double alpha, double (&v2)[2]) {
        constexpr bool CHECK = decltype(check_tag)::value;
        const unsigned b0 = (unsigned)opaque((int)(bt + j0 * T));
        const unsigned long long bg = opaque64(bits);
        double blo[kResGroup], bhi[kResGroup], xv[kResGroup], dv[kResGroup];
#pragma unroll
        for (int u = 0; u < kResGroup; u++) {
            blo[u] = at(a.band, b0 + u * T);
            bhi[u] = at(a.band, b0 + u * T + 1);
            xv[u] = at(a.x, b0 + u * T);
            if (CHECK) dv[u] = at(a.dvec, b0 + u * T);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < kResGroup; u++) {
            const int j = j0 + u;
            const double q = qrow(j, blo[u], bhi[u], bg);
            at(a.x, b0 + u * T) = fma(alpha, p[j], xv[u]);
            r[j] = fma(-alpha, q, r[j]);
            const double r2 = r[j] * r[j];
            v2[0] += r2;
            if (CHECK) v2[1] = fma(dv[u], r2, v2[1]);
            __builtin_amdgcn_sched_barrier(0);
        }
    };

    if (!s.started) {
        refresh();
        refresh();
#pragma unroll
        for (int j = 0; j < R; j++) {
            const unsigned ls = lslot(j);
            if ((bits >> j) & 1ull) st_coh(&a.rpub[sbase + ls], r[j]);
        }
        double none[1] = {0.0};
        if (!grid_reduce<1>(none, a, m++, red, &flag)) s.done = 4;
        s.started = 1;
    }

    for (int it = 0; it < chunk && !s.done; it++) {
        const double beta = s.beta;
        const double *pold = a.ppub + (size_t)((s.k + 1) & 1) * S;  // p_{k-1}
        double *pcur = a.ppub + (size_t)(s.k & 1) * S;               // p_k
        // ---- p = r^ + beta p; q = A^ p; p.q
        refresh();
#pragma unroll
        for (int j = 0; j < R; j++) {
            p[j] = fma(beta, p[j], r[j]);
            const unsigned ls = lslot(j);
            if ((bits >> j) & 1ull) st_coh(&pcur[sbase + ls], p[j]);
            if (lane == 0) firstv[j * W + w] = p[j];
            if (lane == kWave - 1) lastv[j * W + w + 1] = p[j];
        }
        for (int i = (int)t; i < nsl; i += T) acc[i] = 0.0;
        if (t == 0) lastv[0] = sprev >= 0 ? fma(beta, ld_coh(&pold[sprev]), ld_coh(&a.rpub[sprev])) : 0.0;
        if (t == kWave)
            firstv[NSEG] = snext >= 0 ? fma(beta, ld_coh(&pold[snext]), ld_coh(&a.rpub[snext])) : 0.0;
        __syncthreads();
        entries(t, [&](int nb) { return fma(beta, ld_coh(&pold[nb]), ld_coh(&a.rpub[nb])); });
        __syncthreads();
        double v1[1] = {0.0};
        refresh();
#pragma unroll
        for (int j0 = 0; j0 < R; j0 += kResGroup) pq_group(j0, v1[0]);
        if (!grid_reduce<1>(v1, a, m++, red, &flag)) {
            s.done = 4;
            break;
        }
        const double pq = v1[0];
        const double alpha = s.rz / pq;
        if (!(pq > 0.0) || !isfinite(alpha)) {
            s.done = 3;
            break;
        }
        // ---- x^ += alpha p; r^ -= alpha q (q recomputed from the published p_k)
        refresh();
        for (int i = (int)t; i < nsl; i += T) acc[i] = 0.0;
        if (t == 0) lastv[0] = sprev >= 0 ? ld_coh(&pcur[sprev]) : 0.0;
        if (t == kWave) firstv[NSEG] = snext >= 0 ? ld_coh(&pcur[snext]) : 0.0;
        __syncthreads();
        entries(t, [&](int nb) { return ld_coh(&pcur[nb]); });
        __syncthreads();
        const bool check = ((s.k + 1) % kResCheck == 0) || (s.k + 1 >= s.max_iter);
        double v2[2] = {0.0, 0.0};
        if (check) {
#pragma unroll
            for (int j0 = 0; j0 < R; j0 += kResGroup)
                update_group(std::true_type{}, j0, alpha, v2);
        } else {
#pragma unroll
            for (int j0 = 0; j0 < R; j0 += kResGroup)
                update_group(std::false_type{}, j0, alpha, v2);
        }
        refresh();
#pragma unroll
        for (int j = 0; j < R; j++) {
            const unsigned ls = lslot(j);
            if ((bits >> j) & 1ull) st_coh(&a.rpub[sbase + ls], r[j]);
        }
        if (!grid_reduce<2>(v2, a, m++, red, &flag)) {
            s.done = 4;
            break;
        }
        s.beta = v2[0] / s.rz;
        s.rz = v2[0];
        s.k++;
        if (check) {
            s.dr2 = v2[1];
            if (v2[1] <= s.stop2) {
                s.done = 1;
                break;
            }
        }
        if (s.k >= s.max_iter) s.done = 2;
    }
    refresh();
#pragma unroll
    for (int j = 0; j < R; j++) {
        at(a.rs, bt + j * T) = r[j];
        at(a.ps, bt + j * T) = p[j];
    }
    if (g == 0 && t == 0) *sp = s;
}

// x^ = D^1/2 x0, r^ = D^-1/2 r0, p = 0 (rows past n: zero); rz = r^.r^, r0.r0 -> out2
__global__ __launch_bounds__(kThreads) void k_res_init(int n, int npad, const double *__restrict__ d,
                                                       const double *__restrict__ x0,
                                                       const double *__restrict__ r0,
                                                       double *__restrict__ xh,
                                                       double *__restrict__ rh,
                                                       double *__restrict__ ph,
                                                       double *__restrict__ out2,
                                                       unsigned int *__restrict__ counter,
                                                       RedBuf rbuf) {
    __shared__ double red[8];
    __shared__ int flag;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < npad; i += gridDim.x * kThreads) {
        double xv = 0.0, rv = 0.0;
        if (i < n) {
            const double sd = sqrt(d[i]);
            xv = x0[i] * sd;
            rv = r0[i] / sd;
            v[0] += rv * rv;
            v[1] += r0[i] * r0[i];
        }
        xh[i] = xv;
        rh[i] = rv;
        ph[i] = 0.0;
    }
    Blk k;
    k.b = blockIdx.x;
    k.sys = 0;
    k.r0 = 0;
    k.nrows = 0;
    k.first = 0;
    k.count = gridDim.x;
    double out[2];
    if (reduce_sys<2>(v, k, rbuf, counter, red, &flag, out) && threadIdx.x == 0) {
        out2[0] = out[0];
        out2[1] = out[1];
    }
}

// r <- b - ax
__global__ __launch_bounds__(kThreads) void k_res_b_minus(int n, double *__restrict__ r,
                                                         const double *__restrict__ b,
                                                         const double *__restrict__ ax) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

// x = D^-1/2 x^, r = D^1/2 r^
__global__ __launch_bounds__(kThreads) void k_res_finish(int n, const double *__restrict__ d,
                                                         const double *__restrict__ xh,
                                                         const double *__restrict__ rh,
                                                         double *__restrict__ x,
                                                         double *__restrict__ r) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        const double sd = sqrt(d[i]);
        x[i] = xh[i] / sd;
        r[i] = rh[i] * sd;
    }
}

template <int R>
const void *res_kernel() {
    return reinterpret_cast<const void *>(&k_cg_resident<R>);
}

const void *res_kernel_for(int R) {
    switch (R) {
        case 4: return res_kernel<4>();
        case 8: return res_kernel<8>();
        case 16: return res_kernel<16>();
        case 24: return res_kernel<24>();
        case 32: return res_kernel<32>();
        case 40: return res_kernel<40>();
        case 44: return res_kernel<44>();
        case 48: return res_kernel<48>();
        default: return nullptr;
    }
}

// work-groups of the resident kernel for R that the device keeps resident at once
int res_capacity(cicy_ctx *ctx, int R) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, res_kernel_for(R), kResThreads,
                                                     0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
        hipSuccess)
        return 0;
    return per_cu * cus;
}

template <typename T>
bool res_alloc(cicy_ctx *ctx, cicy_res *z, T **p, size_t count) {
    if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
        set_error(ctx, "hipMalloc of %zu bytes failed (resident layout)", count * sizeof(T));
        return false;
    }
    z->bytes += count * sizeof(T);
    return true;
}

template <typename T>
bool res_h2d(cicy_ctx *ctx, T *dst, const std::vector<T> &src) {
    if (src.empty()) return true;
    return hip_ok(ctx, hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T),
                                      hipMemcpyHostToDevice, ctx->stream),
                  "H2D (resident layout)");
}

}  // namespace

void resident_free(cicy_res *z) {
    if (!z) return;
    void *ptrs[] = {z->band, z->dvec, z->x, z->rs, z->ps, z->rpub, z->ppub, z->ent_val, z->part,
                    z->slot_base, z->ent_off, z->ent_slot, z->ent_nbr, z->jbits, z->ctl, z->st};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (z->h_st) (void)hipHostFree(z->h_st);
    delete z;
}

// Builds the resident layout of a band-format system; *why says why not when it returns
// nullptr without an error (too large for the resident grid, too many junction rows).
cicy_res *resident_setup(cicy_ctx *ctx, const BlockedHost &h, const BandHost &bh, int R_force,
                         const char **why, bool *failed) {
    *failed = false;
    *why = nullptr;
    if (h.n == 0) {
        *why = "empty system";
        return nullptr;
    }
    int R = 0, cap = 0;
    for (int cand : kResRows) {
        if (R_force && cand != R_force) continue;
        const int G = (h.n + cand * kResThreads - 1) / (cand * kResThreads);
        cap = res_capacity(ctx, cand);
        if (G <= cap) {
            R = cand;
            break;
        }
    }
    if (!R) {
        *why = R_force ? "the forced rows per thread do not fit the device's resident grid"
                       : "system larger than the registers of the resident grid";
        return nullptr;
    }
    ResHost L;
    if (!build_resident_host(h, bh, R, L, why)) return nullptr;
    cicy_res *z = new cicy_res();
    z->R = L.R;
    z->G = L.G;
    z->S = L.S;
    z->P = L.P;
    z->npad = L.npad;
    z->nent = bh.m;
    z->max_slots = L.max_slots;
    bool ok = res_alloc(ctx, z, &z->band, L.npad + 1) && res_alloc(ctx, z, &z->dvec, L.npad) &&
              res_alloc(ctx, z, &z->x, L.npad) && res_alloc(ctx, z, &z->rs, L.npad) &&
              res_alloc(ctx, z, &z->ps, L.npad) && res_alloc(ctx, z, &z->rpub, (size_t)L.S) &&
              res_alloc(ctx, z, &z->ppub, 2 * (size_t)L.S) &&
              res_alloc(ctx, z, &z->ent_val, (size_t)bh.m) &&
              res_alloc(ctx, z, &z->slot_base, (size_t)L.G + 1) &&
              res_alloc(ctx, z, &z->ent_off, (size_t)L.G + 1) &&
              res_alloc(ctx, z, &z->ent_slot, (size_t)bh.m) &&
              res_alloc(ctx, z, &z->ent_nbr, (size_t)bh.m) &&
              res_alloc(ctx, z, &z->jbits, (size_t)L.G * kResThreads) &&
              res_alloc(ctx, z, &z->ctl, 1) && res_alloc(ctx, z, &z->st, 1) &&
              res_alloc(ctx, z, &z->part, 4 * (size_t)L.G) &&
              hip_ok(ctx, hipHostMalloc((void **)&z->h_st, sizeof(ResState), hipHostMallocDefault),
                     "hipHostMalloc");
    ok = ok && res_h2d(ctx, z->band, L.band) && res_h2d(ctx, z->dvec, L.dvec) &&
         res_h2d(ctx, z->ent_val, L.ent_val) && res_h2d(ctx, z->slot_base, L.slot_base) &&
         res_h2d(ctx, z->ent_off, L.ent_off) && res_h2d(ctx, z->ent_slot, L.ent_slot) &&
         res_h2d(ctx, z->ent_nbr, L.ent_nbr) && res_h2d(ctx, z->jbits, L.jbits) &&
         hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!ok) {
        resident_free(z);
        *failed = true;
        return nullptr;
    }
    return z;
}

// One resident solve from s->x0 / s->r0 (restarted from b - A x while the true residual
// misses the REL_B target); result in s->x and s->r.
bool solve_resident(cicy_sys *s, double tol, long long max_iter, int rel_to,
                    cicy_solve_stats *stats) {
    cicy_ctx *ctx = s->ctx;
    cicy_res *z = s->res;
    if (!z) {
        set_error(ctx, "no resident layout for this system (%s)",
                  s->res_why ? s->res_why : "not built");
        if (stats) stats->status = CICY_BAD_INPUT;
        return false;
    }
    if (rel_to == CICY_REL_B && !s->has_b && s->ref2_override < 0.0) {
        set_error(ctx, "REL_B stopping rule needs b (cicy_sys_set_rhs)");
        return false;
    }
    const double max_seconds = [] {
        const char *e = getenv("CICY_MAX_SOLVE_SECONDS");
        return e ? atof(e) : 0.0;
    }();
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = ctx->stream;
    const int n = s->n, npad = (int)z->npad;
    ResDev a{n, z->G, z->S, z->band, z->dvec, z->x, z->rs, z->ps, z->rpub, z->ppub,
             z->slot_base, z->jbits, z->ent_off, z->ent_slot, z->ent_nbr, z->ent_val, z->ctl,
             z->part};
    const void *kern = res_kernel_for(z->R);
    bool ok = hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "event");
    long long total = 0;
    int restarts = 0, status = CICY_MAXITER;
    double rel_true = -1.0, rel_rec = 0.0;
    const double *x0 = s->x0, *r0 = s->r0;
    while (ok) {
        // initial state
        // grid-stride over the padded rows with at most one reduction slot per row block
        const int ib = std::max(1, std::min(s->nblocks, 1024));
        hipLaunchKernelGGL(k_res_init, dim3(ib), dim3(kThreads), 0, st, n, npad, s->diag, x0,
                           r0, z->x, z->rs, z->ps, s->scal, s->counter2, redbuf(s));
        double nrm[2];
        ok = ok && hip_ok(ctx, hipMemsetAsync(z->ppub, 0, 2 * (size_t)z->S * sizeof(double), st),
                          "memset") &&
             d2h_sync(ctx, nrm, s->scal, sizeof(nrm));
        if (!ok) break;
        const double ref2 = s->ref2_override >= 0.0 ? s->ref2_override
                            : rel_to == CICY_REL_B  ? s->bnorm2
                                                    : nrm[1];
        ResState h;
        memset(&h, 0, sizeof(h));
        h.max_iter = max_iter - total;
        h.rz = nrm[0];
        h.dr2 = nrm[1];
        h.stop2 = tol * tol * ref2;
        if (nrm[1] <= h.stop2) h.done = 1;  // already converged (zero residual included)
        ok = hip_ok(ctx, hipMemcpyAsync(z->st, &h, sizeof(h), hipMemcpyHostToDevice, st), "H2D");
        while (ok && !h.done) {
            int chunk = kResChunk;
            void *args[] = {(void *)&a, (void *)&z->st, (void *)&chunk};
            ok = hip_ok(ctx, hipMemsetAsync(z->ctl, 0, sizeof(ResCtl), st), "memset") &&
                 hip_ok(ctx, hipLaunchCooperativeKernel(kern, dim3(z->G), dim3(kResThreads), args,
                                                        0, st),
                        "hipLaunchCooperativeKernel (resident CG)") &&
                 d2h_sync(ctx, &h, z->st, sizeof(h));
            if (ok && max_seconds > 0 &&
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
                    max_seconds)
                break;
        }
        if (!ok) break;
        total += h.k;
        hipLaunchKernelGGL(k_res_finish, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                           n, s->diag, z->x, z->rs, s->x, s->r);
        rel_rec = ref2 > 0 ? std::sqrt(h.dr2 / ref2) : std::sqrt(h.dr2);
        if (h.done == 4) {
            set_error(ctx, "resident CG: grid barrier timed out (work-groups not co-resident)");
            status = CICY_DEVICE_ERROR;
            ok = false;
            break;
        }
        if (h.done == 3) {
            status = CICY_BREAKDOWN;
            break;
        }
        status = h.done == 1 ? CICY_OK : CICY_MAXITER;
        if (!s->has_b) break;
        ok = sys_true_residual(s, s->x, &rel_true);
        if (!ok || rel_to != CICY_REL_B || status != CICY_OK || rel_true <= tol ||
            restarts >= 4 || total >= max_iter)
            break;
        // restart from x with r = b - A x (q holds A x from the true-residual check); the
        // caller's x0 / r0 stay untouched
        restarts++;
        hipLaunchKernelGGL(k_res_b_minus, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                           st, n, s->r, s->b, s->q);
        x0 = s->x;
        r0 = s->r;
    }
    float ms = 0.f;
    if (hipEventRecord(ctx->ev_stop, st) == hipSuccess &&
        hipEventSynchronize(ctx->ev_stop) == hipSuccess)
        (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    if (stats) {
        stats->iters = total;
        stats->status = ok ? status : CICY_DEVICE_ERROR;
        stats->refinements = restarts;
        stats->rel_res = rel_rec;
        stats->true_rel_res = rel_true;
        stats->solve_ms = ms;
        stats->fp64_iters = total;
    }
    return ok;
}

}  // namespace cicy

// ------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------

extern "C" {

int cicy_set_resident_rows(cicy_ctx *ctx, int rows_per_thread) {
    if (!ctx) return 0;
    if (rows_per_thread != 0 && !cicy::res_kernel_for(rows_per_thread)) {
        cicy::set_error(ctx, "rows per thread must be 0 (auto) or one of 4 8 16 24 32 40 44 48");
        return 0;
    }
    ctx->resident_rows = rows_per_thread;
    return 1;
}

// Host emulation of one product of the resident kernel: out = A^ (r + beta pold), with
// A^ = D^-1/2 A D^-1/2, computed by walking the layout exactly as the kernel does (lanes,
// waves, published slots; phase 0 recomputes remote p from the published r and p_old,
// phase 1 reads the published p).  For layout tests without a GPU.
int cicy_host_resident_apply(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                             const double *values, const double *diagonal, int64_t nnz,
                             int rows_per_thread, int phase, const double *r, const double *pold,
                             double beta, double *out) {
    using namespace cicy;
    cicy_ctx dummy;
    BlockedHost h;
    if (!res_kernel_for(rows_per_thread) || n <= 0) return 0;
    if (!build_blocked(&dummy, n, nnz, row_ptr, col_ind, values, diagonal, 1024, h)) return 0;
    BandHost bh;
    build_band(h, bh);
    ResHost L;
    const char *why = nullptr;
    if (!build_resident_host(h, bh, rows_per_thread, L, &why)) return 0;
    const int T = kResThreads, W = kResWaves, R = L.R, P = L.P;
    // published arrays as the previous iteration left them
    std::vector<double> rpub((size_t)L.S), pold_pub((size_t)L.S), pcur_pub((size_t)L.S);
    std::vector<double> rv(L.npad, 0.0), pv(L.npad, 0.0), pn(L.npad, 0.0);
    for (int64_t i = 0; i < n; i++) {
        rv[i] = r[i];
        pv[i] = pold[i];
        pn[i] = std::fma(beta, pold[i], r[i]);
    }
    for (int g = 0; g < L.G; g++) {
        // slot offsets as the kernel forms them: (j, wave) order, lane rank inside
        std::vector<int> woff((size_t)R * W, 0);
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kWave; l++)
                    woff[j * W + w] += (int)((L.jbits[(size_t)g * T + w * kWave + l] >> j) & 1ull);
        int o = 0;
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++) {
                const int c = woff[j * W + w];
                woff[j * W + w] = o;
                o += c;
            }
        if (o != L.slot_base[g + 1] - L.slot_base[g]) return 0;
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++) {
                int rank = 0;
                for (int l = 0; l < kWave; l++) {
                    const int t = w * kWave + l;
                    if ((L.jbits[(size_t)g * T + t] >> j) & 1ull) {
                        const int gs = L.slot_base[g] + woff[j * W + w] + rank++;
                        const size_t row = (size_t)g * P + (size_t)j * T + t;
                        rpub[gs] = rv[row];
                        pold_pub[gs] = pv[row];
                        pcur_pub[gs] = pn[row];
                    }
                }
            }
    }
    for (int g = 0; g < L.G; g++) {
        const int sbase = L.slot_base[g], nsl = L.slot_base[g + 1] - sbase;
        auto remote = [&](int gs) {
            return phase == 0 ? std::fma(beta, pold_pub[gs], rpub[gs]) : pcur_pub[gs];
        };
        std::vector<double> acc((size_t)nsl, 0.0);
        for (int e = L.ent_off[g]; e < L.ent_off[g + 1]; e++)
            acc[L.ent_slot[e]] += L.ent_val[e] * remote(L.ent_nbr[e]);
        const double bc0 = g > 0 ? remote(sbase - 1) : 0.0;
        const double bc1 = g + 1 < L.G ? remote(L.slot_base[g + 1]) : 0.0;
        std::vector<int> woff((size_t)R * W, 0);
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++)
                for (int l = 0; l < kWave; l++)
                    woff[j * W + w] += (int)((L.jbits[(size_t)g * T + w * kWave + l] >> j) & 1ull);
        int o = 0;
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++) {
                const int c = woff[j * W + w];
                woff[j * W + w] = o;
                o += c;
            }
        // wave edges as the kernel stores them in LDS
        std::vector<double> e_first((size_t)W * R), e_last((size_t)W * R);
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++) {
                const size_t r0 = (size_t)g * P + (size_t)j * T + (size_t)w * kWave;
                e_first[(size_t)w * R + j] = pn[r0];
                e_last[(size_t)w * R + j] = pn[r0 + kWave - 1];
            }
        for (int j = 0; j < R; j++)
            for (int w = 0; w < W; w++) {
                int rank = 0;
                for (int l = 0; l < kWave; l++) {
                    const int t = w * kWave + l;
                    const size_t row = (size_t)g * P + (size_t)j * T + t;
                    const double pj = pn[row];
                    // the kernel's neighbour sources: lane shuffle, LDS wave edge, remote
                    double pm, pp;
                    if (l > 0) pm = pn[row - 1];
                    else if (w > 0) pm = e_last[(size_t)(w - 1) * R + j];
                    else pm = j > 0 ? e_last[(size_t)(W - 1) * R + j - 1] : bc0;
                    if (l < kWave - 1) pp = pn[row + 1];
                    else if (w < W - 1) pp = e_first[(size_t)(w + 1) * R + j];
                    else pp = j < R - 1 ? e_first[(size_t)j + 1] : bc1;
                    double q = std::fma(L.band[row], pm, pj);
                    q = std::fma(L.band[row + 1], pp, q);
                    if ((L.jbits[(size_t)g * T + t] >> j) & 1ull) q += acc[woff[j * W + w] + rank++];
                    if ((int64_t)row < n) out[row] = q;
                }
            }
    }
    return 1;
}

}  // extern "C"

int cicy::resident_info(const cicy_res *z, int what) {
    if (!z) return 0;
    switch (what) {
        case 0: return z->R;
        case 1: return z->G;
        case 2: return z->S;
        default: return 0;
    }
}
