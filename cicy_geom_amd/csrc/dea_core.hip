// dea_core.hip -- block-preconditioned persistent CG for the core of the condensed solve.
//
// After the condensed solve's exact eliminations (dea_condensed.hip) the core system is
// small (1-5 % of the rows) and one CG iteration over it moves only a few tens of MB: the
// two-kernel PCG iteration is then bound by launch and reduction latency, not bandwidth.
// Nanoplatelets make it worse numerically: a plate's triangulation edges conduct ~1e4 S
// against ~1e-3 S for its junctions, so each plate contributes a nearly rigid voltage mode
// that Jacobi scaling leaves at an eigenvalue ~1e-7 and CG needs tens of thousands of
// iterations.  Here the core is solved by
//
//   * a Chronopoulos-Gear formulation of PCG (one fused pair of dot products per
//     reduction, two grid barriers per iteration) run as one cooperative launch per chunk
//     of iterations, every work-group resident on its own CU;
//   * block-Jacobi preconditioning: row groups given by the caller (cicy_sys_set_blocks;
//     the nodes of one nanoplatelet, which the reference numbers consecutively in
//     LM_matrix_for_gnps, Direct_Electrifying.cpp:307-349) are inverted exactly (dense
//     FP64 inverse of the core's diagonal block, computed once per system), the other rows
//     use the reference's Jacobi preconditioner (Direct_Electrifying.cpp:1662-1676).
//     Each work-group owns whole blocks, so the block solve is a work-group-local dense
//     product with the block's residual staged in LDS;
//   * vectors in HBM: the preconditioned residual u, which other work-groups gather for the
//     product w = A u, is written and read with device-coherent accesses (dea_grid.h); x,
//     r, p, s, w are private to the owning work-group.
//
// Iteration (u = M^-1 r, w = A u, gamma = r.u, delta = w.u; Chronopoulos & Gear 1989):
//   p = u + beta p;  s = w + beta s;  x += alpha p;  r -= alpha s;  u = M^-1 r
//   [barrier: gamma' = r.u, r.r]  stop when r.r <= tol^2 ||ref||^2
//   w = A u
//   [barrier: delta = w.u]
//   beta = gamma' / gamma;  alpha = gamma' / (delta - beta gamma' / alpha)
// Sums over work-groups are taken in a fixed order, so solves are bit-reproducible.
#include "dea_grid.h"
#include "dea_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

using namespace cicy;

namespace cicy {

constexpr int kCoreThreads = 512;
constexpr int kCoreWaves = kCoreThreads / kWave;
constexpr int kCoreMaxBlock = 128;     // rows of a dense block (larger groups are split)
constexpr int kCoreChunk = 4096;       // iterations per cooperative launch (at most)
constexpr int kCoreRowsPerGroup = 192; // work-groups: about one per this many rows

struct CoreState {
    long long k;         // iterations done in this solve
    long long max_iter;
    double gamma, alpha, beta, rr, stop2;
    int done;            // 0 running, 1 converged, 2 iteration cap, 3 breakdown, 4 barrier timeout
    int started;         // the prologue (u0 = M^-1 r0, w0 = A u0) has run
};

struct CoreDev {
    int m, G;
    const int *seg;       // [G+1] rows of each work-group (blocks are never split)
    const int *rp;        // [m+1] full symmetric CSR without the diagonal
    const int *col;
    const double *val;
    const double *diag;   // [m]
    const int *blk_of;    // [m] block of each row, -1 for Jacobi rows
    const int *wblk;      // [G+1] blocks of each work-group
    const int *blk_row0;  // [nb]
    const int *blk_n;     // [nb]
    const long long *blk_off;  // [nb] offset of the block's inverse (row-major, symmetric)
    const double *binv;
    double *x, *r, *u, *w, *p, *s;
    GridCtl *ctl;
    double *part;         // [2][G][2]
};

}  // namespace cicy

struct cicy_core {
    int m = 0, G = 0, nb = 0, blocked_rows = 0;
    long long nnz2 = 0, inv_entries = 0;
    int *seg = nullptr, *rp = nullptr, *col = nullptr, *blk_of = nullptr, *wblk = nullptr;
    int *blk_row0 = nullptr, *blk_n = nullptr;
    long long *blk_off = nullptr;
    double *val = nullptr, *diag = nullptr, *binv = nullptr;
    double *x = nullptr, *r = nullptr, *u = nullptr, *w = nullptr, *p = nullptr, *s = nullptr;
    double *part = nullptr;
    cicy::GridCtl *ctl = nullptr;
    cicy::CoreState *st = nullptr;
    size_t bytes = 0;
};

namespace cicy {
namespace {

// ------------------------------------------------------------------------------------
// host layout
// ------------------------------------------------------------------------------------

struct CoreHost {
    int m = 0, G = 1;
    std::vector<int> rp, col, blk_of, seg, wblk, blk_row0, blk_n;
    std::vector<double> val, diag, binv;
    std::vector<long long> blk_off;
};

// In-place Cholesky of the dense SPD matrix a (n x n, row-major) and its inverse into inv.
// Returns false when a pivot is not positive.
bool dense_spd_inverse(int n, std::vector<double> &a, double *inv) {
    for (int j = 0; j < n; j++) {
        double d = a[(size_t)j * n + j];
        for (int k = 0; k < j; k++) d -= a[(size_t)j * n + k] * a[(size_t)j * n + k];
        if (!(d > 0.0) || !std::isfinite(d)) return false;
        const double l = std::sqrt(d);
        a[(size_t)j * n + j] = l;
        for (int i = j + 1; i < n; i++) {
            double v = a[(size_t)i * n + j];
            for (int k = 0; k < j; k++) v -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
            a[(size_t)i * n + j] = v / l;
        }
    }
    // columns of the inverse: L y = e_c, L^T z = y
    std::vector<double> y((size_t)n);
    for (int c = 0; c < n; c++) {
        for (int i = 0; i < n; i++) {
            double v = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; k++) v -= a[(size_t)i * n + k] * y[k];
            y[i] = v / a[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double v = y[i];
            for (int k = i + 1; k < n; k++) v -= a[(size_t)k * n + i] * y[k];
            y[i] = v / a[(size_t)i * n + i];
        }
        for (int i = 0; i < n; i++) inv[(size_t)i * n + c] = y[i];
    }
    // exact symmetry (the product reads the inverse by columns)
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            const double v = 0.5 * (inv[(size_t)i * n + j] + inv[(size_t)j * n + i]);
            inv[(size_t)i * n + j] = v;
            inv[(size_t)j * n + i] = v;
        }
    return true;
}

// Core layout from the core's strictly-lower SSS triangle (crp/ccol/cval), its diagonal,
// the full-system row of each core row (ascending) and the caller's row groups
// (ranges[2 b], ranges[2 b + 1]) of the full system.  G work-groups at most.
bool build_core_host(int m, const std::vector<int64_t> &crp, const std::vector<int64_t> &ccol,
                     const std::vector<double> &cval, const std::vector<double> &cdiag,
                     const std::vector<int> &orig, int64_t nblocks, const int64_t *ranges,
                     int max_groups, CoreHost &H, const char **why) {
    H.m = m;
    // full symmetric CSR, columns ascending within each row
    std::vector<int> cnt((size_t)m + 1, 0);
    for (int r = 0; r < m; r++)
        for (int64_t k = crp[r]; k < crp[r + 1]; k++) {
            cnt[r + 1]++;
            cnt[(size_t)ccol[k] + 1]++;
        }
    H.rp.assign((size_t)m + 1, 0);
    for (int r = 0; r < m; r++) H.rp[r + 1] = H.rp[r] + cnt[r + 1];
    H.col.assign((size_t)H.rp[m], 0);
    H.val.assign((size_t)H.rp[m], 0.0);
    std::vector<int> fill(H.rp.begin(), H.rp.end() - 1);
    for (int r = 0; r < m; r++)
        for (int64_t k = crp[r]; k < crp[r + 1]; k++) {
            const int c = (int)ccol[k];
            H.col[fill[r]] = c;
            H.val[fill[r]++] = cval[k];
            H.col[fill[c]] = r;
            H.val[fill[c]++] = cval[k];
        }
    std::vector<std::pair<int, double>> row;
    for (int r = 0; r < m; r++) {
        row.clear();
        for (int k = H.rp[r]; k < H.rp[r + 1]; k++) row.emplace_back(H.col[k], H.val[k]);
        std::sort(row.begin(), row.end());
        for (size_t q = 0; q < row.size(); q++) {
            H.col[(size_t)H.rp[r] + q] = row[q].first;
            H.val[(size_t)H.rp[r] + q] = row[q].second;
        }
    }
    H.diag = cdiag;
    // row groups -> core row ranges (split at kCoreMaxBlock rows)
    std::vector<std::pair<int, int>> groups;
    for (int64_t b = 0; b < nblocks; b++) {
        const int64_t s0 = ranges[2 * b], s1 = ranges[2 * b + 1];
        if (s0 < 0 || s1 < s0) {
            *why = "a row group has a negative start or ends before it starts";
            return false;
        }
        const int q0 = (int)(std::lower_bound(orig.begin(), orig.end(), s0) - orig.begin());
        const int q1 = (int)(std::lower_bound(orig.begin(), orig.end(), s1) - orig.begin());
        for (int q = q0; q < q1; q += kCoreMaxBlock) {
            const int e = std::min(q1, q + kCoreMaxBlock);
            if (e - q >= 2) groups.emplace_back(q, e);
        }
    }
    std::sort(groups.begin(), groups.end());
    for (size_t g = 1; g < groups.size(); g++)
        if (groups[g].first < groups[g - 1].second) {
            *why = "row groups overlap";
            return false;
        }
    // dense inverses of the diagonal blocks (in parallel over blocks)
    const size_t ng = groups.size();
    std::vector<long long> off(ng + 1, 0);
    for (size_t g = 0; g < ng; g++) {
        const long long nb = groups[g].second - groups[g].first;
        off[g + 1] = off[g] + nb * nb;
    }
    std::vector<double> inv((size_t)off[ng]);
    std::vector<char> ok(ng, 0);
    auto work = [&](size_t g0, size_t g1) {
        std::vector<double> a;
        for (size_t g = g0; g < g1; g++) {
            const int q0 = groups[g].first, nb = groups[g].second - groups[g].first;
            a.assign((size_t)nb * nb, 0.0);
            for (int i = 0; i < nb; i++) {
                a[(size_t)i * nb + i] = H.diag[q0 + i];
                for (int k = H.rp[q0 + i]; k < H.rp[q0 + i + 1]; k++) {
                    const int c = H.col[k] - q0;
                    if (c >= 0 && c < nb) a[(size_t)i * nb + c] = H.val[k];
                }
            }
            ok[g] = dense_spd_inverse(nb, a, &inv[(size_t)off[g]]) ? 1 : 0;
        }
    };
    {
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        const size_t per = (ng + nt - 1) / nt;
        for (unsigned t = 0; t < nt; t++) {
            const size_t g0 = t * per, g1 = std::min(ng, g0 + per);
            if (g0 < g1) th.emplace_back(work, g0, g1);
        }
        for (auto &t : th) t.join();
    }
    // keep the blocks that factorised (a block that did not stays on Jacobi)
    H.blk_of.assign((size_t)m, -1);
    H.blk_row0.clear();
    H.blk_n.clear();
    H.blk_off.clear();
    H.binv.clear();
    for (size_t g = 0; g < ng; g++) {
        if (!ok[g]) continue;
        const int q0 = groups[g].first, nb = groups[g].second - groups[g].first;
        const int id = (int)H.blk_row0.size();
        for (int i = 0; i < nb; i++) H.blk_of[q0 + i] = id;
        H.blk_row0.push_back(q0);
        H.blk_n.push_back(nb);
        H.blk_off.push_back((long long)H.binv.size());
        H.binv.insert(H.binv.end(), inv.begin() + off[g], inv.begin() + off[g + 1]);
    }
    // work-group segments: equal shares of (rows + entries + dense block work), cut only
    // between rows that are not inside the same block
    H.G = std::max(1, std::min(max_groups, (m + kCoreRowsPerGroup - 1) / kCoreRowsPerGroup));
    std::vector<double> cost((size_t)m);
    double total = 0.0;
    for (int r = 0; r < m; r++) {
        cost[r] = 4.0 + (H.rp[r + 1] - H.rp[r]);
        const int b = H.blk_of[r];
        if (b >= 0) cost[r] += 0.25 * H.blk_n[b];
        total += cost[r];
    }
    H.seg.assign(1, 0);
    double acc = 0.0;
    for (int r = 0; r < m; r++) {
        acc += cost[r];
        const bool cut_ok = r + 1 == m || H.blk_of[r] < 0 || H.blk_of[r + 1] != H.blk_of[r];
        if (cut_ok && (int)H.seg.size() < H.G && acc >= total * (double)H.seg.size() / H.G)
            H.seg.push_back(r + 1);
    }
    if (H.seg.back() != m) H.seg.push_back(m);
    H.G = (int)H.seg.size() - 1;
    H.wblk.assign((size_t)H.G + 1, 0);
    int b = 0;
    for (int g = 0; g < H.G; g++) {
        H.wblk[g] = b;
        while (b < (int)H.blk_row0.size() && H.blk_row0[b] < H.seg[g + 1]) b++;
    }
    H.wblk[H.G] = (int)H.blk_row0.size();
    return true;
}

// The preconditioner of the layout on the host: out = M^-1 v.
void host_core_precond(const CoreHost &H, const double *v, double *out) {
    for (int i = 0; i < H.m; i++)
        if (H.blk_of[i] < 0) out[i] = (1.0 / H.diag[i]) * v[i];
    for (size_t b = 0; b < H.blk_row0.size(); b++) {
        const int q0 = H.blk_row0[b], nb = H.blk_n[b];
        const double *M = &H.binv[(size_t)H.blk_off[b]];
        for (int i = 0; i < nb; i++) {
            double z = 0.0;
            for (int j = 0; j < nb; j++) z = fma(M[(size_t)j * nb + i], v[q0 + j], z);
            out[q0 + i] = z;
        }
    }
}

void host_core_apply(const CoreHost &H, const double *v, double *out) {
    for (int i = 0; i < H.m; i++) {
        double acc = H.diag[i] * v[i];
        for (int k = H.rp[i]; k < H.rp[i + 1]; k++) acc = fma(H.val[k], v[H.col[k]], acc);
        out[i] = acc;
    }
}

// The kernel's recurrences on the host (sequential sums): solves A x = rhs from x = 0 and
// returns the iteration count (-1 on breakdown).
long long host_core_cg(const CoreHost &H, const double *rhs, double tol, double ref2,
                       long long max_iter, double *x, int *status) {
    const int m = H.m;
    std::vector<double> r(rhs, rhs + m), u(m), w(m), p(m, 0.0), s(m, 0.0);
    std::fill(x, x + m, 0.0);
    const double stop2 = tol * tol * ref2;
    auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
        double t = 0.0;
        for (int i = 0; i < m; i++) t += a[i] * b[i];
        return t;
    };
    host_core_precond(H, r.data(), u.data());
    double gamma = dot(r, u), rr = dot(r, r);
    *status = CICY_OK;
    if (rr <= stop2) return 0;
    host_core_apply(H, u.data(), w.data());
    double delta = dot(w, u);
    double alpha = gamma / delta, beta = 0.0;
    long long k = 0;
    while (true) {
        for (int i = 0; i < m; i++) {
            p[i] = fma(beta, p[i], u[i]);
            s[i] = fma(beta, s[i], w[i]);
            x[i] = fma(alpha, p[i], x[i]);
            r[i] = fma(-alpha, s[i], r[i]);
        }
        host_core_precond(H, r.data(), u.data());
        const double g2 = dot(r, u);
        rr = dot(r, r);
        k++;
        if (rr <= stop2) return k;
        if (!std::isfinite(rr) || !(g2 > 0.0)) {
            *status = CICY_BREAKDOWN;
            return k;
        }
        if (k >= max_iter) {
            *status = CICY_MAXITER;
            return k;
        }
        host_core_apply(H, u.data(), w.data());
        delta = dot(w, u);
        beta = g2 / gamma;
        const double den = delta - beta * g2 / alpha;
        if (!(den > 0.0) || !std::isfinite(den)) {
            *status = CICY_BREAKDOWN;
            return k;
        }
        alpha = g2 / den;
        gamma = g2;
    }
}

// ------------------------------------------------------------------------------------
// the persistent kernel
// ------------------------------------------------------------------------------------

// One chunk of iterations.  All work-groups hold the same scalars (the grid sums are
// bit-identical everywhere), so they take the same branches and leave together.
__global__ __launch_bounds__(kCoreThreads) void k_core_cg(CoreDev a, CoreState *st, int chunk) {
    __shared__ double rb[kCoreWaves][kCoreMaxBlock];
    __shared__ double lds[2 * kCoreWaves];
    __shared__ int flag;
    const int g = blockIdx.x, tid = threadIdx.x;
    const int wave = tid / kWave, lane = tid & (kWave - 1);
    const int r0 = a.seg[g], r1 = a.seg[g + 1];
    const int b0 = a.wblk[g], b1 = a.wblk[g + 1];
    long long k = st->k;
    const long long max_iter = st->max_iter;
    double gamma = st->gamma, alpha = st->alpha, beta = st->beta, rr = st->rr;
    const double stop2 = st->stop2;
    int done = st->done;
    const bool started = st->started != 0;
    unsigned bar = 0;

    // u = M^-1 r for the rows of this work-group (after the vector update when UPDATE),
    // partial r.u and r.r
    auto phase_u = [&](bool update, double &pg, double &pr) {
        pg = 0.0;
        pr = 0.0;
        for (int i = r0 + tid; i < r1; i += kCoreThreads) {
            if (a.blk_of[i] >= 0) continue;
            double ri = a.r[i];
            if (update) {
                const double pi = fma(beta, a.p[i], ld_coh(&a.u[i]));
                const double si = fma(beta, a.s[i], a.w[i]);
                a.p[i] = pi;
                a.s[i] = si;
                a.x[i] = fma(alpha, pi, a.x[i]);
                ri = fma(-alpha, si, ri);
                a.r[i] = ri;
            }
            const double ui = (1.0 / a.diag[i]) * ri;
            st_coh(&a.u[i], ui);
            pg += ri * ui;
            pr += ri * ri;
        }
        for (int base = b0; base < b1; base += kCoreWaves) {
            const int bi = base + wave;
            const bool act = bi < b1;
            const int nbk = act ? a.blk_n[bi] : 0;
            const int q0 = act ? a.blk_row0[bi] : 0;
            for (int j = lane; j < nbk; j += kWave) {
                const int i = q0 + j;
                double ri = a.r[i];
                if (update) {
                    const double pi = fma(beta, a.p[i], ld_coh(&a.u[i]));
                    const double si = fma(beta, a.s[i], a.w[i]);
                    a.p[i] = pi;
                    a.s[i] = si;
                    a.x[i] = fma(alpha, pi, a.x[i]);
                    ri = fma(-alpha, si, ri);
                    a.r[i] = ri;
                }
                rb[wave][j] = ri;
            }
            __syncthreads();
            if (act) {
                const double *M = a.binv + a.blk_off[bi];
                const bool h0 = lane < nbk, h1 = lane + kWave < nbk;
                double z0 = 0.0, z1 = 0.0;
                for (int j = 0; j < nbk; j++) {
                    const double rj = rb[wave][j];
                    const double *Mj = M + (size_t)j * nbk;
                    if (h0) z0 = fma(Mj[lane], rj, z0);
                    if (h1) z1 = fma(Mj[lane + kWave], rj, z1);
                }
                if (h0) {
                    const double ri = rb[wave][lane];
                    st_coh(&a.u[q0 + lane], z0);
                    pg += ri * z0;
                    pr += ri * ri;
                }
                if (h1) {
                    const double ri = rb[wave][lane + kWave];
                    st_coh(&a.u[q0 + lane + kWave], z1);
                    pg += ri * z1;
                    pr += ri * ri;
                }
            }
            __syncthreads();
        }
    };
    // w = A u for the rows of this work-group, partial w.u
    auto phase_w = [&]() {
        double pd = 0.0;
        for (int i = r0 + tid; i < r1; i += kCoreThreads) {
            const double ui = ld_coh(&a.u[i]);
            double acc = a.diag[i] * ui;
            const int e1 = a.rp[i + 1];
            for (int e = a.rp[i]; e < e1; e++) acc = fma(a.val[e], ld_coh(&a.u[a.col[e]]), acc);
            a.w[i] = acc;
            pd += acc * ui;
        }
        return pd;
    };

    if (!done && !started) {
        double v[2];
        phase_u(false, v[0], v[1]);
        if (!grid_sum<2, kCoreWaves>(v, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
        } else {
            gamma = v[0];
            rr = v[1];
            if (rr <= stop2) {
                done = 1;
            } else {
                double d[1] = {phase_w()};
                if (!grid_sum<1, kCoreWaves>(d, a.G, a.part, a.ctl, bar++, lds, &flag)) {
                    done = 4;
                } else {
                    alpha = gamma / d[0];
                    beta = 0.0;
                    if (!(d[0] > 0.0) || !isfinite(alpha)) done = 3;
                }
            }
        }
    }
    for (int it = 0; it < chunk && !done; it++) {
        double v[2];
        phase_u(true, v[0], v[1]);
        if (!grid_sum<2, kCoreWaves>(v, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
            break;
        }
        const double g2 = v[0];
        rr = v[1];
        k++;
        if (rr <= stop2) {
            done = 1;
            break;
        }
        if (!isfinite(rr) || !(g2 > 0.0)) {
            done = 3;
            break;
        }
        if (k >= max_iter) {
            done = 2;
            break;
        }
        double d[1] = {phase_w()};
        if (!grid_sum<1, kCoreWaves>(d, a.G, a.part, a.ctl, bar++, lds, &flag)) {
            done = 4;
            break;
        }
        beta = g2 / gamma;
        const double den = d[0] - beta * g2 / alpha;
        if (!(den > 0.0) || !isfinite(den)) {
            done = 3;
            break;
        }
        alpha = g2 / den;
        gamma = g2;
    }
    if (g == 0 && tid == 0) {
        st->k = k;
        st->gamma = gamma;
        st->alpha = alpha;
        st->beta = beta;
        st->rr = rr;
        st->done = done;
        st->started = 1;
    }
}

template <typename T>
bool up_vec(cicy_ctx *ctx, cicy_core *z, T **dst, const std::vector<T> &src) {
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    if (!hip_ok(ctx, hipMalloc((void **)dst, bytes), "hipMalloc (core)")) return false;
    z->bytes += bytes;
    if (src.empty()) return true;
    return hip_ok(ctx, hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T),
                                      hipMemcpyHostToDevice, ctx->stream),
                  "H2D (core)");
}

bool zeros(cicy_ctx *ctx, cicy_core *z, double **dst, size_t count) {
    const size_t bytes = std::max<size_t>(1, count) * sizeof(double);
    if (!hip_ok(ctx, hipMalloc((void **)dst, bytes), "hipMalloc (core)")) return false;
    z->bytes += bytes;
    return hip_ok(ctx, hipMemsetAsync(*dst, 0, bytes, ctx->stream), "memset (core)");
}

}  // namespace

void core_free(cicy_core *z) {
    if (!z) return;
    int *ip[] = {z->seg, z->rp, z->col, z->blk_of, z->wblk, z->blk_row0, z->blk_n};
    for (int *q : ip)
        if (q) (void)hipFree(q);
    double *dp[] = {z->val, z->diag, z->binv, z->x, z->r, z->u, z->w, z->p, z->s, z->part};
    for (double *q : dp)
        if (q) (void)hipFree(q);
    if (z->blk_off) (void)hipFree(z->blk_off);
    if (z->ctl) (void)hipFree(z->ctl);
    if (z->st) (void)hipFree(z->st);
    delete z;
}

int core_info(const cicy_core *z, int what) {
    if (!z) return 0;
    switch (what) {
        case 0: return z->nb;
        case 1: return z->blocked_rows;
        case 2: return z->G;
        case 3: return (int)z->inv_entries;
        default: return 0;
    }
}

cicy_core *core_setup(cicy_ctx *ctx, int m, const std::vector<int64_t> &crp,
                      const std::vector<int64_t> &ccol, const std::vector<double> &cval,
                      const std::vector<double> &cdiag, const std::vector<int> &orig,
                      int64_t nblocks, const int64_t *ranges) {
    int cus = 0, per_cu = 0;
    if (!hip_ok(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                           ctx->device),
                "hipDeviceGetAttribute") ||
        !hip_ok(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &per_cu, reinterpret_cast<const void *>(k_core_cg), kCoreThreads, 0),
                "hipOccupancyMaxActiveBlocksPerMultiprocessor"))
        return nullptr;
    if (per_cu < 1) {
        set_error(ctx, "core CG kernel cannot be resident");
        return nullptr;
    }
    CoreHost H;
    const char *why = nullptr;
    if (!build_core_host(m, crp, ccol, cval, cdiag, orig, nblocks, ranges, cus, H, &why)) {
        set_error(ctx, "core blocks: %s", why);
        return nullptr;
    }
    cicy_core *z = new cicy_core();
    z->m = m;
    z->G = H.G;
    z->nb = (int)H.blk_row0.size();
    z->nnz2 = (long long)H.col.size();
    z->inv_entries = (long long)H.binv.size();
    for (int v : H.blk_n) z->blocked_rows += v;
    bool ok = up_vec(ctx, z, &z->seg, H.seg) && up_vec(ctx, z, &z->rp, H.rp) &&
              up_vec(ctx, z, &z->col, H.col) && up_vec(ctx, z, &z->val, H.val) &&
              up_vec(ctx, z, &z->diag, H.diag) && up_vec(ctx, z, &z->blk_of, H.blk_of) &&
              up_vec(ctx, z, &z->wblk, H.wblk) && up_vec(ctx, z, &z->blk_row0, H.blk_row0) &&
              up_vec(ctx, z, &z->blk_n, H.blk_n) && up_vec(ctx, z, &z->blk_off, H.blk_off) &&
              up_vec(ctx, z, &z->binv, H.binv) && zeros(ctx, z, &z->x, m) &&
              zeros(ctx, z, &z->r, m) && zeros(ctx, z, &z->u, m) && zeros(ctx, z, &z->w, m) &&
              zeros(ctx, z, &z->p, m) && zeros(ctx, z, &z->s, m) &&
              zeros(ctx, z, &z->part, 4 * (size_t)H.G) &&
              hip_ok(ctx, hipMalloc((void **)&z->ctl, sizeof(GridCtl)), "hipMalloc (core)") &&
              hip_ok(ctx, hipMalloc((void **)&z->st, sizeof(CoreState)), "hipMalloc (core)") &&
              hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!ok) {
        core_free(z);
        return nullptr;
    }
    return z;
}

// Solves A x = rhs (core rows, on the device) from x = 0 until ||r|| <= tol sqrt(ref2);
// the solution goes to x_out.  stats: iters and status.
bool solve_core(cicy_ctx *ctx, cicy_core *z, const double *rhs, double tol, long long max_iter,
                double ref2, double *x_out, cicy_solve_stats *stats) {
    hipStream_t stream = ctx->stream;
    const size_t vb = (size_t)z->m * sizeof(double);
    const double max_seconds = [] {
        const char *e = getenv("CICY_MAX_SOLVE_SECONDS");
        return e ? atof(e) : 0.0;
    }();
    auto t0 = std::chrono::steady_clock::now();
    CoreState h;
    memset(&h, 0, sizeof(h));
    h.max_iter = max_iter > 0 ? max_iter : 1;
    h.stop2 = tol * tol * ref2;
    bool ok = hip_ok(ctx, hipMemcpyAsync(z->r, rhs, vb, hipMemcpyDeviceToDevice, stream), "D2D") &&
              hip_ok(ctx, hipMemsetAsync(z->x, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemsetAsync(z->p, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemsetAsync(z->s, 0, vb, stream), "memset") &&
              hip_ok(ctx, hipMemcpyAsync(z->st, &h, sizeof(h), hipMemcpyHostToDevice, stream),
                     "H2D");
    CoreDev a{z->m, z->G, z->seg, z->rp, z->col, z->val, z->diag, z->blk_of, z->wblk,
              z->blk_row0, z->blk_n, z->blk_off, z->binv, z->x, z->r, z->u, z->w, z->p, z->s,
              z->ctl, z->part};
    while (ok && !h.done) {
        int chunk = kCoreChunk;
        void *args[] = {(void *)&a, (void *)&z->st, (void *)&chunk};
        ok = hip_ok(ctx, hipMemsetAsync(z->ctl, 0, sizeof(GridCtl), stream), "memset") &&
             hip_ok(ctx, hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_core_cg),
                                                    dim3(z->G), dim3(kCoreThreads), args, 0,
                                                    stream),
                    "hipLaunchCooperativeKernel (core CG)") &&
             d2h_sync(ctx, &h, z->st, sizeof(h));
        if (ok && !h.done && max_seconds > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
                max_seconds)
            break;
    }
    if (ok) ok = hip_ok(ctx, hipMemcpyAsync(x_out, z->x, vb, hipMemcpyDeviceToDevice, stream), "D2D");
    memset(stats, 0, sizeof(*stats));
    stats->iters = h.k;
    stats->status = !ok            ? CICY_DEVICE_ERROR
                    : h.done == 1 ? CICY_OK
                    : h.done == 3 ? CICY_BREAKDOWN
                    : h.done == 4 ? CICY_DEVICE_ERROR
                                  : CICY_MAXITER;
    if (ok && h.done == 4) {
        set_error(ctx, "core CG: grid barrier timed out (work-groups not co-resident)");
        return false;
    }
    return ok;
}

}  // namespace cicy

// ------------------------------------------------------------------------------------
// C-ABI test hook: the layout and the kernel's recurrences on the host
// ------------------------------------------------------------------------------------

extern "C" int cicy_host_core_block_cg(int64_t m, const int64_t *row_ptr, const int64_t *col_ind,
                                       const double *values, const double *diagonal,
                                       int64_t nnz, int64_t nblocks, const int64_t *ranges,
                                       const double *rhs, double tol, double ref2,
                                       int64_t max_iter, double *x, int64_t *iters,
                                       int64_t *blocked_rows) {
    using namespace cicy;
    if (m <= 0 || !row_ptr || !diagonal || !rhs || !x || row_ptr[m] != nnz) return 0;
    std::vector<int64_t> crp(row_ptr, row_ptr + m + 1), ccol(col_ind, col_ind + nnz);
    std::vector<double> cval(values, values + nnz), cdiag(diagonal, diagonal + m);
    std::vector<int> orig((size_t)m);
    for (int64_t i = 0; i < m; i++) orig[i] = (int)i;
    CoreHost H;
    const char *why = nullptr;
    if (!build_core_host((int)m, crp, ccol, cval, cdiag, orig, nblocks, ranges, 256, H, &why))
        return 0;
    int status = CICY_OK;
    const long long k = host_core_cg(H, rhs, tol, ref2, max_iter > 0 ? max_iter : 10 * m + 100,
                                     x, &status);
    if (iters) *iters = k;
    if (blocked_rows) {
        long long br = 0;
        for (int v : H.blk_n) br += v;
        *blocked_rows = br;
    }
    return status == CICY_OK ? 1 : 0;
}
