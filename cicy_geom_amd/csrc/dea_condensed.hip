// dea_condensed.hip -- exact static condensation of the DEA system before CG.
//
// The reference solves the whole SSS system with Jacobi-PCG
// (Direct_Electrifying.cpp:1538-1659).  In a CNT network most rows are interior points of
// a tube: their only couplings are the two neighbouring points of the same tube, which the
// tube-by-tube numbering puts on the sub-diagonal (Direct_Electrifying.cpp:286-331 assigns
// the nodes of a CNT consecutively).  Such a row has no junction (non-sub-diagonal)
// entries.  Here those rows are eliminated exactly before CG:
//
//   * anchors J = rows with at least one junction entry; every other row lies on a "path",
//     a maximal run of consecutive non-anchor rows joined by sub-diagonal entries, coupled
//     to at most one anchor on each side (L = s-1, R = e+1) through the band values;
//   * the path block T of A is tridiagonal and is LU-factorised once (Thomas algorithm,
//     no pivoting: T is a diagonally dominant M-matrix); the Schur complement
//     S = A_JJ - A_JI T^-1 A_IJ then gains, per path, a diagonal term at L and R and one
//     coupling (R, L) -- series elimination of the path;
//   * a solve A x = b in correction form (x = x0 + d, A d = r0) condenses the right-hand
//     side (y = T^-1 r0_I by one sweep per path, r0'_J = r0_J - A_JI y), runs the FP64
//     Jacobi-PCG on S d_J = r0'_J, and recovers d_I = T^-1 (r0_I - A_IJ d_J) by one more
//     sweep per path.  The interior equations then hold to rounding, so the full residual
//     equals the reduced one and the reduced solve stops on the full system's reference
//     norm (||b|| for CICY_REL_B, ||r0|| for CICY_REL_R0).  The full residual is checked
//     afterwards with the FP64 SpMV of the whole system and the correction is repeated
//     from it if needed.
// The anchors keep their order, so consecutive anchors of one tube are adjacent rows of S
// and the path couplings land on its sub-diagonal: S is again band + remainder.
#include "dea_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>

using namespace cicy;

#include <vector>


namespace cicy {
namespace {

constexpr int kCondThreads = 256;
constexpr int kFloating = -2;  // pl marker of a floating path (same value as kFloatingPath)
constexpr int kPathChunk = 8;  // path rows whose loads a sweep issues together

// One sweep of T z = f per path, f = rhs on the path rows, minus the couplings to the
// anchors when x is given (back-substitution): forward elimination with the stored
// multipliers, then back-substitution with the pivots.  One thread per path.
__global__ __launch_bounds__(kCondThreads) void k_path_solve(
    int npath, const int *__restrict__ ps, const int *__restrict__ pe,
    const int *__restrict__ pl, const int *__restrict__ pr, const double *__restrict__ lfac,
    const double *__restrict__ mfac, const double *__restrict__ band,
    const double *__restrict__ rhs, const double *__restrict__ xanchor, double *__restrict__ z) {
    const int k = blockIdx.x * kCondThreads + threadIdx.x;
    if (k >= npath) return;
    const int s = ps[k], e = pe[k];
    if (pl[k] == kFloating) {
        // singular chain: grounded at its last row, then the zero-mean solution
        double g = rhs[s];
        z[s] = g;
        for (int i = s + 1; i < e; i++) {
            g = rhs[i] - lfac[i] * g;
            z[i] = g;
        }
        double zn = 0.0, mean = 0.0;
        z[e] = 0.0;
        for (int i = e - 1; i >= s; i--) {
            zn = (z[i] - band[i + 1] * zn) / mfac[i];
            z[i] = zn;
            mean += zn;
        }
        mean /= (double)(e - s + 1);
        for (int i = s; i <= e; i++) z[i] -= mean;
        return;
    }
    // the same recurrences with the loads of kPathChunk rows issued before their arithmetic
    // (a path can be a whole tube: its rows would otherwise be a chain of dependent misses)
    double g = rhs[s];
    if (xanchor && pl[k] >= 0) g -= band[s] * xanchor[pl[k]];
    const double fe = (xanchor && pr[k] >= 0) ? band[e + 1] * xanchor[pr[k]] : 0.0;
    if (s == e) g -= fe;
    z[s] = g;
    for (int i0 = s + 1; i0 <= e; i0 += kPathChunk) {
        double f[kPathChunk], l[kPathChunk];
#pragma unroll
        for (int t = 0; t < kPathChunk; t++) {
            const int i = min(i0 + t, e);
            f[t] = rhs[i];
            l[t] = lfac[i];
        }
#pragma unroll
        for (int t = 0; t < kPathChunk; t++) {
            const int i = i0 + t;
            if (i > e) break;
            const double fi = i == e ? f[t] - fe : f[t];
            g = fi - l[t] * g;
            z[i] = g;
        }
    }
    double zn = z[e] / mfac[e];
    z[e] = zn;
    for (int i0 = e - 1; i0 >= s; i0 -= kPathChunk) {
        double zz[kPathChunk], bb[kPathChunk], mm[kPathChunk];
#pragma unroll
        for (int t = 0; t < kPathChunk; t++) {
            const int i = max(i0 - t, s);
            zz[t] = z[i];
            bb[t] = band[i + 1];
            mm[t] = mfac[i];
        }
#pragma unroll
        for (int t = 0; t < kPathChunk; t++) {
            const int i = i0 - t;
            if (i < s) break;
            zn = (zz[t] - bb[t] * zn) / mm[t];
            z[i] = zn;
        }
    }
}

// Reduced right-hand side: r'_j = rhs[a] - a(a, a-1) y[a-1] - a(a+1, a) y[a+1] over the
// paths that end next to anchor a.
__global__ __launch_bounds__(kCondThreads) void k_anchor_rhs(int m, int n,
                                                             const int *__restrict__ anchors,
                                                             const int *__restrict__ anchor_of,
                                                             const double *__restrict__ band,
                                                             const double *__restrict__ rhs,
                                                             const double *__restrict__ y,
                                                             double *__restrict__ out) {
    const int j = blockIdx.x * kCondThreads + threadIdx.x;
    if (j >= m) return;
    const int a = anchors[j];
    double v = rhs[a];
    if (a > 0 && anchor_of[a - 1] < 0 && band[a] != 0.0) v -= band[a] * y[a - 1];
    if (a + 1 < n && anchor_of[a + 1] < 0 && band[a + 1] != 0.0) v -= band[a + 1] * y[a + 1];
    out[j] = v;
}

__global__ __launch_bounds__(kCondThreads) void k_scatter(int m, const int *__restrict__ anchors,
                                                          const double *__restrict__ xj,
                                                          double *__restrict__ x) {
    const int j = blockIdx.x * kCondThreads + threadIdx.x;
    if (j < m) x[anchors[j]] = xj[j];
}

// one round of the forward sweep: f_u -= a_uk f_k / d_k for every neighbour u of an
// eliminated node k.  Nodes of a round are pairwise non-adjacent, so f_k is final; a node u
// next to several of them gathers their contributions itself, in record order (one thread
// per receiving node, no atomics), so the sweep is deterministic and equals the host form
// bit for bit.
__global__ __launch_bounds__(kCondThreads) void k_series_fwd(int b, int e, const int *__restrict__ rv_node,
                                                             const int *__restrict__ rv_off,
                                                             const int *__restrict__ rv_rec,
                                                             const double *__restrict__ rv_a,
                                                             const int *__restrict__ node,
                                                             const double *__restrict__ dd,
                                                             double *__restrict__ f) {
    const int j = b + blockIdx.x * kCondThreads + threadIdx.x;
    if (j >= e) return;
    const int u = rv_node[j];
    double r = f[u];
    for (int c = rv_off[j]; c < rv_off[j + 1]; c++) {
        const int k = rv_rec[c];
        r -= rv_a[c] * (f[node[k]] / dd[k]);
    }
    f[u] = r;
}

// one round of the back-substitution: x_k = (f_k - sum_u a_uk x_u) / d_k
__global__ __launch_bounds__(kCondThreads) void k_series_back(int b, int e, const int *__restrict__ node,
                                                              const int *__restrict__ off,
                                                              const int *__restrict__ nbr,
                                                              const double *__restrict__ a,
                                                              const double *__restrict__ dd,
                                                              const double *__restrict__ f,
                                                              double *__restrict__ x) {
    const int k = b + blockIdx.x * kCondThreads + threadIdx.x;
    if (k >= e) return;
    double r = f[node[k]];
    for (int i = off[k]; i < off[k + 1]; i++) r -= a[i] * x[nbr[i]];
    x[node[k]] = r / dd[k];
}

__global__ __launch_bounds__(kCondThreads) void k_gather(int mc, const int *__restrict__ idx,
                                                         const double *__restrict__ src,
                                                         double *__restrict__ dst) {
    const int i = blockIdx.x * kCondThreads + threadIdx.x;
    if (i < mc) dst[i] = src[idx[i]];
}

// x += d; r = rhs - ad; out = (r.r, ...) via the fence-free reduction
__global__ __launch_bounds__(kThreads) void k_correct(int n, double *__restrict__ x,
                                                      const double *__restrict__ d,
                                                      const double *__restrict__ rhs,
                                                      const double *__restrict__ ad,
                                                      double *__restrict__ r,
                                                      double *__restrict__ out,
                                                      unsigned int *__restrict__ counter,
                                                      RedBuf rbuf) {
    __shared__ double red[8];
    __shared__ int flag;
    double v[1] = {0.0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        x[i] += d[i];
        const double ri = rhs[i] - ad[i];
        r[i] = ri;
        v[0] += ri * ri;
    }
    Blk k;
    k.b = blockIdx.x;
    k.sys = 0;
    k.r0 = 0;
    k.nrows = 0;
    k.first = 0;
    k.count = gridDim.x;
    double o[1];
    if (reduce_sys<1>(v, k, rbuf, counter, red, &flag, o) && threadIdx.x == 0) out[0] = o[0];
}

template <typename T>
bool calloc_dev(cicy_ctx *ctx, T **p, size_t count) {
    *p = static_cast<T *>(ctx->pool.get(std::max<size_t>(count, 1) * sizeof(T)));
    if (!*p) {
        set_error(ctx, "hipMalloc of %zu bytes failed (condensed system)", count * sizeof(T));
        return false;
    }
    return hip_ok(ctx, hipMemsetAsync(*p, 0, std::max<size_t>(count, 1) * sizeof(T), ctx->stream),
                  "memset");
}

template <typename T>
bool up(cicy_ctx *ctx, T *dst, const std::vector<T> &src) {
    if (src.empty()) return true;
    return hip_ok(ctx, hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T),
                                      hipMemcpyHostToDevice, ctx->stream),
                  "H2D (condensed system)");
}

int blocks(long long n, int threads) { return (int)std::max<long long>(1, (n + threads - 1) / threads); }

}  // namespace

void condensed_free(cicy_cond *c) {
    if (!c) return;
    if (c->red) cicy_sys_destroy(c->red);
    if (c->blk) core_free(c->blk);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    void *ptrs[] = {c->ps, c->pe, c->pl, c->pr, c->anchor_of, c->anchors, c->lfac, c->mfac,
                    c->band, c->y, c->delta, c->rhs, c->norm, c->counter, c->part, c->gpart,
                    c->gcount, c->e_node, c->e_off, c->e_nbr, c->e_a, c->e_d,
                    c->core_nodes, c->fS, c->xS, c->rv_node, c->rv_off, c->rv_rec, c->rv_a,
                    c->core_b, c->core_x};
    for (void *p : ptrs) dev_release(c->ctx, p);
    condensed_free_device(c);
    delete c;
}

int condensed_info(const cicy_cond *c, int what) {
    if (!c) return 0;
    switch (what) {
        case 0: return c->m;                                 // rows after the path stage
        case 1: return (int)c->nnz_red;
        case 2: return c->mc;                                // core rows (the CG system)
        case 3: return (int)c->nnz_core;
        case 4: return (int)c->round_off.size() - 1;         // series rounds
        case 5: return core_info(c->blk, 1);                 // core rows in dense blocks
        case 6: return core_info(c->blk, 0);                 // dense blocks
        case 7: return core_info(c->blk, 2);                 // core CG work-groups
        case 8: return core_info(c->blk, 3);                 // dense block entries
        case 9: return c->blk ? 1 : 0;                       // block core built
        default: return 0;
    }
}

// Host form of the condensation: anchors, paths with their LU factors, the reduced SSS
// matrix S.  Returns false with *why set when condensing is not worth it, or with an error
// (failed) when a pivot is not positive.
// pl value of a path that touches no anchor and whose block is singular (floating tube)
constexpr int kFloatingPath = -2;

struct CondHost {
    int n = 0, m = 0;
    std::vector<int> anchor_of, anchors, ps, pe, pl, pr;
    std::vector<double> lfac, mfac, band;
    std::vector<int64_t> rrp, rcol;
    std::vector<double> rval, sdiag;
};

bool build_condensed_host(cicy_ctx *ctx, const BlockedHost &h, const BandHost &bh, CondHost &C,
                          const char **why, bool *failed) {
    *failed = false;
    const int n = h.n;
    C.n = n;
    if (n == 0) {
        *why = "empty system";
        return false;
    }
    StageClock clk;
    std::vector<int> &anchor_of = C.anchor_of, &anchors = C.anchors;
    // anchors: the rows with an entry off the sub-diagonal (a half-entry of the remainder),
    // numbered in row order; per-thread counts then a prefix, on host threads
    anchor_of.resize((size_t)n);
    const int nat = host_threads(n, 1 << 16);
    std::vector<int> acount((size_t)nat + 1, 0);
    parallel_for(nat, [&](int t) {
        const int i0 = (int)((long long)n * t / nat), i1 = (int)((long long)n * (t + 1) / nat);
        int c = 0;
        for (int i = i0; i < i1; i++) c += bh.xrp[(size_t)i + 1] > bh.xrp[(size_t)i];
        acount[(size_t)t + 1] = c;
    });
    for (int t = 0; t < nat; t++) acount[(size_t)t + 1] += acount[(size_t)t];
    const int m = acount[(size_t)nat];
    C.m = m;
    anchors.resize((size_t)m);
    parallel_for(nat, [&](int t) {
        const int i0 = (int)((long long)n * t / nat), i1 = (int)((long long)n * (t + 1) / nat);
        int a = acount[(size_t)t];
        for (int i = i0; i < i1; i++) {
            if (bh.xrp[(size_t)i + 1] > bh.xrp[(size_t)i]) {
                anchor_of[i] = a;
                anchors[(size_t)a++] = i;
            } else {
                anchor_of[i] = -1;
            }
        }
    });
    if ((long long)m * 2 > n) {
        *why = "more than half of the rows carry junction entries: nothing worth condensing";
        return false;
    }
    // paths, their LU factors and their Schur complement terms
    std::vector<int> &ps = C.ps, &pe = C.pe, &pl = C.pl, &pr = C.pr;
    std::vector<double> &lfac = C.lfac, &mfac = C.mfac, &sdiag = C.sdiag;
    lfac.resize((size_t)n);
    mfac.resize((size_t)n);
    C.band.resize(bh.band.size());
    parallel_for(nat, [&](int t) {
        const size_t i0 = (size_t)n * t / nat, i1 = (size_t)n * (t + 1) / nat;
        std::fill(lfac.begin() + i0, lfac.begin() + i1, 0.0);
        std::fill(mfac.begin() + i0, mfac.begin() + i1, 1.0);
        std::copy(bh.band.begin() + i0, bh.band.begin() + i1, C.band.begin() + i0);
    });
    C.band[(size_t)n] = bh.band[(size_t)n];
    sdiag.resize((size_t)m);
    for (int j = 0; j < m; j++) sdiag[j] = h.diag[anchors[j]];
    clk.mark("paths: anchors");
    struct Fill {
        int r, c;
        double v;
    };
    const std::vector<double> &band = bh.band;  // [n + 1], band[i] = a(i, i-1)
    // A path starts at a non-anchor row whose predecessor is an anchor or not coupled to it.
    auto path_start = [&](int i) {
        return anchor_of[i] < 0 && (i == 0 || anchor_of[i - 1] >= 0 || band[i] == 0.0);
    };
    // The rows are cut into one range per host thread at anchors or path starts, so no path
    // crosses a cut.  Each thread factorises its paths (lfac/mfac rows are its own) and
    // collects, in row order, its paths, its Schur terms on the anchors' diagonals (an anchor
    // can border paths of two threads) and its fill; these are merged in thread order, which
    // is the sequential order, so the result does not depend on the thread count.
    const int nth = host_threads(n, 1 << 16);
    std::vector<int> cut((size_t)nth + 1, n);
    cut[0] = 0;
    for (int t = 1; t < nth; t++) {
        int b = (int)((long long)n * t / nth);
        b = std::max(b, cut[t - 1]);
        while (b < n && anchor_of[b] < 0 && !path_start(b)) b++;
        cut[t] = b;
    }
    struct Part {
        std::vector<int> ps, pe, pl, pr;
        std::vector<std::pair<int, double>> dsub;  // (anchor, term subtracted from sdiag)
        std::vector<Fill> fill, jfill;  // path fill; couplings of the original matrix
        const char *why = nullptr;
    };
    std::vector<Part> parts((size_t)nth);
    parallel_for(nth, [&](int t) {
        Part &P = parts[(size_t)t];
        std::vector<double> g;
        for (int i = cut[t]; i < cut[t + 1];) {
            if (anchor_of[i] >= 0) {
                i++;
                continue;
            }
            const int s = i;
            int e = s;
            while (e + 1 < n && anchor_of[e + 1] < 0 && band[e + 1] != 0.0) e++;
            const int L = (s > 0 && band[s] != 0.0) ? s - 1 : -1;
            const int R = (e + 1 < n && band[e + 1] != 0.0) ? e + 1 : -1;
            P.ps.push_back(s);
            P.pe.push_back(e);
            P.pl.push_back(L >= 0 ? anchor_of[L] : -1);
            P.pr.push_back(R >= 0 ? anchor_of[R] : -1);
            mfac[s] = h.diag[s];
            for (int k = s + 1; k <= e; k++) {
                lfac[k] = band[k] / mfac[k - 1];
                mfac[k] = h.diag[k] - lfac[k] * band[k];
            }
            // A pivot can only vanish on a path that touches no anchor and no electrode: a
            // floating tube (for instance one whose only junction reaches an electrode and was
            // dropped, Direct_Electrifying.cpp:817-837), singular on its own (its rows of A
            // sum to zero).  Its block is singular but consistent: r0 = b - A x0 = -A x0
            // there.  CG's correction stays orthogonal to the constant null vector, so the
            // tube ends at the mean of its initial guess; the condensed solve gives the same:
            // the path is solved grounded at its last row (the leading block keeps its
            // pivots) and the correction's mean is removed.
            const double pivot_floor = 1e-12 * h.diag[e];
            bool singular = false;
            for (int k = s; k <= e; k++)
                if (!(mfac[k] > pivot_floor) || !std::isfinite(mfac[k])) singular = true;
            if (singular) {
                if (L >= 0 || R >= 0) {
                    P.why = "non-positive pivot on a path coupled to the rest (not an M-matrix)";
                    return;
                }
                for (int k = s; k < e; k++)
                    if (!(mfac[k] > pivot_floor) || !std::isfinite(mfac[k])) {
                        P.why = "non-positive pivot inside a floating path";
                        return;
                    }
                P.pl.back() = kFloatingPath;
                lfac[e] = 0.0;
                mfac[e] = 1.0;
                i = e + 1;
                continue;
            }
            if (L >= 0) {
                // u = T^-1 (a(s, L) e_s): forward sweep from g_s = a(s, L), then
                // back-substitution
                g.assign((size_t)(e - s + 1), 0.0);
                g[0] = band[s];
                for (int k = s + 1; k <= e; k++) g[k - s] = -lfac[k] * g[k - 1 - s];
                double u = g[e - s] / mfac[e];
                const double ue = u;
                for (int k = e - 1; k >= s; k--) u = (g[k - s] - band[k + 1] * u) / mfac[k];
                P.dsub.push_back({anchor_of[L], band[s] * u});  // u is u_s here
                if (R >= 0) P.fill.push_back({anchor_of[R], anchor_of[L], -band[e + 1] * ue});
            }
            if (R >= 0) {
                // v = T^-1 (a(R, e) e_e): only its last entry is needed
                P.dsub.push_back({anchor_of[R], band[e + 1] * (band[e + 1] / mfac[e])});
            }
            i = e + 1;
        }
        // couplings among this range's anchors and earlier anchors in the original matrix
        // (junctions, adjacent anchors)
        // (from the band layout: the band entry to the previous row, then the row's own
        // half-entries in stored order; an entry whose value is exactly zero is absent)
        for (int r = cut[t]; r < cut[t + 1]; r++) {
            if (anchor_of[r] < 0) continue;
            if (r > 0 && band[r] != 0.0 && anchor_of[r - 1] >= 0)
                P.jfill.push_back({anchor_of[r], anchor_of[r - 1], band[r]});
            for (int k = bh.xrp[r]; k < bh.xrp[r + 1]; k++) {
                const int c = bh.xnbr[k];
                if (c < r && anchor_of[c] >= 0)
                    P.jfill.push_back({anchor_of[r], anchor_of[c], bh.xval[k]});
            }
        }
    });
    clk.mark("paths: anchors + factors");
    for (const Part &P : parts)
        if (P.why) {
            *why = P.why;
            return false;
        }
    size_t nfill = 0;
    for (const Part &P : parts) {
        ps.insert(ps.end(), P.ps.begin(), P.ps.end());
        pe.insert(pe.end(), P.pe.begin(), P.pe.end());
        pl.insert(pl.end(), P.pl.begin(), P.pl.end());
        pr.insert(pr.end(), P.pr.begin(), P.pr.end());
        for (const auto &d : P.dsub) sdiag[d.first] -= d.second;
        nfill += P.fill.size() + P.jfill.size();
    }
    // S in rows: the path fill of every thread in order, then the couplings of every thread
    // in order (the sequential order), placed by a stable counting pass on the row and a
    // stable sort on the column within each row; entries with the same (row, column) are
    // summed in that order
    std::vector<int64_t> &rrp = C.rrp, &rcol = C.rcol;
    std::vector<double> &rval = C.rval;
    std::vector<int> rcnt((size_t)m + 1, 0);
    for (const Part &P : parts) {
        for (const Fill &f : P.fill) rcnt[(size_t)f.r + 1]++;
        for (const Fill &f : P.jfill) rcnt[(size_t)f.r + 1]++;
    }
    for (int j = 0; j < m; j++) rcnt[j + 1] += rcnt[j];
    std::vector<std::pair<int, double>> byrow(nfill);
    {
        std::vector<int> at(rcnt.begin(), rcnt.end() - 1);
        for (const Part &P : parts)
            for (const Fill &f : P.fill) byrow[(size_t)at[f.r]++] = {f.c, f.v};
        for (const Part &P : parts)
            for (const Fill &f : P.jfill) byrow[(size_t)at[f.r]++] = {f.c, f.v};
    }
    // each row sorted (stable insertion sort: the rows hold a few entries) and its equal
    // columns summed in order, counted then written, on host threads over rows
    rrp.assign((size_t)m + 1, 0);
    const int nrt = host_threads(m, 1 << 14);
    auto sort_row = [&](int j) {
        auto b = byrow.begin() + rcnt[j], e = byrow.begin() + rcnt[j + 1];
        for (auto q = b + (b != e); q < e; q++)
            for (auto w = q; w > b && (w - 1)->first > w->first; w--) std::iter_swap(w - 1, w);
        int u = 0;
        for (auto q = b; q != e; q++)
            if (q == b || q->first != (q - 1)->first) u++;
        rrp[(size_t)j + 1] = u;
    };
    parallel_for(nrt, [&](int t) {
        const int j0 = (int)((long long)m * t / nrt), j1 = (int)((long long)m * (t + 1) / nrt);
        for (int j = j0; j < j1; j++) sort_row(j);
    });
    for (int j = 0; j < m; j++) rrp[j + 1] += rrp[j];
    rcol.resize((size_t)rrp[m]);
    rval.resize((size_t)rrp[m]);
    parallel_for(nrt, [&](int t) {
        const int j0 = (int)((long long)m * t / nrt), j1 = (int)((long long)m * (t + 1) / nrt);
        for (int j = j0; j < j1; j++) {
            auto b = byrow.begin() + rcnt[j], e = byrow.begin() + rcnt[j + 1];
            int64_t o = rrp[j];
            for (auto q = b; q != e;) {
                auto w = q;
                double v = 0.0;
                while (w != e && w->first == q->first) v += (w++)->second;
                rcol[(size_t)o] = q->first;
                rval[(size_t)o++] = v;
                q = w;
            }
        }
    });
    clk.mark("paths: merge + reduced rows");
    return true;
}

// Stage 2: low-degree elimination on S.  A node k of S with at most kElimDegree neighbours
// is eliminated exactly: each neighbour u loses a_uk^2 / d_k from its diagonal and every
// pair of neighbours (u, v) gains the coupling -a_uk a_vk / d_k (for two neighbours this is
// series reduction of the resistor network, for more the star-mesh transform).  This can
// leave other nodes with few neighbours.  Nodes are eliminated in rounds; each round takes
// an independent set of such nodes (greedy, in elim_priority order, a fixed hash of the
// node), so a chain of L nodes collapses in O(log L) rounds and the nodes of a round are
// processed in parallel in the per-solve sweeps.  The cap bounds the fill (a node of degree D adds at most
// D (D - 1) / 2 couplings).  What remains is the core system.

struct SeriesHost {
    int m = 0, mc = 0;
    std::vector<int> round_off;               // [rounds + 1] into the elimination records
    std::vector<int> e_node;                   // S index of each eliminated node
    std::vector<int> e_off;                    // [records + 1] into e_nbr / e_a
    std::vector<int> e_nbr;                    // neighbours at elimination (S indices)
    std::vector<double> e_a, e_d;              // couplings to them, and the pivot
    std::vector<int> core_nodes;               // [mc] S index of each core row
    std::vector<int64_t> crp, ccol;            // core SSS
    std::vector<double> cval, cdiag;
    // receivers of each round's records, ascending, each with its contributions in record
    // order: rv_node[rv_round_off[q] ..], contributions rv_off[j] .. rv_off[j + 1] of
    // (record rv_rec, coupling rv_a)
    std::vector<int> rv_round_off, rv_node, rv_off, rv_rec;
    std::vector<double> rv_a;
};

// Stage 2 on the host, the device set-up's rounds (dea_setup.hip) in sequential form.  S is
// kept as a symmetric CSR with ascending columns.  Each round: the candidates (alive, at most
// max_degree neighbours, a usable pivot) are taken greedily in elim_priority order into an
// independent set; its nodes, ascending, are the round's records (their rows before the
// round); every node next to a record is a receiver, whose diagonal loses a_uk^2 / d_k and
// whose row gains the fill -a_uk a_vk / d_k, record after record; equal columns are summed
// after the old value, in record order.
void build_series_host(const CondHost &C, SeriesHost &Z, int max_degree = elim_degree()) {
    const int m = C.m;
    Z.m = m;
    StageClock clk;
    // symmetric CSR: each row's lower entries, then the rows below it that couple to it
    std::vector<int> arp((size_t)m + 1, 0), acol;
    std::vector<double> aval;
    for (int r = 0; r < m; r++)
        for (int64_t k = C.rrp[r]; k < C.rrp[r + 1]; k++) {
            arp[(size_t)r + 1]++;
            arp[(size_t)C.rcol[k] + 1]++;
        }
    for (int r = 0; r < m; r++) arp[(size_t)r + 1] += arp[(size_t)r];
    acol.resize((size_t)arp[(size_t)m]);
    aval.resize((size_t)arp[(size_t)m]);
    {
        std::vector<int> at(arp.begin(), arp.end() - 1);
        for (int r = 0; r < m; r++)
            for (int64_t k = C.rrp[r]; k < C.rrp[r + 1]; k++) {
                const int c = (int)C.rcol[k];
                acol[(size_t)at[r]] = c;
                aval[(size_t)at[r]++] = C.rval[k];
                acol[(size_t)at[c]] = r;
                aval[(size_t)at[c]++] = C.rval[k];
            }
    }
    std::vector<double> d(C.sdiag.begin(), C.sdiag.end());
    std::vector<char> alive((size_t)m, 1), in((size_t)m, 0);
    std::vector<int> rec_of((size_t)m, -1);
    long long left = m;  // nodes not eliminated yet
    Z.round_off.assign(1, 0);
    Z.e_off.assign(1, 0);
    Z.rv_round_off.assign(1, 0);
    Z.rv_off.assign(1, 0);
    std::vector<int> cand, nrp, ncol;
    std::vector<double> nval;
    std::vector<std::pair<int, double>> fill;
    clk.mark("series: adjacency");
    for (;;) {
        cand.clear();
        for (int k = 0; k < m; k++)
            if (alive[k] && arp[k + 1] - arp[k] <= max_degree &&
                d[k] > 1e-12 * std::fabs(C.sdiag[k]) && std::isfinite(d[k]))
                cand.push_back(k);
        std::sort(cand.begin(), cand.end(), [](int a, int b) {
            return elim_priority(a) < elim_priority(b);
        });
        int nsel = 0;
        for (int k : cand) {
            bool free = true;
            for (int t = arp[k]; t < arp[k + 1]; t++)
                if (in[acol[t]]) free = false;
            if (free) {
                in[k] = 1;
                nsel++;
            }
        }
        // a round that would take fewer than 1 % of the nodes left is not taken: the rounds
        // stop (as on the device)
        if (nsel == 0 || (long long)nsel * min_round_fraction() < left) {
            std::fill(in.begin(), in.end(), 0);
            break;
        }
        left -= nsel;
        // records, ascending
        for (int k = 0; k < m; k++) {
            if (!in[k]) continue;
            rec_of[k] = (int)Z.e_node.size();
            Z.e_node.push_back(k);
            Z.e_d.push_back(d[k]);
            for (int t = arp[k]; t < arp[k + 1]; t++) {
                Z.e_nbr.push_back(acol[t]);
                Z.e_a.push_back(aval[t]);
            }
            Z.e_off.push_back((int)Z.e_nbr.size());
        }
        // receivers, diagonals and the new rows
        nrp.assign((size_t)m + 1, 0);
        ncol.clear();
        nval.clear();
        for (int u = 0; u < m; u++) {
            if (alive[u] && !in[u]) {
                bool target = false;
                for (int t = arp[u]; t < arp[u + 1]; t++)
                    if (in[acol[t]]) target = true;
                if (!target) {
                    for (int t = arp[u]; t < arp[u + 1]; t++) {
                        ncol.push_back(acol[t]);
                        nval.push_back(aval[t]);
                    }
                } else {
                    Z.rv_node.push_back(u);
                    fill.clear();
                    double du = d[u];
                    for (int t = arp[u]; t < arp[u + 1]; t++) {
                        const int j = acol[t];
                        if (!in[j]) continue;
                        const int rec = rec_of[j];
                        const double a = aval[t], dk = Z.e_d[rec];
                        Z.rv_rec.push_back(rec);
                        Z.rv_a.push_back(a);
                        du = du - a * a / dk;
                        for (int q = Z.e_off[rec]; q < Z.e_off[rec + 1]; q++) {
                            const int v = Z.e_nbr[q];
                            if (v != u) fill.push_back({v, -a * Z.e_a[q] / dk});
                        }
                    }
                    d[u] = du;
                    Z.rv_off.push_back((int)Z.rv_rec.size());
                    std::stable_sort(fill.begin(), fill.end(),
                                     [](const std::pair<int, double> &x,
                                        const std::pair<int, double> &y) {
                                         return x.first < y.first;
                                     });
                    size_t p = 0;
                    int t = arp[u];
                    const int te = arp[u + 1];
                    for (;;) {
                        while (t < te && in[acol[t]]) t++;
                        if (t >= te && p >= fill.size()) break;
                        int c;
                        double v;
                        if (p >= fill.size() || (t < te && acol[t] < fill[p].first)) {
                            c = acol[t];
                            v = aval[t++];
                        } else if (t >= te || fill[p].first < acol[t]) {
                            c = fill[p].first;
                            v = fill[p++].second;
                            while (p < fill.size() && fill[p].first == c) v += fill[p++].second;
                        } else {
                            c = acol[t];
                            v = aval[t++];
                            while (p < fill.size() && fill[p].first == c) v += fill[p++].second;
                        }
                        ncol.push_back(c);
                        nval.push_back(v);
                    }
                }
            }
            nrp[(size_t)u + 1] = (int)ncol.size();
        }
        for (int k = 0; k < m; k++)
            if (in[k]) {
                alive[k] = 0;
                in[k] = 0;
            }
        arp.swap(nrp);
        acol.swap(ncol);
        aval.swap(nval);
        Z.round_off.push_back((int)Z.e_node.size());
        Z.rv_round_off.push_back((int)Z.rv_node.size());
    }
    clk.mark("series: rounds");
    // core: the nodes left, in S order, with their lower entries
    std::vector<int> core_of((size_t)m, -1);
    for (int i = 0; i < m; i++)
        if (alive[i]) {
            core_of[i] = (int)Z.core_nodes.size();
            Z.core_nodes.push_back(i);
        }
    Z.mc = (int)Z.core_nodes.size();
    Z.crp.assign((size_t)Z.mc + 1, 0);
    Z.cdiag.resize((size_t)Z.mc);
    for (int q = 0; q < Z.mc; q++) {
        const int i = Z.core_nodes[q];
        // a node left without neighbours and with no usable pivot is what the rounds leave
        // of a floating group of tubes (its rows sum to zero, so its diagonal comes out as
        // rounding noise): it is decoupled, diagonal 1, and the group takes the voltage its
        // zero right-hand side gives, one of the singular block's solutions
        const bool isolated = arp[i] == arp[i + 1];
        Z.cdiag[q] = (isolated && std::isfinite(d[i]) && !(d[i] > 1e-12 * std::fabs(C.sdiag[i])))
                         ? 1.0
                         : d[i];
        for (int t = arp[i]; t < arp[i + 1]; t++) {
            const int c = core_of[acol[t]];
            if (c < q) {
                Z.ccol.push_back(c);
                Z.cval.push_back(aval[t]);
            }
        }
        Z.crp[(size_t)q + 1] = (int64_t)Z.ccol.size();
    }
    clk.mark("series: core rows");
}

// Stage-2 sweeps on the host (the device kernels do the same per round).
void host_series_forward(const SeriesHost &Z, double *f) {
    for (size_t k = 0; k < Z.e_node.size(); k++) {
        const double fk = f[Z.e_node[k]] / Z.e_d[k];
        for (int i = Z.e_off[k]; i < Z.e_off[k + 1]; i++) f[Z.e_nbr[i]] -= Z.e_a[i] * fk;
    }
}

void host_series_back(const SeriesHost &Z, const double *f, double *x) {
    for (size_t k = Z.e_node.size(); k-- > 0;) {
        double v = f[Z.e_node[k]];
        for (int i = Z.e_off[k]; i < Z.e_off[k + 1]; i++) v -= Z.e_a[i] * x[Z.e_nbr[i]];
        x[Z.e_node[k]] = v / Z.e_d[k];
    }
}

// The sweeps of k_path_solve and k_anchor_rhs on the host (for layout tests).
void host_path_solve(const CondHost &C, const double *rhs, const double *xanchor, double *z) {
    for (size_t k = 0; k < C.ps.size(); k++) {
        const int s = C.ps[k], e = C.pe[k];
        if (C.pl[k] == kFloatingPath) {
            double g = rhs[s];
            z[s] = g;
            for (int i = s + 1; i < e; i++) {
                g = rhs[i] - C.lfac[i] * g;
                z[i] = g;
            }
            double zn = 0.0, mean = 0.0;
            z[e] = 0.0;
            for (int i = e - 1; i >= s; i--) {
                zn = (z[i] - C.band[i + 1] * zn) / C.mfac[i];
                z[i] = zn;
                mean += zn;
            }
            mean /= (double)(e - s + 1);
            for (int i = s; i <= e; i++) z[i] -= mean;
            continue;
        }
        double g = rhs[s];
        if (xanchor && C.pl[k] >= 0) g -= C.band[s] * xanchor[C.pl[k]];
        if (s == e && xanchor && C.pr[k] >= 0) g -= C.band[e + 1] * xanchor[C.pr[k]];
        z[s] = g;
        for (int i = s + 1; i <= e; i++) {
            double f = rhs[i];
            if (i == e && xanchor && C.pr[k] >= 0) f -= C.band[e + 1] * xanchor[C.pr[k]];
            g = f - C.lfac[i] * g;
            z[i] = g;
        }
        double zn = z[e] / C.mfac[e];
        z[e] = zn;
        for (int i = e - 1; i >= s; i--) {
            zn = (z[i] - C.band[i + 1] * zn) / C.mfac[i];
            z[i] = zn;
        }
    }
}

void host_anchor_rhs(const CondHost &C, const double *rhs, const double *y, double *out) {
    for (int j = 0; j < C.m; j++) {
        const int a = C.anchors[j];
        double v = rhs[a];
        if (a > 0 && C.anchor_of[a - 1] < 0 && C.band[a] != 0.0) v -= C.band[a] * y[a - 1];
        if (a + 1 < C.n && C.anchor_of[a + 1] < 0 && C.band[a + 1] != 0.0)
            v -= C.band[a + 1] * y[a + 1];
        out[j] = v;
    }
}

cicy_cond *condensed_setup(cicy_ctx *ctx, const BlockedHost &h, const BandHost &bh,
                           const char **why, bool *failed) {
    StageClock clk;
    CondHost C;
    if (!build_condensed_host(ctx, h, bh, C, why, failed)) return nullptr;
    clk.mark("condense: paths");
    const int n = C.n, m = C.m;
    const std::vector<int> &ps = C.ps, &pe = C.pe, &pl = C.pl, &pr = C.pr, &anchor_of = C.anchor_of,
                           &anchors = C.anchors;
    const std::vector<double> &lfac = C.lfac, &mfac = C.mfac, &band = C.band, &rval = C.rval;
    cicy_cond *c = new cicy_cond();
    c->ctx = ctx;
    c->n = n;
    c->m = m;
    c->npath = (int)ps.size();
    c->nnz_red = (long long)rval.size();
    c->nb = std::max(1, std::min(h.nblocks, 1024));
    bool ok = calloc_dev(ctx, &c->ps, ps.size()) && calloc_dev(ctx, &c->pe, pe.size()) &&
              calloc_dev(ctx, &c->pl, pl.size()) && calloc_dev(ctx, &c->pr, pr.size()) &&
              calloc_dev(ctx, &c->anchor_of, (size_t)n) && calloc_dev(ctx, &c->anchors, (size_t)m) &&
              calloc_dev(ctx, &c->lfac, (size_t)n) && calloc_dev(ctx, &c->mfac, (size_t)n) &&
              calloc_dev(ctx, &c->band, (size_t)n + 1) && calloc_dev(ctx, &c->y, (size_t)n) &&
              calloc_dev(ctx, &c->delta, (size_t)n) && calloc_dev(ctx, &c->rhs, (size_t)n) &&
              calloc_dev(ctx, &c->norm, 2) && calloc_dev(ctx, &c->counter, 1) &&
              calloc_dev(ctx, &c->part, 2 * (size_t)c->nb) &&
              calloc_dev(ctx, &c->gpart, 2 * (size_t)c->nb) &&
              calloc_dev(ctx, &c->gcount, (size_t)c->nb) &&
              hip_ok(ctx, hipEventCreate(&c->ev0), "hipEventCreate") &&
              hip_ok(ctx, hipEventCreate(&c->ev1), "hipEventCreate");
    ok = ok && up(ctx, c->ps, ps) && up(ctx, c->pe, pe) && up(ctx, c->pl, pl) &&
         up(ctx, c->pr, pr) && up(ctx, c->anchor_of, anchor_of) && up(ctx, c->anchors, anchors) &&
         up(ctx, c->lfac, lfac) && up(ctx, c->mfac, mfac) && up(ctx, c->band, band) &&
         hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    // stage 2: series rounds over S, then the core system
    clk.mark("condense: path upload");
    SeriesHost Z;
    if (ok && m > 0) build_series_host(C, Z);
    clk.mark("condense: series rounds");
    c->mc = Z.mc;
    c->nrec = (int)Z.e_node.size();
    c->round_off = Z.round_off;
    // receivers of each round, each with its contributions in record order (built with the
    // rounds)
    const std::vector<int> &rv_node = Z.rv_node, &rv_off = Z.rv_off, &rv_rec = Z.rv_rec;
    const std::vector<double> &rv_a = Z.rv_a;
    c->rv_round_off = Z.rv_round_off.empty() ? std::vector<int>(1, 0) : Z.rv_round_off;
    clk.mark("condense: receiver lists");
    c->nnz_core = (long long)Z.cval.size();
    if (ok && m > 0) {
        ok = calloc_dev(ctx, &c->e_node, Z.e_node.size()) &&
             calloc_dev(ctx, &c->e_off, Z.e_off.size()) && calloc_dev(ctx, &c->e_nbr, Z.e_nbr.size()) &&
             calloc_dev(ctx, &c->e_a, Z.e_a.size()) && calloc_dev(ctx, &c->e_d, Z.e_d.size()) &&
             calloc_dev(ctx, &c->core_nodes, Z.core_nodes.size()) &&
             calloc_dev(ctx, &c->fS, (size_t)m) && calloc_dev(ctx, &c->xS, (size_t)m) &&
             up(ctx, c->e_node, Z.e_node) && up(ctx, c->e_off, Z.e_off) &&
             up(ctx, c->e_nbr, Z.e_nbr) && up(ctx, c->e_a, Z.e_a) && up(ctx, c->e_d, Z.e_d) &&
             up(ctx, c->core_nodes, Z.core_nodes) &&
             calloc_dev(ctx, &c->rv_node, rv_node.size()) &&
             calloc_dev(ctx, &c->rv_off, rv_off.size()) &&
             calloc_dev(ctx, &c->rv_rec, rv_rec.size()) &&
             calloc_dev(ctx, &c->rv_a, rv_a.size()) && up(ctx, c->rv_node, rv_node) &&
             up(ctx, c->rv_off, rv_off) && up(ctx, c->rv_rec, rv_rec) && up(ctx, c->rv_a, rv_a) &&
             hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
        if (ok && Z.mc > 0) {
            // a core row with a vanishing diagonal that still has neighbours is refused by
            // the layout check: then there is simply no condensed system
            bool positive = true;
            for (double v : Z.cdiag)
                if (!(v > 0.0) || !std::isfinite(v)) positive = false;
            if (!positive) {
                condensed_free(c);
                *why = "the core system has a non-positive diagonal (floating component)";
                return nullptr;
            }
            // the core's own vectors (right-hand side and solution of the block solver; the
            // two-kernel and resident solvers use those of the core cicy_sys)
            ok = calloc_dev(ctx, &c->core_b, (size_t)Z.mc) &&
                 calloc_dev(ctx, &c->core_x, (size_t)Z.mc);
            c->h_orig.resize((size_t)Z.mc);
            for (int q = 0; q < Z.mc; q++) c->h_orig[q] = anchors[Z.core_nodes[q]];
            c->h_crp = std::move(Z.crp);
            c->h_ccol = std::move(Z.ccol);
            c->h_cval = std::move(Z.cval);
            c->h_cdiag = std::move(Z.cdiag);
        }
    }
    if (!ok) {
        condensed_free(c);
        *failed = true;
        return nullptr;
    }
    return c;
}

// Solves A d = rhs (rhs on the device, n rows) into c->delta with the reduced system;
// returns the reduced solve's statistics.
static bool condensed_correction(cicy_sys *s, int inner, const double *rhs, double tol,
                                 long long max_iter, double ref2, cicy_solve_stats *inner_st) {
    cicy_cond *c = s->cond;
    cicy_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    memset(inner_st, 0, sizeof(*inner_st));
    inner_st->status = CICY_OK;
    // y = T^-1 rhs on the paths
    hipLaunchKernelGGL(k_path_solve, dim3(blocks(c->npath, kCondThreads)), dim3(kCondThreads), 0,
                       st, c->npath, c->ps, c->pe, c->pl, c->pr, c->lfac, c->mfac, c->band, rhs,
                       (const double *)nullptr, c->y);
    if (c->m > 0) {
        // reduced right-hand side on S, pushed through the series rounds
        hipLaunchKernelGGL(k_anchor_rhs, dim3(blocks(c->m, kCondThreads)), dim3(kCondThreads), 0,
                           st, c->m, c->n, c->anchors, c->anchor_of, c->band, rhs, c->y, c->fS);
        const int nr = (int)c->round_off.size() - 1;
        for (int q = 0; q < nr; q++) {
            const int b = c->rv_round_off[q], e = c->rv_round_off[q + 1];
            if (e > b)
                hipLaunchKernelGGL(k_series_fwd, dim3(blocks(e - b, kCondThreads)),
                                   dim3(kCondThreads), 0, st, b, e, c->rv_node, c->rv_off,
                                   c->rv_rec, c->rv_a, c->e_node, c->e_d, c->fS);
        }
        if (c->mc > 0) {
            bool ok;
            const double *core_x;
            if (inner == CICY_METHOD_CONDENSED_BLOCK) {
                hipLaunchKernelGGL(k_gather, dim3(blocks(c->mc, kCondThreads)),
                                   dim3(kCondThreads), 0, st, c->mc, c->core_nodes, c->fS,
                                   c->core_b);
                ok = (c->blk || condensed_set_blocks(s, 0, nullptr)) &&
                     solve_core(ctx, c->blk, c->core_b, tol, max_iter, ref2, c->core_x, inner_st);
                core_x = c->core_x;
            } else {
                if (!condensed_core_sys(s, inner)) return false;
                cicy_sys *red = c->red;
                hipLaunchKernelGGL(k_gather, dim3(blocks(c->mc, kCondThreads)),
                                   dim3(kCondThreads), 0, st, c->mc, c->core_nodes, c->fS,
                                   red->r0);
                if (!hip_ok(ctx, hipMemsetAsync(red->x0, 0, (size_t)c->mc * sizeof(double), st),
                            "memset"))
                    return false;
                red->has_b = false;
                red->ref2_override = ref2;
                ok = inner == CICY_METHOD_RESIDENT
                         ? solve_resident(red, tol, max_iter, CICY_REL_B, inner_st)
                         : solve_pcg64(red, tol, max_iter, CICY_REL_B, inner_st);
                core_x = red->x;
            }
            if (!ok) return false;
            hipLaunchKernelGGL(k_scatter, dim3(blocks(c->mc, kCondThreads)), dim3(kCondThreads), 0,
                               st, c->mc, c->core_nodes, core_x, c->xS);
        }
        for (int q = nr - 1; q >= 0; q--) {
            const int b = c->round_off[q], e = c->round_off[q + 1];
            hipLaunchKernelGGL(k_series_back, dim3(blocks(e - b, kCondThreads)), dim3(kCondThreads),
                               0, st, b, e, c->e_node, c->e_off, c->e_nbr, c->e_a, c->e_d, c->fS,
                               c->xS);
        }
        hipLaunchKernelGGL(k_scatter, dim3(blocks(c->m, kCondThreads)), dim3(kCondThreads), 0, st,
                           c->m, c->anchors, c->xS, c->delta);
    }
    // d_I = T^-1 (rhs_I - A_IJ d_J)
    hipLaunchKernelGGL(k_path_solve, dim3(blocks(c->npath, kCondThreads)), dim3(kCondThreads), 0,
                       st, c->npath, c->ps, c->pe, c->pl, c->pr, c->lfac, c->mfac, c->band, rhs,
                       (const double *)(c->m > 0 ? c->xS : nullptr), c->delta);
    return true;
}

// host copy of the core's strictly-lower rows, diagonal and full-system rows (kept on the
// device by the device set-up)
static bool core_to_host(cicy_ctx *ctx, cicy_cond *c) {
    if (!c->on_device || !c->h_crp.empty()) return true;
    const int mc = c->mc;
    std::vector<int> frp((size_t)mc + 1), orig((size_t)mc);
    c->h_cdiag.resize((size_t)mc);
    if (!d2h_sync(ctx, frp.data(), c->d_frp, frp.size() * sizeof(int))) return false;
    std::vector<int> fcol((size_t)frp[(size_t)mc]);
    std::vector<double> fval(fcol.size());
    if (!d2h_sync(ctx, fcol.data(), c->d_fcol, fcol.size() * sizeof(int)) ||
        !d2h_sync(ctx, fval.data(), c->d_fval, fval.size() * sizeof(double)) ||
        !d2h_sync(ctx, c->h_cdiag.data(), c->d_cdiag, (size_t)mc * sizeof(double)) ||
        !d2h_sync(ctx, orig.data(), c->d_orig, (size_t)mc * sizeof(int)))
        return false;
    c->h_crp.assign((size_t)mc + 1, 0);
    for (int q = 0; q < mc; q++) {
        for (int k = frp[(size_t)q]; k < frp[(size_t)q + 1]; k++)
            if (fcol[(size_t)k] < q) {
                c->h_ccol.push_back(fcol[(size_t)k]);
                c->h_cval.push_back(fval[(size_t)k]);
            }
        c->h_crp[(size_t)q + 1] = (int64_t)c->h_ccol.size();
    }
    c->h_orig = std::move(orig);
    return true;
}

bool condensed_core_sys(cicy_sys *s, int inner) {
    cicy_cond *c = s->cond;
    if (!c || c->mc <= 0) return true;
    if (!core_to_host(s->ctx, c)) return false;
    if (!c->red) {
        if (!sys_create(s->ctx, c->mc, (int64_t)c->h_cval.size(), c->h_crp.data(),
                        c->h_ccol.data(), c->h_cval.data(), c->h_cdiag.data(), &c->red))
            return false;
        c->red->resident_rows_req = s->resident_rows_req;
    }
    return sys_prepare(c->red, inner);
}

int condensed_set_blocks(cicy_sys *s, int64_t nblocks, const int64_t *ranges) {
    cicy_ctx *ctx = s->ctx;
    cicy_cond *c = s->cond;
    if (!c || c->mc <= 0) {
        set_error(ctx, "no condensed core (%s)", s->cond_why ? s->cond_why : "not built");
        return 0;
    }
    if (nblocks < 0 || (nblocks > 0 && !ranges)) {
        set_error(ctx, "cicy_sys_set_blocks: invalid block list");
        return 0;
    }
    for (int64_t b = 0; b < nblocks; b++)
        if (ranges[2 * b] < 0 || ranges[2 * b + 1] > s->n || ranges[2 * b + 1] < ranges[2 * b]) {
            set_error(ctx, "cicy_sys_set_blocks: row group %lld outside [0, %d)", (long long)b, s->n);
            return 0;
        }
    if (c->blk) {
        core_free(c->blk);
        c->blk = nullptr;
    }
    c->blk = c->on_device
                 ? core_setup_device(ctx, c->mc, c->d_frp, c->d_fcol, c->d_fval, c->d_cdiag,
                                     c->d_orig, nblocks, ranges)
                 : core_setup(ctx, c->mc, c->h_crp, c->h_ccol, c->h_cval, c->h_cdiag, c->h_orig,
                              nblocks, ranges);
    return c->blk ? 1 : 0;
}

bool solve_condensed(cicy_sys *s, int inner, double tol, long long max_iter, int rel_to,
                     cicy_solve_stats *stats) {
    cicy_ctx *ctx = s->ctx;
    cicy_cond *c = s->cond;
    if (!c) {
        set_error(ctx, "no condensed system (%s)", s->cond_why ? s->cond_why : "not built");
        if (stats) stats->status = CICY_BAD_INPUT;
        return false;
    }
    if (inner == CICY_METHOD_RESIDENT && c->mc > 0 &&
        (!condensed_core_sys(s, CICY_METHOD_RESIDENT) || !c->red->res)) {
        set_error(ctx, "the reduced system has no resident layout (%s)",
                  c->red->res_why ? c->red->res_why : "not built");
        if (stats) stats->status = CICY_BAD_INPUT;
        return false;
    }
    if (rel_to == CICY_REL_B && !s->has_b) {
        set_error(ctx, "CICY_REL_B needs b (cicy_sys_set_rhs)");
        return false;
    }
    const double ref2 = rel_to == CICY_REL_B ? s->bnorm2 : s->r0norm2;
    hipStream_t st = ctx->stream;
    const size_t vb = (size_t)s->n * sizeof(double);
    bool ok = hip_ok(ctx, hipEventRecord(c->ev0, st), "event") &&
              hip_ok(ctx, hipMemcpyAsync(s->x, s->x0, vb, hipMemcpyDeviceToDevice, st), "copy x0") &&
              hip_ok(ctx, hipMemcpyAsync(c->rhs, s->r0, vb, hipMemcpyDeviceToDevice, st), "copy r0");
    long long total = 0;
    int passes = 0, status = CICY_MAXITER;
    double rel = -1.0, kernel_ms = 0.0;
    // core iterations allowed per correction pass (default: the remaining cap)
    long long pass_cap = 0;
    if (const char *e = getenv("CICY_CORE_PASS_ITERS")) pass_cap = atoll(e);
    while (ok) {
        cicy_solve_stats ist;
        const long long left = max_iter - total;
        ok = condensed_correction(s, inner, c->rhs, tol,
                                  pass_cap > 0 ? std::min(pass_cap, left) : left, ref2, &ist);
        if (!ok) break;
        total += ist.iters;
        kernel_ms += ist.kernel_ms;
        // x += d; r = rhs - A d (the residual of the whole system)
        sys_spmv64(s, c->delta, s->q);
        hipLaunchKernelGGL(k_correct, dim3(c->nb), dim3(kThreads), 0, st, s->n, s->x, c->delta,
                           c->rhs, s->q, s->r, c->norm, c->counter,
                           RedBuf{c->part, c->gpart, c->gcount});
        double rr = 0.0;
        ok = d2h_sync(ctx, &rr, c->norm, sizeof(double));
        if (!ok) break;
        rel = ref2 > 0 ? std::sqrt(rr / ref2) : std::sqrt(rr);
        if (rel <= tol) {
            status = CICY_OK;
            break;
        }
        if (ist.status == CICY_BREAKDOWN) {
            status = CICY_BREAKDOWN;
            break;
        }
        // the block method's pipelined recursion may stall on a badly preconditioned core
        // and is then corrected from the true residual; the others are expected to converge
        const int max_passes = inner == CICY_METHOD_CONDENSED_BLOCK ? 10 : 3;
        if (passes >= max_passes || total >= max_iter) break;
        passes++;
        ok = hip_ok(ctx, hipMemcpyAsync(c->rhs, s->r, vb, hipMemcpyDeviceToDevice, st), "copy r");
    }
    // Correction passes used up short of the tolerance and of the iteration cap (a stalled
    // core recursion): continue from x with the reference's algorithm, FP64 Jacobi-PCG on
    // the whole system (s->r = b - A x after the last pass), stopping against the same
    // reference norm, until it converges or the remaining iterations of the cap
    // (300 * nodes, Direct_Electrifying.cpp:1559) are spent.
    long long fallback_iters = 0;
    if (ok && status == CICY_MAXITER && total < max_iter) {
        cicy_solve_stats fst;
        s->ref2_override = ref2;
        ok = solve_pcg64(s, tol, max_iter - total, rel_to, &fst, /*keep_x=*/true);
        s->ref2_override = -1.0;
        if (ok) {
            fallback_iters = fst.iters;
            total += fst.iters;
            status = fst.status;
            rel = fst.rel_res;
        }
    }
    double tr = -1.0;
    if (ok && s->has_b) ok = sys_true_residual(s, s->x, &tr);
    float ms = 0.f;
    if (hipEventRecord(c->ev1, st) == hipSuccess && hipEventSynchronize(c->ev1) == hipSuccess)
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iters = total;
        stats->status = ok ? status : CICY_DEVICE_ERROR;
        stats->refinements = passes;
        stats->rel_res = rel;
        stats->true_rel_res = tr;
        stats->solve_ms = ms;
        stats->fp64_iters = fallback_iters;
        stats->kernel_ms = kernel_ms;
    }
    return ok;
}

}  // namespace cicy

// ------------------------------------------------------------------------------------
// C-ABI: host forms of the condensation, for tests without a GPU
// ------------------------------------------------------------------------------------

extern "C" {

int cicy_host_condense(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                       const double *values, const double *diagonal, int64_t nnz,
                       const double *rhs, int64_t *m_out, int64_t *nnz_out, int64_t *anchors,
                       int64_t *red_row_ptr, int64_t *red_col_ind, double *red_values,
                       double *red_diagonal, double *red_rhs) {
    using namespace cicy;
    cicy_ctx dummy;
    BlockedHost h;
    if (!build_blocked(&dummy, n, nnz, row_ptr, col_ind, values, diagonal, 1024, h)) return 0;
    BandHost bh;
    build_band(h, bh);
    CondHost C;
    const char *why = nullptr;
    bool failed = false;
    if (!build_condensed_host(&dummy, h, bh, C, &why, &failed)) return 0;
    *m_out = C.m;
    *nnz_out = (int64_t)C.rval.size();
    for (int j = 0; j < C.m; j++) {
        anchors[j] = C.anchors[j];
        red_diagonal[j] = C.sdiag[j];
    }
    for (int j = 0; j <= C.m; j++) red_row_ptr[j] = C.rrp[j];
    for (size_t k = 0; k < C.rval.size(); k++) {
        red_col_ind[k] = C.rcol[k];
        red_values[k] = C.rval[k];
    }
    if (rhs && red_rhs) {
        std::vector<double> y((size_t)n, 0.0);
        host_path_solve(C, rhs, nullptr, y.data());
        host_anchor_rhs(C, rhs, y.data(), red_rhs);
    }
    return 1;
}

int cicy_host_expand(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                     const double *values, const double *diagonal, int64_t nnz,
                     const double *rhs, const double *x_anchor, double *x) {
    using namespace cicy;
    cicy_ctx dummy;
    BlockedHost h;
    if (!build_blocked(&dummy, n, nnz, row_ptr, col_ind, values, diagonal, 1024, h)) return 0;
    BandHost bh;
    build_band(h, bh);
    CondHost C;
    const char *why = nullptr;
    bool failed = false;
    if (!build_condensed_host(&dummy, h, bh, C, &why, &failed)) return 0;
    for (int j = 0; j < C.m; j++) x[C.anchors[j]] = x_anchor[j];
    host_path_solve(C, rhs, x_anchor, x);
    return 1;
}

int cicy_sys_core(cicy_sys *s, cicy_sys **core) {
    if (!s || !core) return 0;
    (void)hipSetDevice(s->ctx->device);
    if (!cicy::sys_prepare(s, CICY_METHOD_CONDENSED)) return 0;
    *core = (s->cond && s->cond->red) ? s->cond->red : nullptr;
    return *core ? 1 : 0;
}

static bool host_both(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                      const double *values, const double *diagonal, int64_t nnz,
                      cicy::CondHost &C, cicy::SeriesHost &Z) {
    using namespace cicy;
    cicy_ctx dummy;
    BlockedHost h;
    if (!build_blocked(&dummy, n, nnz, row_ptr, col_ind, values, diagonal, 1024, h)) return false;
    BandHost bh;
    build_band(h, bh);
    const char *why = nullptr;
    bool failed = false;
    if (!build_condensed_host(&dummy, h, bh, C, &why, &failed)) return false;
    build_series_host(C, Z);
    return true;
}

int cicy_host_core(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                   const double *values, const double *diagonal, int64_t nnz, const double *rhs,
                   int64_t *m_out, int64_t *nnz_out, int64_t *rounds_out, int64_t *core_rows,
                   int64_t *core_row_ptr, int64_t *core_col_ind, double *core_values,
                   double *core_diagonal, double *core_rhs) {
    using namespace cicy;
    CondHost C;
    SeriesHost Z;
    if (!host_both(n, row_ptr, col_ind, values, diagonal, nnz, C, Z)) return 0;
    *m_out = Z.mc;
    *nnz_out = (int64_t)Z.cval.size();
    *rounds_out = (int64_t)Z.round_off.size() - 1;
    for (int q = 0; q < Z.mc; q++) {
        core_rows[q] = C.anchors[Z.core_nodes[q]];
        core_diagonal[q] = Z.cdiag[q];
    }
    for (int q = 0; q <= Z.mc; q++) core_row_ptr[q] = Z.crp[q];
    for (size_t k = 0; k < Z.cval.size(); k++) {
        core_col_ind[k] = Z.ccol[k];
        core_values[k] = Z.cval[k];
    }
    if (rhs && core_rhs) {
        std::vector<double> y((size_t)n, 0.0), f((size_t)C.m, 0.0);
        host_path_solve(C, rhs, nullptr, y.data());
        host_anchor_rhs(C, rhs, y.data(), f.data());
        host_series_forward(Z, f.data());
        for (int q = 0; q < Z.mc; q++) core_rhs[q] = f[Z.core_nodes[q]];
    }
    return 1;
}

int cicy_host_core_expand(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                          const double *values, const double *diagonal, int64_t nnz,
                          const double *rhs, const double *x_core, double *x) {
    using namespace cicy;
    CondHost C;
    SeriesHost Z;
    if (!host_both(n, row_ptr, col_ind, values, diagonal, nnz, C, Z)) return 0;
    std::vector<double> y((size_t)n, 0.0), f((size_t)C.m, 0.0), xs((size_t)C.m, 0.0);
    host_path_solve(C, rhs, nullptr, y.data());
    host_anchor_rhs(C, rhs, y.data(), f.data());
    host_series_forward(Z, f.data());
    for (int q = 0; q < Z.mc; q++) xs[Z.core_nodes[q]] = x_core[q];
    host_series_back(Z, f.data(), xs.data());
    for (int j = 0; j < C.m; j++) x[C.anchors[j]] = xs[j];
    host_path_solve(C, rhs, xs.data(), x);
    return 1;
}

}  // extern "C"
