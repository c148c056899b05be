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
#include <cmath>
#include <cstring>
#include <numeric>

using namespace cicy;

struct cicy_cond {
    int n = 0, m = 0, npath = 0;
    long long nnz_red = 0;
    int *ps = nullptr, *pe = nullptr, *pl = nullptr, *pr = nullptr;  // [npath]
    int *anchor_of = nullptr;   // [n] reduced index or -1
    int *anchors = nullptr;     // [m] full index
    double *lfac = nullptr, *mfac = nullptr;  // [n] Thomas multipliers and pivots (path rows)
    double *band = nullptr;     // [n + 1] a(i, i-1) (0 where absent)
    double *y = nullptr, *delta = nullptr, *rhs = nullptr;  // [n]
    double *norm = nullptr;     // [2]
    unsigned *counter = nullptr;
    double *part = nullptr, *gpart = nullptr;
    unsigned *gcount = nullptr;
    int nb = 0;                 // work-groups of the norm kernel
    cicy_sys *red = nullptr;    // the reduced system S on the device
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // timing (the inner solve uses the context's)
};

namespace cicy {
namespace {

constexpr int kCondThreads = 256;

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
    double g = rhs[s];
    if (xanchor && pl[k] >= 0) g -= band[s] * xanchor[pl[k]];
    if (s == e && xanchor && pr[k] >= 0) g -= band[e + 1] * xanchor[pr[k]];
    z[s] = g;
    for (int i = s + 1; i <= e; i++) {
        double f = rhs[i];
        if (i == e && xanchor && pr[k] >= 0) f -= band[e + 1] * xanchor[pr[k]];
        g = f - lfac[i] * g;
        z[i] = g;
    }
    double zn = z[e] / mfac[e];
    z[e] = zn;
    for (int i = e - 1; i >= s; i--) {
        zn = (z[i] - band[i + 1] * zn) / mfac[i];
        z[i] = zn;
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
    if (hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
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
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    void *ptrs[] = {c->ps, c->pe, c->pl, c->pr, c->anchor_of, c->anchors, c->lfac, c->mfac,
                    c->band, c->y, c->delta, c->rhs, c->norm, c->counter, c->part, c->gpart,
                    c->gcount};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete c;
}

int condensed_info(const cicy_cond *c, int what) {
    if (!c) return 0;
    return what == 0 ? c->m : (int)c->nnz_red;
}

// Host form of the condensation: anchors, paths with their LU factors, the reduced SSS
// matrix S.  Returns false with *why set when condensing is not worth it, or with an error
// (failed) when a pivot is not positive.
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
    std::vector<int> &anchor_of = C.anchor_of, &anchors = C.anchors;
    anchor_of.assign((size_t)n, -1);
    anchors.clear();
    {
        std::vector<char> is_anchor((size_t)n, 0);
        for (int k = 0; k < bh.m; k++) is_anchor[bh.xrow[k]] = 1;
        for (int i = 0; i < n; i++)
            if (is_anchor[i]) {
                anchor_of[i] = (int)anchors.size();
                anchors.push_back(i);
            }
    }
    const int m = (int)anchors.size();
    C.m = m;
    if ((long long)m * 2 > n) {
        *why = "more than half of the rows carry junction entries: nothing worth condensing";
        return false;
    }
    // paths, their LU factors and their Schur complement terms
    std::vector<int> &ps = C.ps, &pe = C.pe, &pl = C.pl, &pr = C.pr;
    std::vector<double> &lfac = C.lfac, &mfac = C.mfac, &sdiag = C.sdiag;
    lfac.assign((size_t)n, 0.0);
    mfac.assign((size_t)n, 1.0);
    sdiag.resize((size_t)m);
    C.band = bh.band;
    for (int j = 0; j < m; j++) sdiag[j] = h.diag[anchors[j]];
    struct Fill {
        int r, c;
        double v;
    };
    std::vector<Fill> fill;
    std::vector<double> g;
    const std::vector<double> &band = bh.band;  // [n + 1], band[i] = a(i, i-1)
    for (int i = 0; i < n;) {
        if (anchor_of[i] >= 0) {
            i++;
            continue;
        }
        const int s = i;
        int e = s;
        while (e + 1 < n && anchor_of[e + 1] < 0 && band[e + 1] != 0.0) e++;
        const int L = (s > 0 && band[s] != 0.0) ? s - 1 : -1;
        const int R = (e + 1 < n && band[e + 1] != 0.0) ? e + 1 : -1;
        ps.push_back(s);
        pe.push_back(e);
        pl.push_back(L >= 0 ? anchor_of[L] : -1);
        pr.push_back(R >= 0 ? anchor_of[R] : -1);
        mfac[s] = h.diag[s];
        for (int k = s + 1; k <= e; k++) {
            lfac[k] = band[k] / mfac[k - 1];
            mfac[k] = h.diag[k] - lfac[k] * band[k];
        }
        for (int k = s; k <= e; k++)
            if (!(mfac[k] > 0.0) || !std::isfinite(mfac[k])) {
                set_error(ctx, "condensation: non-positive pivot on the path of rows %d..%d", s, e);
                *failed = true;
                return false;
            }
        if (L >= 0) {
            // u = T^-1 (a(s, L) e_s): forward sweep from g_s = a(s, L), then back-substitution
            g.assign((size_t)(e - s + 1), 0.0);
            g[0] = band[s];
            for (int k = s + 1; k <= e; k++) g[k - s] = -lfac[k] * g[k - 1 - s];
            double u = g[e - s] / mfac[e];
            const double ue = u;
            for (int k = e - 1; k >= s; k--) u = (g[k - s] - band[k + 1] * u) / mfac[k];
            sdiag[anchor_of[L]] -= band[s] * u;  // u is u_s here
            if (R >= 0) fill.push_back({anchor_of[R], anchor_of[L], -band[e + 1] * ue});
        }
        if (R >= 0) {
            // v = T^-1 (a(R, e) e_e): only its last entry is needed
            sdiag[anchor_of[R]] -= band[e + 1] * (band[e + 1] / mfac[e]);
        }
        i = e + 1;
    }
    // couplings among anchors in the original matrix (junctions, adjacent anchors)
    for (int r = 0; r < n; r++) {
        if (anchor_of[r] < 0) continue;
        for (int k = h.rp[r]; k < h.rp[r + 1]; k++) {
            const int c = h.col[k];
            if (anchor_of[c] >= 0) fill.push_back({anchor_of[r], anchor_of[c], h.val[k]});
        }
    }
    std::sort(fill.begin(), fill.end(),
              [](const Fill &a, const Fill &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    std::vector<int64_t> &rrp = C.rrp, &rcol = C.rcol;
    std::vector<double> &rval = C.rval;
    rrp.assign((size_t)m + 1, 0);
    rcol.clear();
    rval.clear();
    for (size_t k = 0; k < fill.size();) {
        size_t q = k;
        double v = 0.0;
        while (q < fill.size() && fill[q].r == fill[k].r && fill[q].c == fill[k].c) v += fill[q++].v;
        rcol.push_back(fill[k].c);
        rval.push_back(v);
        rrp[(size_t)fill[k].r + 1]++;
        k = q;
    }
    for (int j = 0; j < m; j++) rrp[j + 1] += rrp[j];

    return true;
}

// The sweeps of k_path_solve and k_anchor_rhs on the host (for layout tests).
void host_path_solve(const CondHost &C, const double *rhs, const double *xanchor, double *z) {
    for (size_t k = 0; k < C.ps.size(); k++) {
        const int s = C.ps[k], e = C.pe[k];
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
    CondHost C;
    if (!build_condensed_host(ctx, h, bh, C, why, failed)) return nullptr;
    const int n = C.n, m = C.m;
    const std::vector<int> &ps = C.ps, &pe = C.pe, &pl = C.pl, &pr = C.pr, &anchor_of = C.anchor_of,
                           &anchors = C.anchors;
    const std::vector<double> &lfac = C.lfac, &mfac = C.mfac, &band = C.band, &rval = C.rval,
                              &sdiag = C.sdiag;
    const std::vector<int64_t> &rrp = C.rrp, &rcol = C.rcol;
    cicy_cond *c = new cicy_cond();
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
    if (ok && m > 0)
        ok = sys_create(ctx, m, (int64_t)rval.size(), rrp.data(), rcol.data(), rval.data(),
                        sdiag.data(), kExtraResident, &c->red) != 0;
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
        cicy_sys *red = c->red;
        hipLaunchKernelGGL(k_anchor_rhs, dim3(blocks(c->m, kCondThreads)), dim3(kCondThreads), 0,
                           st, c->m, c->n, c->anchors, c->anchor_of, c->band, rhs, c->y, red->r0);
        if (!hip_ok(ctx, hipMemsetAsync(red->x0, 0, (size_t)c->m * sizeof(double), st), "memset"))
            return false;
        red->has_b = false;
        red->ref2_override = ref2;
        bool ok;
        if (inner == CICY_METHOD_RESIDENT)
            ok = solve_resident(red, tol, max_iter, CICY_REL_B, inner_st);
        else
            ok = solve_pcg64(red, tol, max_iter, CICY_REL_B, inner_st);
        if (!ok) return false;
        hipLaunchKernelGGL(k_scatter, dim3(blocks(c->m, kCondThreads)), dim3(kCondThreads), 0, st,
                           c->m, c->anchors, red->x, c->delta);
    }
    // d_I = T^-1 (rhs_I - A_IJ d_J)
    hipLaunchKernelGGL(k_path_solve, dim3(blocks(c->npath, kCondThreads)), dim3(kCondThreads), 0,
                       st, c->npath, c->ps, c->pe, c->pl, c->pr, c->lfac, c->mfac, c->band, rhs,
                       (const double *)(c->m > 0 ? c->red->x : nullptr), c->delta);
    return true;
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
    if (inner == CICY_METHOD_RESIDENT && c->m > 0 && !c->red->res) {
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
    double rel = -1.0;
    while (ok) {
        cicy_solve_stats ist;
        ok = condensed_correction(s, inner, c->rhs, tol, max_iter - total, ref2, &ist);
        if (!ok) break;
        total += ist.iters;
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
        if (passes >= 3 || total >= max_iter) break;
        passes++;
        ok = hip_ok(ctx, hipMemcpyAsync(c->rhs, s->r, vb, hipMemcpyDeviceToDevice, st), "copy r");
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
        stats->fp64_iters = total;
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

}  // extern "C"
