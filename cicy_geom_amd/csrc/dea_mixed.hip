// dea_mixed.hip -- mixed-precision Jacobi-PCG for the DEA conductance system.
//
// Not a reference algorithm: the reference solves in FP64 only
// (Direct_Electrifying.cpp:1538-1659).  This path returns the same solution to the same
// stopping rule, measured on the FP64 true residual:
//
//   * the matrix is Jacobi-scaled once, A^ = D^-1/2 A D^-1/2 (unit diagonal, |a^| <= 1 for
//     the conductance M-matrix), and stored in FP32; Jacobi-PCG on A is plain CG on A^;
//   * CG runs in FP32 on A^ (the SSS kernel of dea_kernels.h with UNIT=true): r^, p, q and
//     the correction x^ are FP32 vectors, reductions accumulate in FP64;
//   * reliable updates (Sleijpen & van der Vorst, "Reliable updated residuals in hybrid
//     Bi-CG methods", Computing 56, 1996): when the FP32 residual has dropped by a factor
//     delta since the last update, or far enough that the stopping rule should hold, the
//     correction is folded into the FP64 solution x += D^-1/2 x^, the true residual
//     r = b - A x is recomputed with the FP64 SSS kernel, and r^ = D^-1/2 r replaces the
//     recursive residual without restarting the search direction;
//   * the solve stops only on the FP64 true residual, ||b - A x|| <= tol * ||ref||.
// Updates are evaluated once per captured chunk of iterations.
#include "dea_internal.h"

#include <cmath>
#include <cstring>

using namespace cicy;

namespace {

constexpr BatchDev kNoBatch = {nullptr, nullptr, nullptr, nullptr, nullptr};

// r = b - ax
__global__ __launch_bounds__(kThreads) void k_sub_from(int n, double *__restrict__ r,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ ax) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

__global__ __launch_bounds__(kThreads) void k_add(int n, double *__restrict__ out,
                                                  const double *__restrict__ a,
                                                  const double *__restrict__ b) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

template <int RB>
void launch_chunk(cicy_sys *s) {
    hipStream_t st = s->ctx->stream;
    dim3 g(s->nblocks), b(kThreads);
    for (int it = 0; it < kChunk; it++) {
        float *pold = (it & 1) ? s->ps1 : s->ps0;
        float *pnew = (it & 1) ? s->ps0 : s->ps1;
        launch_apply<float, RB, true, true, false>(mat32(s), s->nblocks, st, kNoBatch, pold,
                                                   s->rs, pnew, s->qs, s->st, redbuf(s),
                                                   nullptr);
        hipLaunchKernelGGL((k_update_mixed<RB, false>), g, b, 0, st, s->n, kNoBatch, s->xs,
                           s->rs, pnew, s->qs, s->st, s->mix, redbuf(s));
    }
    hipLaunchKernelGGL((k_fold<RB, false>), g, b, 0, st, s->n, kNoBatch, s->diag, s->x, s->xs,
                       s->mix, 0);
    launch_apply<double, RB, false, false, false>(mat64(s), s->nblocks, st, kNoBatch, s->x,
                                                  nullptr, nullptr, s->q, nullptr, RedBuf{},
                                                  s->mix);
    hipLaunchKernelGGL((k_reliable_resid<RB, false>), g, b, 0, st, s->n, kNoBatch, s->diag,
                       s->b, s->q, s->r, s->rs, s->st, s->mix, redbuf(s), kStallLimit);
}

template <int RB>
void launch_init(cicy_sys *s, double tol, double stop_ref2, long long max_iter) {
    hipLaunchKernelGGL((k_mixed_init<RB, false>), dim3(s->nblocks), dim3(kThreads), 0,
                       s->ctx->stream, s->n, kNoBatch, s->diag, s->r0, s->rs, s->xs, s->ps0,
                       s->ps1, s->st, s->mix, redbuf(s), tol, nullptr, stop_ref2, max_iter,
                       kDelta * kDelta);
}

template <int RB>
void launch_fold_always(cicy_sys *s) {
    hipLaunchKernelGGL((k_fold<RB, false>), dim3(s->nblocks), dim3(kThreads), 0, s->ctx->stream,
                       s->n, kNoBatch, s->diag, s->x, s->xs, s->mix, 1);
}

template <int RB>
void launch_spmv64(cicy_sys *s, const double *v, double *out) {
    launch_apply<double, RB, false, false, false>(mat64(s), s->nblocks, s->ctx->stream,
                                                  kNoBatch, v, nullptr, nullptr, out, nullptr,
                                                  RedBuf{}, nullptr);
}

template <int RB>
void launch_spmv32(cicy_sys *s) {
    launch_apply<float, RB, false, true, false>(mat32(s), s->nblocks, s->ctx->stream, kNoBatch,
                                                s->ps0, nullptr, nullptr, s->qs, nullptr, RedBuf{},
                                                nullptr);
}

bool ensure_graph_mixed(cicy_sys *s) {
    if (s->graph_exec_mixed) return true;
    cicy_ctx *ctx = s->ctx;
    hipGraph_t g;
    if (!hip_ok(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture"))
        return false;
    CICY_RB_SWITCH(s->rb, launch_chunk<RB>(s));
    if (!hip_ok(ctx, hipStreamEndCapture(ctx->stream, &g), "hipStreamEndCapture")) return false;
    hipError_t e = hipGraphInstantiate(&s->graph_exec_mixed, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return hip_ok(ctx, e, "hipGraphInstantiate");
}

}  // namespace

namespace cicy {

bool mixed_setup(cicy_sys *s, const BlockedHost &h, const BandHost &bh) {
    cicy_ctx *ctx = s->ctx;
    const size_t nv = (size_t)h.n;
    auto alloc = [&](void **p, size_t bytes) {
        if (!hip_ok(ctx, hipMalloc(p, std::max<size_t>(bytes, 4)), "hipMalloc")) return false;
        s->device_bytes += std::max<size_t>(bytes, 4);
        return true;
    };
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return !bytes || hip_ok(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                                    ctx->stream), "H2D");
    };
    std::vector<float> a32, b32;
    bool ok = true;
    if (s->format == kFormatSss) {
        scaled_values(h, a32, b32);
        ok = alloc((void **)&s->val32, 4 * (size_t)h.nnz) &&
             alloc((void **)&s->uval32, 4 * (size_t)h.nremote) &&
             up(s->val32, a32.data(), 4 * a32.size()) && up(s->uval32, b32.data(), 4 * b32.size());
    } else {
        scaled_band(h, bh, a32, b32);
        ok = alloc((void **)&s->band32, 4 * (nv + 1)) &&
             alloc((void **)&s->xval32, 4 * (size_t)bh.m) &&
             up(s->band32, a32.data(), 4 * a32.size()) && up(s->xval32, b32.data(), 4 * b32.size());
    }
    ok = ok && alloc((void **)&s->rs, 4 * nv) && alloc((void **)&s->xs, 4 * nv) &&
         alloc((void **)&s->ps0, 4 * nv) && alloc((void **)&s->ps1, 4 * nv) &&
         alloc((void **)&s->qs, 4 * nv) && alloc((void **)&s->mix, sizeof(MixState));
    if (!ok) return false;
    ok = hip_ok(ctx, hipMemsetAsync(s->mix, 0, sizeof(MixState), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->ps0, 0, 4 * std::max<size_t>(nv, 1), ctx->stream), "memset") &&
         hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return ok;
}

// a^ = a * s_hi * s_lo, s = 1/sqrt(d), hi = the larger index, so both stored copies of an
// entry round to the same FP32 value
void scaled_values(const BlockedHost &h, std::vector<float> &v32, std::vector<float> &u32) {
    std::vector<double> sc((size_t)h.n);
    for (int i = 0; i < h.n; i++) sc[i] = 1.0 / std::sqrt(h.diag[i]);
    v32.resize((size_t)h.nnz);
    u32.resize((size_t)h.nremote);
    for (int r = 0; r < h.n; r++) {
        for (int j = h.rp[r]; j < h.rp[r + 1]; j++) v32[j] = (float)(h.val[j] * sc[r] * sc[h.col[j]]);
        for (int j = h.urp[r]; j < h.urp[r + 1]; j++)
            u32[j] = (float)(h.uval[j] * sc[h.ucol[j]] * sc[r]);
    }
}

void mixed_free(cicy_sys *s) {
    if (s->graph_exec_mixed) (void)hipGraphExecDestroy(s->graph_exec_mixed);
    void *ptrs[] = {s->val32, s->uval32, s->band32, s->xval32, s->rs, s->xs, s->ps0, s->ps1,
                    s->qs, s->mix};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

void mixed_spmv32(cicy_sys *s) { CICY_RB_SWITCH(s->rb, launch_spmv32<RB>(s)); }

bool solve_mixed(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats) {
    cicy_ctx *ctx = s->ctx;
    if (s->n == 0) {
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->status = CICY_OK;
        }
        return true;
    }
    if (rel_to == CICY_REL_B && !s->has_b) {
        set_error(ctx, "CICY_REL_B needs b (cicy_sys_set_rhs)");
        return false;
    }
    if (!ensure_graph_mixed(s)) return false;
    hipStream_t st = ctx->stream;
    const size_t vb = (size_t)s->n * sizeof(double);
    if (!hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "hipEventRecord")) return false;
    if (!hip_ok(ctx, hipMemcpyAsync(s->x, s->x0, vb, hipMemcpyDeviceToDevice, st), "copy x0") ||
        !hip_ok(ctx, hipMemsetAsync(s->st, 0, sizeof(PcgState), st), "memset state"))
        return false;
    if (!s->has_b) {
        // the true residual needs b = r0 + A x0
        CICY_RB_SWITCH(s->rb, launch_spmv64<RB>(s, s->x0, s->q));
        hipLaunchKernelGGL(k_add, dim3((s->n + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                           s->n, s->b, s->r0, s->q);
    }
    double stop_ref2 = rel_to == CICY_REL_B ? s->bnorm2 : -1.0;
    CICY_RB_SWITCH(s->rb, launch_init<RB>(s, tol, stop_ref2, max_iter));
    if (!run_chunks(s, s->graph_exec_mixed)) return false;
    CICY_RB_SWITCH(s->rb, launch_fold_always<RB>(s));
    PcgState h;
    MixState m;
    if (!d2h_sync(ctx, &h, s->st, sizeof(h)) ||
        !d2h_sync(ctx, &m, s->mix, sizeof(m)))
        return false;
    // final residual r = b - A x (the R the caller gets back)
    CICY_RB_SWITCH(s->rb, launch_spmv64<RB>(s, s->x, s->q));
    hipLaunchKernelGGL(k_sub_from, dim3((s->n + kThreads - 1) / kThreads), dim3(kThreads), 0, st,
                       s->n, s->r, s->b, s->q);
    long long iters = h.iters;
    int status = h.status;
    int fallback_iters = 0;
    const bool had_b = s->has_b;
    s->has_b = true;  // b is valid here (given or formed above)
    if (status != CICY_OK && iters < max_iter) {
        // the FP32 recursion stopped making progress, or broke down (p.q no longer positive
        // in FP32 once the scaled residual nears FP32's range): finish in FP64 from x,
        // r = b - A x; a breakdown of the FP64 iteration itself is reported as such
        // stopping against the solve's reference norm (||b||, or ||r0|| of the given start),
        // not against the residual the continuation starts from
        cicy_solve_stats fs;
        s->ref2_override = rel_to == CICY_REL_B ? s->bnorm2 : s->r0norm2;
        const bool ok = solve_pcg64(s, tol, max_iter - iters, rel_to, &fs, true);
        s->ref2_override = -1.0;
        if (!ok) {
            s->has_b = had_b;
            return false;
        }
        iters += fs.iters;
        fallback_iters = (int)fs.iters;
        status = fs.status;
    }
    if (!hip_ok(ctx, hipEventRecord(ctx->ev_stop, st), "hipEventRecord") ||
        !hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "hipEventSynchronize")) {
        s->has_b = had_b;
        return false;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    double tr = -1.0, b2 = 0.0;
    const bool ok = sys_true_residual(s, s->x, &tr, &b2);
    s->has_b = had_b;
    if (!ok) return false;
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iters = iters;
        stats->status = status;
        stats->refinements = m.refinements;
        stats->rel_res = h.rr_init > 0 ? tr * std::sqrt(b2 / h.rr_init) : 0.0;
        stats->solve_ms = ms;
        stats->true_rel_res = tr;
        stats->fp64_iters = fallback_iters;
    }
    return hip_ok(ctx, hipStreamSynchronize(st), "sync");
}

}  // namespace cicy
