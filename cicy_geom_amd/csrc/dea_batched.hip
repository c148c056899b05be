// dea_batched.hip -- many DEA systems of uneven size solved together.
//
// The reference solves the clusters and directions of a window one after another, each
// with a fresh Direct_Electrifying object (Electrical_analysis.cpp:34-105, :249-283).  Here
// a batch stores all systems back to back (block-diagonal, blocked-SSS layout of
// dea_kernels.h) with one PcgState/MixState per system.  Each launch covers the row blocks
// of every unfinished system, so one hipGraph chunk advances all systems by kChunk
// iterations; every system keeps its own alpha/beta, stopping test and iteration count.
// After each chunk the host reads the per-system flags; when the blocks of finished
// systems make up a large part of the grid, the active block list is compacted (double
// buffered, so the chunk already queued keeps its list) and the chunk graph is captured
// again for the smaller grid.
#include "dea_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace cicy;

struct cicy_batch {
    cicy_ctx *ctx = nullptr;
    int nsys = 0, rb = 256;
    int n = 0, nnz = 0, nblocks = 0, nremote = 0;
    std::vector<int> h_sys_row0, h_sys_blk0, h_blk_sys, h_blk_row0;
    std::vector<double> bnorm2, r0norm2;
    // device: matrix (FP64 + Jacobi-scaled FP32 values), tables, vectors, state
    int *rp = nullptr, *col = nullptr, *urp = nullptr, *ucol = nullptr;
    double *val = nullptr, *diag = nullptr, *uval = nullptr;
    float *val32 = nullptr, *uval32 = nullptr;
    int format = kFormatSss, nx = 0;
    int *xblk = nullptr, *xrp = nullptr, *xnbr = nullptr;
    double *band = nullptr, *xval = nullptr;
    float *band32 = nullptr, *xval32 = nullptr;
    int *blk_sys = nullptr, *blk_row0 = nullptr, *sys_row0 = nullptr, *sys_blk0 = nullptr;
    int *active[2] = {nullptr, nullptr};
    int *h_active[2] = {nullptr, nullptr};  // pinned staging
    double *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    double *x0 = nullptr, *r0 = nullptr, *b = nullptr, *stop_ref2 = nullptr;
    float *rs = nullptr, *xs = nullptr, *ps0 = nullptr, *ps1 = nullptr, *qs = nullptr;
    double *partials = nullptr, *norms = nullptr, *gpart = nullptr;
    unsigned int *gcount = nullptr;
    PcgState *st = nullptr;
    MixState *mix = nullptr;
    unsigned int *ncount = nullptr;
    PcgState *h_st[2] = {nullptr, nullptr};  // pinned poll slots
    hipGraphExec_t exec = nullptr;
    int exec_method = -1, exec_grid = -1, exec_buf = -1;
    size_t device_bytes = 0;
};

namespace {

// per-system ||b - A x||^2 and ||b||^2 (ax = A x formed before)
template <int RB>
__global__ __launch_bounds__(kThreads) void k_batch_norms(int n, BatchDev bd,
                                                          const double *__restrict__ bvec,
                                                          const double *__restrict__ ax,
                                                          double *__restrict__ out,
                                                          unsigned int *__restrict__ counters,
                                                          RedBuf rb) {
    const Blk k = block_info<RB, true>(bd, n);
    __shared__ double red[8];
    __shared__ int flag;
    double v[2] = {0.0, 0.0};
    for (int j = threadIdx.x; j < k.nrows; j += kThreads) {
        const int i = k.r0 + j;
        const double d = bvec[i] - ax[i];
        v[0] += d * d;
        v[1] += bvec[i] * bvec[i];
    }
    double tot[2];
    if (reduce_sys<2>(v, k, rb, &counters[k.sys], red, &flag, tot) && threadIdx.x == 0) {
        out[2 * k.sys] = tot[0];
        out[2 * k.sys + 1] = tot[1];
    }
}

RedBuf rbuf(const cicy_batch *B) { return RedBuf{B->partials, B->gpart, B->gcount}; }

BatchDev bdev(const cicy_batch *B, const int *active) {
    return BatchDev{active, B->blk_sys, B->blk_row0, B->sys_row0, B->sys_blk0};
}

DevMat<double> m64(const cicy_batch *B) {
    DevMat<double> m;
    m.format = B->format;
    m.sss = SssDev<double>{B->n, B->nblocks, B->rp, B->col, B->val, B->diag, B->urp, B->ucol,
                           B->uval};
    m.band = BandDev<double>{B->n, B->nblocks, B->band, B->diag, B->xblk, B->xrp, B->xnbr,
                             B->xval};
    return m;
}

DevMat<float> m32(const cicy_batch *B) {
    DevMat<float> m;
    m.format = B->format;
    m.sss = SssDev<float>{B->n, B->nblocks, B->rp, B->col, B->val32, nullptr, B->urp, B->ucol,
                          B->uval32};
    m.band = BandDev<float>{B->n, B->nblocks, B->band32, nullptr, B->xblk, B->xrp, B->xnbr,
                            B->xval32};
    return m;
}

template <int RB>
void chunk_f64(cicy_batch *B, const int *act, int grid) {
    hipStream_t st = B->ctx->stream;
    const BatchDev bd = bdev(B, act);
    for (int it = 0; it < kChunk; it++) {
        double *pold = (it & 1) ? B->p1 : B->p0;
        double *pnew = (it & 1) ? B->p0 : B->p1;
        launch_apply<double, RB, true, false, true>(m64(B), grid, st, bd, pold, B->r, pnew, B->q,
                                                    B->st, rbuf(B), nullptr);
        hipLaunchKernelGGL((k_update_f64<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n, bd,
                           B->diag, B->x, B->r, pnew, B->q, B->st, rbuf(B));
    }
}

template <int RB>
void chunk_mixed(cicy_batch *B, const int *act, int grid) {
    hipStream_t st = B->ctx->stream;
    const BatchDev bd = bdev(B, act);
    for (int it = 0; it < kChunk; it++) {
        float *pold = (it & 1) ? B->ps1 : B->ps0;
        float *pnew = (it & 1) ? B->ps0 : B->ps1;
        launch_apply<float, RB, true, true, true>(m32(B), grid, st, bd, pold, B->rs, pnew, B->qs,
                                                  B->st, rbuf(B), nullptr);
        hipLaunchKernelGGL((k_update_mixed<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n,
                           bd, B->xs, B->rs, pnew, B->qs, B->st, B->mix, rbuf(B));
    }
    hipLaunchKernelGGL((k_fold<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n, bd, B->diag,
                       B->x, B->xs, B->mix, 0);
    launch_apply<double, RB, false, false, true>(m64(B), grid, st, bd, B->x, nullptr, nullptr,
                                                 B->q, nullptr, RedBuf{}, B->mix);
    hipLaunchKernelGGL((k_reliable_resid<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n, bd,
                       B->diag, B->b, B->q, B->r, B->rs, B->st, B->mix, rbuf(B),
                       kStallLimit);
}

template <int RB>
void init_blocks(cicy_batch *B, int method, double tol, long long max_iter, const int *act,
                 int grid) {
    hipStream_t st = B->ctx->stream;
    const BatchDev bd = bdev(B, act);
    if (method == CICY_METHOD_PCG_F64)
        hipLaunchKernelGGL((k_init_f64<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n, bd,
                           B->diag, B->r, B->p0, B->p1, B->st, rbuf(B), tol, B->stop_ref2,
                           -1.0, max_iter);
    else
        hipLaunchKernelGGL((k_mixed_init<RB, true>), dim3(grid), dim3(kThreads), 0, st, B->n, bd,
                           B->diag, B->r0, B->rs, B->xs, B->ps0, B->ps1, B->st, B->mix,
                           rbuf(B), tol, B->stop_ref2, -1.0, max_iter, kDelta * kDelta);
}

// r = b - A x over the blocks of `act` (q = A x formed by the caller)
template <int RB>
__global__ __launch_bounds__(kThreads) void k_batch_b_minus(int n, BatchDev bd,
                                                            const double *__restrict__ b,
                                                            const double *__restrict__ ax,
                                                            double *__restrict__ r) {
    const Blk k = block_info<RB, true>(bd, n);
    for (int t = threadIdx.x; t < k.nrows; t += kThreads) r[k.r0 + t] = b[k.r0 + t] - ax[k.r0 + t];
}

template <int RB>
void fold_all(cicy_batch *B) {
    hipLaunchKernelGGL((k_fold<RB, true>), dim3(B->nblocks), dim3(kThreads), 0, B->ctx->stream,
                       B->n, bdev(B, nullptr), B->diag, B->x, B->xs, B->mix, 1);
}

template <int RB>
void resid_blocks(cicy_batch *B, const int *act, int grid) {
    hipStream_t st = B->ctx->stream;
    const BatchDev bd = bdev(B, act);
    launch_apply<double, RB, false, false, true>(m64(B), grid, st, bd, B->x, nullptr, nullptr,
                                                 B->q, nullptr, RedBuf{}, nullptr);
    hipLaunchKernelGGL((k_batch_b_minus<RB>), dim3(grid), dim3(kThreads), 0, st, B->n, bd, B->b,
                       B->q, B->r);
}

template <int RB>
void final_norms(cicy_batch *B) {
    hipStream_t st = B->ctx->stream;
    const BatchDev bd = bdev(B, nullptr);
    hipLaunchKernelGGL((k_fold<RB, true>), dim3(B->nblocks), dim3(kThreads), 0, st, B->n, bd,
                       B->diag, B->x, B->xs, B->mix, 1);
    launch_apply<double, RB, false, false, true>(m64(B), B->nblocks, st, bd, B->x, nullptr,
                                                 nullptr, B->q, nullptr, RedBuf{}, nullptr);
    hipLaunchKernelGGL((k_batch_norms<RB>), dim3(B->nblocks), dim3(kThreads), 0, st, B->n, bd,
                       B->b, B->q, B->norms, B->ncount, rbuf(B));
}

bool capture(cicy_batch *B, int method, int buf, int grid) {
    cicy_ctx *ctx = B->ctx;
    if (B->exec && B->exec_method == method && B->exec_grid == grid && B->exec_buf == buf)
        return true;
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    B->exec = nullptr;
    hipGraph_t g;
    if (!hip_ok(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture"))
        return false;
    const int *act = B->active[buf];
    if (method == CICY_METHOD_PCG_F64) {
        CICY_RB_SWITCH(B->rb, chunk_f64<RB>(B, act, grid));
    } else {
        CICY_RB_SWITCH(B->rb, chunk_mixed<RB>(B, act, grid));
    }
    if (!hip_ok(ctx, hipStreamEndCapture(ctx->stream, &g), "hipStreamEndCapture")) return false;
    hipError_t e = hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (!hip_ok(ctx, e, "hipGraphInstantiate")) return false;
    B->exec_method = method;
    B->exec_grid = grid;
    B->exec_buf = buf;
    return true;
}

// Active block list of the unfinished systems, into pinned slot `buf`; returns its length.
int build_active(cicy_batch *B, const PcgState *hs, int buf) {
    int m = 0;
    for (int s = 0; s < B->nsys; s++) {
        if (hs[s].done) continue;
        for (int b = B->h_sys_blk0[s]; b < B->h_sys_blk0[s + 1]; b++) B->h_active[buf][m++] = b;
    }
    return m;
}

template <typename T>
bool balloc(cicy_batch *B, T **p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (!hip_ok(B->ctx, hipMalloc((void **)p, bytes), "hipMalloc")) return false;
    B->device_bytes += bytes;
    return true;
}

template <typename T>
bool bh2d(cicy_batch *B, T *dst, const T *src, size_t count) {
    if (!count) return true;
    return hip_ok(B->ctx, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice,
                                         B->ctx->stream), "H2D");
}

}  // namespace

extern "C" {

int cicy_batch_create(cicy_ctx *ctx, int nsys, const int64_t *n, const int64_t *nnz,
                      const int64_t *const *row_ptr, const int64_t *const *col_ind,
                      const double *const *values, const double *const *diag,
                      cicy_batch **out) {
    if (!ctx || !out || nsys <= 0 || !n || !nnz || !row_ptr || !col_ind || !values || !diag) {
        if (ctx) set_error(ctx, "cicy_batch_create: invalid arguments");
        return 0;
    }
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    const int rb = ctx->rows_per_block;
    std::vector<BlockedHost> hs((size_t)nsys);
    long long tn = 0, tnnz = 0, tb = 0, trem = 0;
    for (int s = 0; s < nsys; s++) {
        if (!build_blocked(ctx, n[s], nnz[s], row_ptr[s], col_ind[s], values[s], diag[s], rb,
                           hs[s])) {
            char msg[600];
            snprintf(msg, sizeof(msg), "system %d: %s", s, ctx->err);
            set_error(ctx, "%s", msg);
            return 0;
        }
        tn += hs[s].n;
        tnnz += hs[s].nnz;
        tb += hs[s].nblocks;
        trem += hs[s].nremote;
    }
    if (tn >= (1LL << 31) - 1 || tnnz >= (1LL << 31) - 1 || trem >= (1LL << 31) - 1) {
        set_error(ctx, "batch too large for 32-bit device indices");
        return 0;
    }
    cicy_batch *B = new cicy_batch();
    B->ctx = ctx;
    B->nsys = nsys;
    B->rb = rb;
    B->n = (int)tn;
    B->nnz = (int)tnnz;
    B->nblocks = (int)tb;
    B->nremote = (int)trem;
    B->bnorm2.assign((size_t)nsys, -1.0);
    B->r0norm2.assign((size_t)nsys, 0.0);
    // one layout for the whole batch: band when most entries of the batch are sub-diagonal
    // (the band layout unless the blocked SSS layout is asked for, as for single systems)
    B->format = ctx->format == kFormatSss ? kFormatSss : kFormatBand;
    const bool band_fmt = B->format == kFormatBand;
    std::vector<BandHost> bhs(band_fmt ? (size_t)nsys : 0);
    long long tx = 0;
    if (band_fmt) {
        for (int s = 0; s < nsys; s++) {
            build_band(hs[s], bhs[s]);
            tx += bhs[s].m;
        }
        if (tx >= (1LL << 31) - 1) {
            set_error(ctx, "batch too large for 32-bit device indices");
            delete B;
            return 0;
        }
    }
    B->nx = (int)tx;
    const size_t nsss = band_fmt ? 0 : 1;  // SSS arrays only for the SSS layout
    std::vector<int> rp(nsss * ((size_t)tn + 1)), col(nsss * (size_t)tnnz),
        urp(nsss * ((size_t)tn + 1)), ucol(nsss * (size_t)trem);
    std::vector<double> val(nsss * (size_t)tnnz), dg((size_t)tn), uval(nsss * (size_t)trem);
    std::vector<float> val32(nsss * (size_t)tnnz), uval32(nsss * (size_t)trem);
    const size_t nb1 = band_fmt ? 1 : 0;
    std::vector<double> band(nb1 * ((size_t)tn + 1), 0.0), xval(nb1 * (size_t)tx);
    std::vector<float> band32(nb1 * ((size_t)tn + 1), 0.f), xval32(nb1 * (size_t)tx);
    std::vector<int> xblk(nb1 * ((size_t)tb + 1)), xrp(nb1 * ((size_t)tn + 1)),
        xnbr(nb1 * (size_t)tx);
    int xo = 0;
    B->h_sys_row0.assign((size_t)nsys + 1, 0);
    B->h_sys_blk0.assign((size_t)nsys + 1, 0);
    int ro = 0, jo = 0, uo = 0, bo = 0;
    for (int s = 0; s < nsys; s++) {
        const BlockedHost &h = hs[s];
        for (int i = 0; i < h.n; i++) dg[ro + i] = h.diag[i];
        if (!band_fmt) {
            std::vector<float> a32, u32;
            scaled_values(h, a32, u32);
            for (int i = 0; i < h.n; i++) {
                rp[ro + i] = jo + h.rp[i];
                urp[ro + i] = uo + h.urp[i];
            }
            for (int j = 0; j < h.nnz; j++) {
                col[jo + j] = ro + h.col[j];
                val[jo + j] = h.val[j];
                val32[jo + j] = a32[j];
            }
            for (int j = 0; j < h.nremote; j++) {
                ucol[uo + j] = ro + h.ucol[j];
                uval[uo + j] = h.uval[j];
                uval32[uo + j] = u32[j];
            }
        } else {
            const BandHost &bh = bhs[s];
            std::vector<float> a32, x32;
            scaled_band(h, bh, a32, x32);
            for (int i = 0; i < h.n; i++) {
                band[ro + i] = bh.band[i];
                band32[ro + i] = a32[i];
            }
            for (int i = 0; i < h.n; i++) xrp[ro + i] = xo + bh.xrp[i];
            for (int k = 0; k < bh.m; k++) {
                xnbr[xo + k] = ro + bh.xnbr[k];
                xval[xo + k] = bh.xval[k];
                xval32[xo + k] = x32[k];
            }
            for (int k = 0; k < h.nblocks; k++) xblk[bo + k] = xo + bh.xblk[k];
            xo += bh.m;
        }
        for (int k = 0; k < h.nblocks; k++) {
            B->h_blk_sys.push_back(s);
            B->h_blk_row0.push_back(ro + k * rb);
        }
        ro += h.n;
        jo += h.nnz;
        uo += h.nremote;
        bo += h.nblocks;
        B->h_sys_row0[s + 1] = ro;
        B->h_sys_blk0[s + 1] = bo;
    }
    const size_t N = (size_t)tn, NB = (size_t)tb;
    bool ok = balloc(B, &B->diag, N);
    if (!band_fmt) {
        rp[tn] = jo;
        urp[tn] = uo;
        ok = ok && balloc(B, &B->rp, N + 1) && balloc(B, &B->col, (size_t)tnnz) &&
             balloc(B, &B->val, (size_t)tnnz) && balloc(B, &B->urp, N + 1) &&
             balloc(B, &B->ucol, (size_t)trem) && balloc(B, &B->uval, (size_t)trem) &&
             balloc(B, &B->val32, (size_t)tnnz) && balloc(B, &B->uval32, (size_t)trem) &&
             bh2d(B, B->rp, rp.data(), N + 1) && bh2d(B, B->col, col.data(), (size_t)tnnz) &&
             bh2d(B, B->val, val.data(), (size_t)tnnz) && bh2d(B, B->urp, urp.data(), N + 1) &&
             bh2d(B, B->ucol, ucol.data(), (size_t)trem) &&
             bh2d(B, B->uval, uval.data(), (size_t)trem) &&
             bh2d(B, B->val32, val32.data(), (size_t)tnnz) &&
             bh2d(B, B->uval32, uval32.data(), (size_t)trem);
    } else {
        xblk[tb] = xo;
        xrp[tn] = xo;
        ok = ok && balloc(B, &B->band, N + 1) && balloc(B, &B->band32, N + 1) &&
             balloc(B, &B->xblk, NB + 1) && balloc(B, &B->xrp, N + 1) &&
             balloc(B, &B->xnbr, (size_t)tx) && balloc(B, &B->xval, (size_t)tx) &&
             balloc(B, &B->xval32, (size_t)tx) && bh2d(B, B->band, band.data(), N + 1) &&
             bh2d(B, B->band32, band32.data(), N + 1) && bh2d(B, B->xblk, xblk.data(), NB + 1) &&
             bh2d(B, B->xrp, xrp.data(), N + 1) && bh2d(B, B->xnbr, xnbr.data(), (size_t)tx) &&
             bh2d(B, B->xval, xval.data(), (size_t)tx) &&
             bh2d(B, B->xval32, xval32.data(), (size_t)tx);
    }
    ok = ok && balloc(B, &B->blk_sys, NB) &&
              balloc(B, &B->blk_row0, NB) && balloc(B, &B->sys_row0, (size_t)nsys + 1) &&
              balloc(B, &B->sys_blk0, (size_t)nsys + 1) && balloc(B, &B->active[0], NB) &&
              balloc(B, &B->active[1], NB) && balloc(B, &B->x, N) && balloc(B, &B->r, N) &&
              balloc(B, &B->p0, N) && balloc(B, &B->p1, N) && balloc(B, &B->q, N) &&
              balloc(B, &B->x0, N) && balloc(B, &B->r0, N) && balloc(B, &B->b, N) &&
              balloc(B, &B->stop_ref2, (size_t)nsys) && balloc(B, &B->rs, N) &&
              balloc(B, &B->xs, N) && balloc(B, &B->ps0, N) && balloc(B, &B->ps1, N) &&
              balloc(B, &B->qs, N) && balloc(B, &B->partials, 2 * NB) &&
              balloc(B, &B->gpart, 2 * NB) && balloc(B, &B->gcount, NB) &&
              balloc(B, &B->norms, 2 * (size_t)nsys) && balloc(B, &B->st, (size_t)nsys) &&
              balloc(B, &B->mix, (size_t)nsys) && balloc(B, &B->ncount, (size_t)nsys);
    ok = ok && hip_ok(ctx, hipHostMalloc((void **)&B->h_active[0], NB * sizeof(int),
                                         hipHostMallocDefault), "hipHostMalloc") &&
         hip_ok(ctx, hipHostMalloc((void **)&B->h_active[1], NB * sizeof(int),
                                   hipHostMallocDefault), "hipHostMalloc") &&
         hip_ok(ctx, hipHostMalloc((void **)&B->h_st[0], nsys * sizeof(PcgState),
                                   hipHostMallocDefault), "hipHostMalloc") &&
         hip_ok(ctx, hipHostMalloc((void **)&B->h_st[1], nsys * sizeof(PcgState),
                                   hipHostMallocDefault), "hipHostMalloc");
    ok = ok && bh2d(B, B->diag, dg.data(), N) &&
         bh2d(B, B->blk_sys, B->h_blk_sys.data(), NB) &&
         bh2d(B, B->blk_row0, B->h_blk_row0.data(), NB) &&
         bh2d(B, B->sys_row0, B->h_sys_row0.data(), (size_t)nsys + 1) &&
         bh2d(B, B->sys_blk0, B->h_sys_blk0.data(), (size_t)nsys + 1) &&
         hip_ok(ctx, hipMemsetAsync(B->x0, 0, N * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->r0, 0, N * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->b, 0, N * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->xs, 0, N * sizeof(float), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->ncount, 0, nsys * sizeof(unsigned), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->gcount, 0, NB * sizeof(unsigned), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->partials, 0, 2 * NB * sizeof(double), ctx->stream),
                "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->gpart, 0, 2 * NB * sizeof(double), ctx->stream),
                "memset") &&
         hip_ok(ctx, hipMemsetAsync(B->mix, 0, nsys * sizeof(MixState), ctx->stream), "memset") &&
         hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!ok) {
        cicy_batch_destroy(B);
        return 0;
    }
    *out = B;
    return 1;
}

void cicy_batch_destroy(cicy_batch *B) {
    if (!B) return;
    (void)hipSetDevice(B->ctx->device);
    (void)hipStreamSynchronize(B->ctx->stream);
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    void *dev[] = {B->rp, B->col, B->urp, B->ucol, B->val, B->diag, B->uval, B->val32, B->uval32,
                   B->band, B->band32, B->xblk, B->xrp, B->xnbr, B->xval, B->xval32,
                   B->blk_sys, B->blk_row0, B->sys_row0, B->sys_blk0, B->active[0],
                   B->active[1], B->x, B->r, B->p0, B->p1, B->q, B->x0, B->r0, B->b,
                   B->stop_ref2, B->rs, B->xs, B->ps0, B->ps1, B->qs, B->partials, B->norms,
                   B->gpart, B->gcount,
                   B->st, B->mix, B->ncount};
    for (void *p : dev)
        if (p) (void)hipFree(p);
    void *host[] = {B->h_active[0], B->h_active[1], B->h_st[0], B->h_st[1]};
    for (void *p : host)
        if (p) (void)hipHostFree(p);
    delete B;
}

int64_t cicy_batch_size(const cicy_batch *B, int what) {
    if (!B) return -1;
    switch (what) {
        case 0: return B->n;
        case 1: return B->nnz;
        case 2: return B->nremote;
        case 3: return (int64_t)B->device_bytes;
        case 4: return B->nsys;
        case 5: return B->nblocks;
        case 6: return B->format;
        default: return -1;
    }
}

int cicy_batch_set_rhs(cicy_batch *B, int sys, const double *r0, const double *x0,
                       const double *b) {
    if (!B || sys < 0 || sys >= B->nsys || !r0) return 0;
    cicy_ctx *ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    const size_t off = (size_t)B->h_sys_row0[sys];
    const size_t ns = (size_t)(B->h_sys_row0[sys + 1] - B->h_sys_row0[sys]);
    bool ok = bh2d(B, B->r0 + off, r0, ns);
    if (x0) ok = ok && bh2d(B, B->x0 + off, x0, ns);
    else ok = ok && hip_ok(ctx, hipMemsetAsync(B->x0 + off, 0, ns * sizeof(double), ctx->stream), "memset");
    // b is needed for the true residual; with x0 == 0 it equals r0
    const double *bb = b ? b : (x0 ? nullptr : r0);
    if (!bb) {
        set_error(ctx, "cicy_batch_set_rhs: b is required when x0 is given");
        return 0;
    }
    ok = ok && bh2d(B, B->b + off, bb, ns);
    double nb2 = 0.0, r2 = 0.0;
    for (size_t i = 0; i < ns; i++) nb2 += bb[i] * bb[i];
    for (size_t i = 0; i < ns; i++) r2 += r0[i] * r0[i];
    B->bnorm2[sys] = nb2;
    B->r0norm2[sys] = r2;
    ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return ok ? 1 : 0;
}

}  // extern "C"

namespace {

// Runs chunks of `method` over the blocks in pinned list `buf` (length grid) until every
// system in it is done; compacts the list as systems finish.
bool batch_loop(cicy_batch *B, int method, int buf, int grid) {
    cicy_ctx *ctx = B->ctx;
    hipStream_t st = ctx->stream;
    if (grid <= 0) return true;
    double max_s = 600.0;
    if (const char *e = getenv("CICY_MAX_SOLVE_SECONDS")) max_s = atof(e);
    const auto t0 = std::chrono::steady_clock::now();
    if (!bh2d(B, B->active[buf], B->h_active[buf], (size_t)grid) || !capture(B, method, buf, grid))
        return false;
    // keep one chunk queued ahead of the poll
    int slot = 0;
    if (!hip_ok(ctx, hipGraphLaunch(B->exec, st), "hipGraphLaunch") ||
        !hip_ok(ctx, hipMemcpyAsync(B->h_st[slot], B->st, B->nsys * sizeof(PcgState),
                                    hipMemcpyDeviceToHost, st), "poll") ||
        !hip_ok(ctx, hipEventRecord(ctx->ev_poll[slot], st), "event"))
        return false;
    for (;;) {
        const int nslot = slot ^ 1;
        if (!hip_ok(ctx, hipGraphLaunch(B->exec, st), "hipGraphLaunch") ||
            !hip_ok(ctx, hipMemcpyAsync(B->h_st[nslot], B->st, B->nsys * sizeof(PcgState),
                                        hipMemcpyDeviceToHost, st), "poll") ||
            !hip_ok(ctx, hipEventRecord(ctx->ev_poll[nslot], st), "event"))
            return false;
        if (!hip_ok(ctx, hipEventSynchronize(ctx->ev_poll[slot]), "event sync")) return false;
        const PcgState *hs = B->h_st[slot];
        int live_blocks = 0;
        for (int k = 0; k < grid; k++)
            if (!hs[B->h_blk_sys[B->h_active[buf][k]]].done) live_blocks++;
        if (live_blocks == 0) break;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > max_s) {
            set_error(ctx, "batch solve stopped after %.0f s (CICY_MAX_SOLVE_SECONDS)", max_s);
            break;
        }
        if (live_blocks < grid * 7 / 10) {
            // compact into the other list; the queued chunk still reads the current one
            const int nb = buf ^ 1;
            int m = 0;
            for (int k = 0; k < grid; k++) {
                const int b = B->h_active[buf][k];
                if (!hs[B->h_blk_sys[b]].done) B->h_active[nb][m++] = b;
            }
            grid = m;
            if (!bh2d(B, B->active[nb], B->h_active[nb], (size_t)grid) ||
                !capture(B, method, nb, grid))
                return false;
            buf = nb;
        }
        slot = nslot;
    }
    return hip_ok(ctx, hipStreamSynchronize(st), "sync");
}

}  // namespace

extern "C" {

int cicy_batch_solve(cicy_batch *B, int method, double tol, int64_t max_iter, int rel_to,
                     cicy_solve_stats *per_sys, double *total_ms) {
    if (!B) return 0;
    cicy_ctx *ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    if (method != CICY_METHOD_PCG_F64 && method != CICY_METHOD_MIXED) {
        set_error(ctx, "unknown method %d", method);
        return 0;
    }
    if (max_iter <= 0) max_iter = 300LL * ((long long)B->n + 2);
    hipStream_t st = ctx->stream;
    const size_t N = (size_t)B->n;
    std::vector<double> ref2((size_t)B->nsys, -1.0);
    if (rel_to == CICY_REL_B) ref2 = B->bnorm2;
    if (!hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "event") ||
        !hip_ok(ctx, hipMemcpyAsync(B->stop_ref2, ref2.data(), ref2.size() * sizeof(double),
                                    hipMemcpyHostToDevice, st), "H2D") ||
        !hip_ok(ctx, hipMemcpyAsync(B->r, B->r0, N * sizeof(double), hipMemcpyDeviceToDevice, st),
                "copy r0") ||
        !hip_ok(ctx, hipMemcpyAsync(B->x, B->x0, N * sizeof(double), hipMemcpyDeviceToDevice, st),
                "copy x0") ||
        !hip_ok(ctx, hipMemsetAsync(B->st, 0, B->nsys * sizeof(PcgState), st), "memset"))
        return 0;
    CICY_RB_SWITCH(B->rb, init_blocks<RB>(B, method, tol, max_iter, nullptr, B->nblocks));
    if (!hip_ok(ctx, hipStreamSynchronize(st), "sync")) return 0;  // ref2 is pageable
    if (!d2h_sync(ctx, B->h_st[0], B->st, B->nsys * sizeof(PcgState)))
        return 0;
    int buf = 0;
    int grid = build_active(B, B->h_st[0], buf);
    if (!batch_loop(B, method, buf, grid)) return 0;
    std::vector<PcgState> hs((size_t)B->nsys);
    std::vector<MixState> hm((size_t)B->nsys);
    std::vector<long long> fp64_iters((size_t)B->nsys, 0), iters((size_t)B->nsys, 0);
    std::vector<double> norms(2 * (size_t)B->nsys);
    auto read_state = [&]() {
        return d2h_sync(ctx, hs.data(), B->st, hs.size() * sizeof(PcgState)) &&
               d2h_sync(ctx, hm.data(), B->mix, hm.size() * sizeof(MixState));
    };
    auto read_norms = [&]() {
        CICY_RB_SWITCH(B->rb, final_norms<RB>(B));
        return d2h_sync(ctx, norms.data(), B->norms, norms.size() * sizeof(double));
    };
    auto true_rel = [&](int s) {
        return norms[2 * s + 1] > 0 ? std::sqrt(norms[2 * s] / norms[2 * s + 1])
                                    : std::sqrt(norms[2 * s]);
    };
    // final_norms folds the FP32 corrections (mixed) and forms every ||b - A x||
    if (!read_state() || !read_norms()) return 0;
    for (int s = 0; s < B->nsys; s++) iters[s] = hs[s].iters;
    // Continue in FP64, from x with r = b - A x, every system whose FP32 recursion stalled
    // or broke down (status CICY_MAXITER below the cap, or CICY_BREAKDOWN with a finite
    // residual) or whose true residual misses tol*||b||.
    for (int pass = 0; pass < 3; pass++) {
        int m = 0;
        std::vector<char> redo((size_t)B->nsys, 0);
        for (int s = 0; s < B->nsys; s++) {
            // an FP32 breakdown (p.q no longer positive in FP32) is continued in FP64 too
            const bool stalled = method == CICY_METHOD_MIXED && pass == 0 &&
                                 (hs[s].status == CICY_MAXITER ||
                                  hs[s].status == CICY_BREAKDOWN) &&
                                 iters[s] < max_iter && std::isfinite(true_rel(s));
            const bool short_of_b = rel_to == CICY_REL_B && hs[s].status == CICY_OK &&
                                    true_rel(s) > tol;
            if (!stalled && !short_of_b) continue;
            redo[s] = 1;
            for (int b = B->h_sys_blk0[s]; b < B->h_sys_blk0[s + 1]; b++) B->h_active[0][m++] = b;
        }
        if (m == 0) break;
        // the continuation stops against the same reference norm as the solve (||b|| or the
        // given ||r0||), not against its own starting residual
        if (pass == 0) {
            const std::vector<double> &cref = rel_to == CICY_REL_B ? B->bnorm2 : B->r0norm2;
            if (!bh2d(B, B->stop_ref2, cref.data(), cref.size())) return 0;
        }
        if (!bh2d(B, B->active[0], B->h_active[0], (size_t)m)) return 0;
        CICY_RB_SWITCH(B->rb, resid_blocks<RB>(B, B->active[0], m));
        CICY_RB_SWITCH(B->rb, init_blocks<RB>(B, CICY_METHOD_PCG_F64, tol, max_iter,
                                              B->active[0], m));
        if (!batch_loop(B, CICY_METHOD_PCG_F64, 0, m)) return 0;
        if (!read_state() || !read_norms()) return 0;
        for (int s = 0; s < B->nsys; s++)
            if (redo[s]) {
                iters[s] += hs[s].iters;
                fp64_iters[s] += hs[s].iters;
            }
    }
    if (!hip_ok(ctx, hipEventRecord(ctx->ev_stop, st), "event") ||
        !hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "sync"))
        return 0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    if (total_ms) *total_ms = ms;
    if (per_sys) {
        for (int s = 0; s < B->nsys; s++) {
            cicy_solve_stats &o = per_sys[s];
            memset(&o, 0, sizeof(o));
            const bool mixed = method == CICY_METHOD_MIXED;
            o.iters = iters[s];
            o.status = hs[s].status;
            o.refinements = mixed ? hm[s].refinements : 0;
            o.fp64_iters = fp64_iters[s];
            o.true_rel_res = true_rel(s);
            o.rel_res = (!mixed && fp64_iters[s] == 0 && hs[s].rr_init > 0)
                            ? std::sqrt(hs[s].rr / hs[s].rr_init) : o.true_rel_res;
            o.solve_ms = ms;
        }
    }
    return 1;
}

int cicy_batch_get(cicy_batch *B, int sys, double *x, double *r) {
    if (!B || sys < 0 || sys >= B->nsys) return 0;
    cicy_ctx *ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    const size_t off = (size_t)B->h_sys_row0[sys];
    const size_t ns = (size_t)(B->h_sys_row0[sys + 1] - B->h_sys_row0[sys]);
    bool ok = true;
    if (x && ns)
        ok = ok && hip_ok(ctx, hipMemcpyAsync(x, B->x + off, ns * sizeof(double),
                                              hipMemcpyDeviceToHost, ctx->stream), "D2H");
    if (r && ns)
        ok = ok && hip_ok(ctx, hipMemcpyAsync(r, B->r + off, ns * sizeof(double),
                                              hipMemcpyDeviceToHost, ctx->stream), "D2H");
    ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return ok ? 1 : 0;
}

}  // extern "C"
