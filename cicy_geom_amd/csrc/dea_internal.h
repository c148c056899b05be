// dea_internal.h -- host-side structures shared by the solver translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "cicy_dea.h"
#include "dea_kernels.h"

namespace cicy {

// Host threads for `work` items, at least `grain` items each, at most 16 (or at most
// CICY_HOST_THREADS when set: the set-up's results do not depend on the count).
inline int host_threads(long long work, long long grain) {
    long long hw = std::max(1u, std::thread::hardware_concurrency());
    if (const char *e = getenv("CICY_HOST_THREADS")) hw = std::max(1LL, std::min(hw, atoll(e)));
    return (int)std::max<long long>(1, std::min<long long>({16, hw, work / std::max(1LL, grain)}));
}

// Runs fn(t) for t in [0, nth) on nth host threads (t = 0 on the calling thread).
template <typename F>
void parallel_for(int nth, F &&fn) {
    std::vector<std::thread> th;
    th.reserve((size_t)std::max(0, nth - 1));
    for (int t = 1; t < nth; t++) th.emplace_back([&fn, t]() { fn(t); });
    fn(0);
    for (auto &x : th) x.join();
}

// Wall-clock marks of a system's set-up stages, printed to stderr when CICY_SETUP_TIMING is
// set (a diagnostic of where a fresh system's set-up time goes).
struct StageClock {
    std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
    bool on = getenv("CICY_SETUP_TIMING") != nullptr;
    void mark(const char *what) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        fprintf(stderr, "[cicy setup] %-28s %8.2f ms\n", what,
                std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};

// iterations per captured hipGraph chunk (even: the p double buffer returns to parity 0)
constexpr int kChunk = 32;
// mixed precision: reliable-update trigger r^.r^ <= kDelta^2 * (r^.r^ at the last update),
// and the number of updates in a row without halving ||b - A x|| after which the FP32
// recursion is abandoned (the single-system solver then finishes in FP64)
constexpr double kDelta = 0.1;
constexpr int kStallLimit = 3;

#define CICY_RB_SWITCH(rb, CALL)                              \
    switch (rb) {                                             \
        case 128: { constexpr int RB = 128; CALL; } break;    \
        case 256: { constexpr int RB = 256; CALL; } break;    \
        case 512: { constexpr int RB = 512; CALL; } break;    \
        default: { constexpr int RB = 1024; CALL; } break;    \
    }

struct BlockedHost {
    int n = 0, nnz = 0, rb = 256, nblocks = 0, nremote = 0;
    std::vector<int> rp, col, urp, ucol;
    std::vector<double> val, diag, uval;
};

// Band + remainder layout (BandDev): band[r] = a(r, r-1), the other entries as symmetric
// half-entries sorted by row, xrp[r] = first half-entry of row r, xblk[b] = first
// half-entry of row block b.
struct BandHost {
    int m = 0;
    std::vector<double> band, xval;
    std::vector<int> xblk, xrp, xrow, xnbr;
    std::vector<int> order;  // launch order of the row blocks (BandDev::order)
};

// Host copy of an uploaded system, kept until every layout built on demand from it (FP32
// copy, resident layout, condensation) has been built.
struct SysHost {
    BlockedHost h;
    BandHost bh;
    bool have_band = false;
};

constexpr int kFormatSss = 1;   // blocked SSS with remote transpose list
constexpr int kFormatBand = 2;  // band + remainder

// Device view of one matrix in either layout.
template <typename T>
struct DevMat {
    int format;
    SssDev<T> sss;
    BandDev<T> band;
};

#ifdef __HIPCC__
template <typename T, int RB, bool PCG, bool UNIT, bool BATCH>
inline void launch_apply(const DevMat<T> &m, int grid, hipStream_t st, const BatchDev &bd,
                         const T *src, const T *rvec, T *pnew, T *q, PcgState *stp,
                         const RedBuf &rb, const MixState *guard) {
    if (m.format == kFormatBand)
        hipLaunchKernelGGL((k_apply_band<T, RB, PCG, UNIT, BATCH>), dim3(grid), dim3(kThreads), 0,
                           st, m.band, bd, src, rvec, pnew, q, stp, rb, guard);
    else
        hipLaunchKernelGGL((k_apply<T, RB, PCG, UNIT, BATCH>), dim3(grid), dim3(kThreads), 0, st,
                           m.sss, bd, src, rvec, pnew, q, stp, rb, guard);
}
#endif

}  // namespace cicy

struct cicy_res;   // resident-CG layout (dea_resident.hip)
struct cicy_cond;  // condensed system (dea_condensed.hip)
struct cicy_core;  // block-preconditioned persistent core CG (dea_core.hip)

struct cicy_ctx {
    int device = 0;
    int rows_per_block = 256;
    int format = 0;  // 0 auto, kFormatSss, kFormatBand
    int resident_rows = 0;  // rows per thread of the resident CG layout (0: smallest that fits)
    bool no_graphs = false;  // CICY_NO_GRAPHS set at creation: launch chunks directly
    int dropin_method = CICY_METHOD_PCG_F64;  // method of cicy_solve_dea_cg_sss
    std::vector<int64_t> dropin_blocks;       // row groups for the next block-method call
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_poll[2] = {nullptr, nullptr};
    cicy::PcgState *host_state = nullptr;  // pinned, 2 slots
    double *bc_ws = nullptr;   // boundary-current workspace (grows, kept until destroy)
    double *scratch = nullptr; // this device's copy of the allocation-free reduction words
    size_t bc_words = 0;
    char err[512] = {0};
};

struct cicy_sys {
    cicy_ctx *ctx = nullptr;
    int n = 0, nnz = 0, rb = 256, nblocks = 0, nremote = 0;
    int format = cicy::kFormatSss;
    // blocked SSS, FP64 (format kFormatSss)
    int *rp = nullptr, *col = nullptr, *urp = nullptr, *ucol = nullptr;
    double *val = nullptr, *diag = nullptr, *uval = nullptr;
    // band + remainder (format kFormatBand)
    int nx = 0;
    int *xblk = nullptr, *xrp = nullptr, *xnbr = nullptr;
    double *band = nullptr, *xval = nullptr;
    float *band32 = nullptr, *xval32 = nullptr;
    int *xorder = nullptr;  // launch order of the row blocks
    // FP64 vectors
    double *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    double *x0 = nullptr, *r0 = nullptr, *b = nullptr;
    double *partials = nullptr, *scal = nullptr, *gpart = nullptr;
    unsigned int *gcount = nullptr;
    cicy::PcgState *st = nullptr;
    unsigned int *counter2 = nullptr;
    bool has_b = false;
    double bnorm2 = 0.0;
    bool last_parity_x_valid = false;
    size_t device_bytes = 0;
    hipGraphExec_t graph_exec64 = nullptr;
    // mixed precision: Jacobi-scaled FP32 copy (unit diagonal) + FP32 vectors
    float *val32 = nullptr, *uval32 = nullptr;
    float *rs = nullptr, *xs = nullptr, *ps0 = nullptr, *ps1 = nullptr, *qs = nullptr;
    cicy::MixState *mix = nullptr;
    hipGraphExec_t graph_exec_mixed = nullptr;
    // register-resident FP64 CG (band layout only); res_why says why it is absent
    cicy_res *res = nullptr;
    const char *res_why = nullptr;
    // condensed solve (chain-interior rows eliminated); cond_why says why it is absent
    cicy_cond *cond = nullptr;
    const char *cond_why = nullptr;
    // squared reference norm of the stopping test when >= 0 (set by the condensed solver
    // for its reduced system: the full system's ||b|| or ||r0||)
    double ref2_override = -1.0;
    double r0norm2 = 0.0;  // ||r0||^2 of the last cicy_sys_set_rhs
    // layouts built on first use (cicy_sys_prepare): which have been attempted, the host
    // copy they are built from, and the resident rows per thread asked for at creation
    bool mixed_done = false, res_done = false, cond_done = false;
    cicy::SysHost *host = nullptr;
    int resident_rows_req = 0;
};

namespace cicy {

void set_error(cicy_ctx *ctx, const char *fmt, ...);
bool hip_ok(cicy_ctx *ctx, hipError_t e, const char *what);
bool build_blocked(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                   const int64_t *col_ind, const double *values, const double *diag, int rb,
                   BlockedHost &h);
// `direct` launches one chunk without a graph; used instead of the graph when the context
// was created with CICY_NO_GRAPHS set (profilers that do not follow graph launches)
bool run_chunks(cicy_sys *s, hipGraphExec_t exec, void (*direct)(cicy_sys *) = nullptr);
// device -> host copy ordered after all work queued on the context's stream, then waited
// for (the stream is non-blocking, so a plain hipMemcpy would not wait for it)
bool d2h_sync(cicy_ctx *ctx, void *dst, const void *src, size_t bytes);
bool sys_true_residual(cicy_sys *s, const double *xdev, double *rel, double *b2 = nullptr);
void sys_spmv64(cicy_sys *s, const double *xdev, double *ydev);  // y = A x on the stream
bool solve_pcg64(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats, bool keep_x = false);

// dea_mixed.hip
bool mixed_setup(cicy_sys *s, const BlockedHost &h, const BandHost &bh);
void mixed_free(cicy_sys *s);
bool solve_mixed(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats);
void mixed_spmv32(cicy_sys *s);
void scaled_values(const BlockedHost &h, std::vector<float> &v32, std::vector<float> &u32);
// band layout of a validated blocked system (same row blocks); fraction of lower entries
// on the sub-diagonal
void build_band(const BlockedHost &h, BandHost &b);
double band_fraction(const BlockedHost &h);
void scaled_band(const BlockedHost &h, const BandHost &b, std::vector<float> &band32,
                 std::vector<float> &xval32);
int choose_format(const cicy_ctx *ctx, const BlockedHost &h);
DevMat<double> mat64(const cicy_sys *s);
inline RedBuf redbuf(const cicy_sys *s) { return RedBuf{s->partials, s->gpart, s->gcount}; }
DevMat<float> mat32(const cicy_sys *s);

// dea_resident.hip
cicy_res *resident_setup(cicy_ctx *ctx, const BlockedHost &h, const BandHost &bh, int R_force,
                         const char **why, bool *failed);
void resident_free(cicy_res *z);
bool solve_resident(cicy_sys *s, double tol, long long max_iter, int rel_to,
                    cicy_solve_stats *stats);
int resident_info(const cicy_res *z, int what);

// dea_condensed.hip
cicy_cond *condensed_setup(cicy_ctx *ctx, const BlockedHost &h, const BandHost &bh,
                           const char **why, bool *failed);
void condensed_free(cicy_cond *c);
bool solve_condensed(cicy_sys *s, int inner_method, double tol, long long max_iter, int rel_to,
                     cicy_solve_stats *stats);
int condensed_info(const cicy_cond *c, int what);
int condensed_set_blocks(cicy_sys *s, int64_t nblocks, const int64_t *ranges);

// dea_core.hip
cicy_core *core_setup(cicy_ctx *ctx, int m, const std::vector<int64_t> &crp,
                      const std::vector<int64_t> &ccol, const std::vector<double> &cval,
                      const std::vector<double> &cdiag, const std::vector<int> &orig,
                      int64_t nblocks, const int64_t *ranges);
void core_free(cicy_core *z);
int core_info(const cicy_core *z, int what);
bool solve_core(cicy_ctx *ctx, cicy_core *z, const double *rhs, double tol, long long max_iter,
                double ref2, double *x_out, cicy_solve_stats *stats);
// uploads a system (dea_solver.hip); the layouts of the other methods are built by
// sys_prepare, on the first solve with them at the latest
int sys_create(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
               const int64_t *col_ind, const double *values, const double *diagonal,
               cicy_sys **out);
// builds what `method` needs (idempotent); false on a device or allocation error, or when
// the method cannot run on this system (the reason in the context's error)
// core = false stops after the condensation itself (the core solver's layout is left for
// the next call, e.g. after cicy_sys_set_blocks has chosen its row groups)
bool sys_prepare(cicy_sys *s, int method, bool core = true);
// the core system of a condensed system as an ordinary cicy_sys prepared for `inner`
// (CICY_METHOD_PCG_F64 or CICY_METHOD_RESIDENT); built on first use
bool condensed_core_sys(cicy_sys *s, int inner);

}  // namespace cicy
