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
struct cicy_core;  // block-preconditioned persistent core CG (dea_core.hip)

namespace cicy {

// Device memory reused from one set-up to the next: blocks are handed out best-fit and
// returned when a set-up ends, so repeated set-ups of systems of one size allocate nothing
// after the first.  Held by the context until it is destroyed.
struct DevPool {
    struct Blk {
        void *p;
        size_t bytes;
        bool used;
    };
    std::vector<Blk> blocks;
    void *get(size_t bytes);  // nullptr when hipMalloc fails
    bool put(void *p);        // false when p is not a block of this pool
    void clear();             // frees every block
};

// Nodes of the reduced system with at most this many neighbours are eliminated exactly
// (series reduction / star-mesh transform) before the core CG.
constexpr int kElimDegree = 6;
// the degree cap in effect: kElimDegree, or CICY_ELIM_DEGREE (2..8) when set (a tuning knob;
// host and device forms read the same value)
inline int elim_degree() {
    if (const char *e = getenv("CICY_ELIM_DEGREE")) return std::max(2, std::min(8, atoi(e)));
    return kElimDegree;
}

// The series rounds stop at the first that would eliminate fewer than 1 / this many of the
// nodes left (such a round costs a pass over the whole reduced system for little gain);
// CICY_ROUND_FRACTION (2..1000) overrides it (a tuning knob; both forms read it)
constexpr int kMinRoundFraction = 100;
inline long long min_round_fraction() {
    if (const char *e = getenv("CICY_ROUND_FRACTION"))
        return std::max(2, std::min(1000, atoi(e)));
    return kMinRoundFraction;
}

// Whether the series rounds find their independent set in one cooperative launch
// (CICY_MIS_ASYNC=0 makes them use batches of passes instead; the set is the same).
inline bool mis_async_enabled() {
    const char *e = getenv("CICY_MIS_ASYNC");
    return !(e && e[0] == '0');
}

// Elimination priority of node k of the reduced system in the series rounds: a fixed
// integer hash of k with k itself as tie-break (lower value = earlier).  The rounds take
// the greedy independent set of their candidates in this order, which the device finds in
// a few parallel passes and the host form by one sweep in the same order.
__host__ __device__ inline unsigned long long elim_priority(int k) {
    unsigned int h = (unsigned int)k;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return ((unsigned long long)h << 32) | (unsigned int)k;
}

}  // namespace cicy

// The condensed system (dea_condensed.hip), set up on the device (dea_setup.hip) or, for
// the blocked SSS layout, on the host.
struct cicy_cond {
    cicy_ctx *ctx = nullptr;    // its arrays come from the context's pool
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
    // stage 2 (series rounds over S) and the core system the CG runs on
    int mc = 0, nrec = 0;
    long long nnz_core = 0;
    std::vector<int> round_off;                       // host copy: launch bounds per round
    std::vector<int> rv_round_off;                    // host copy: receivers of each round
    int *e_node = nullptr, *e_off = nullptr, *e_nbr = nullptr, *core_nodes = nullptr;
    // forward sweep by receiver: rv_node[j] gathers rv_a[c] * f_k / d_k over its
    // contributions c in rv_off[j] .. rv_off[j+1] (record rv_rec[c]), in record order
    int *rv_node = nullptr, *rv_off = nullptr, *rv_rec = nullptr;
    double *rv_a = nullptr;
    double *e_a = nullptr, *e_d = nullptr;
    double *fS = nullptr, *xS = nullptr;              // [m] right-hand side / solution on S
    double *core_b = nullptr, *core_x = nullptr;      // [mc] the block solver's core vectors
    cicy_sys *red = nullptr;    // the core system on the device (built on first use)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // timing (the inner solve uses the context's)
    // host copy of the core (strictly-lower SSS, diagonal, full-system row of each core
    // row) for the block-preconditioned core CG, built on demand
    std::vector<int64_t> h_crp, h_ccol;
    std::vector<double> h_cval, h_cdiag;
    std::vector<int> h_orig;
    // the core on the device (device set-up): both triangles with columns ascending
    // (d_frp, d_fcol, d_fval), its diagonal and the full-system row of each core row
    int *d_frp = nullptr, *d_fcol = nullptr, *d_orig = nullptr;
    double *d_fval = nullptr, *d_cdiag = nullptr;
    bool on_device = false;
    cicy_core *blk = nullptr;
};

struct cicy_ctx {
    int device = 0;
    int rows_per_block = 256;
    int format = 0;  // 0 auto, kFormatSss, kFormatBand
    int resident_rows = 0;  // rows per thread of the resident CG layout (0: smallest that fits)
    bool no_graphs = false;  // CICY_NO_GRAPHS set at creation: launch chunks directly
    int dropin_method = CICY_METHOD_PCG_F64;  // method of cicy_solve_dea_cg_sss
    std::vector<int64_t> dropin_blocks;       // row groups for the next block-method call
    hipStream_t stream = nullptr;
    // a second stream (same priority) for set-up work that runs beside the first, joined
    // back through ev_fork / ev_join
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_poll[2] = {nullptr, nullptr};
    cicy::PcgState *host_state = nullptr;  // pinned, 2 slots
    void *readback = nullptr;               // pinned, kReadbackBytes: small read-backs land here
    char *upload = nullptr;                 // pinned, kUploadBytes: staging of small uploads
    size_t upload_used = 0;                 // bytes of `upload` whose copies may be in flight
    double *bc_ws = nullptr;   // boundary-current workspace (grows, kept until destroy)
    double *scratch = nullptr; // this device's copy of the allocation-free reduction words
    size_t bc_words = 0;
    bool setup_on_host = false;  // cicy_set_setup: build layouts and condensation on the host
    cicy::DevPool pool;          // temporaries of the device set-up
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
// host -> device copy on stream st; on the context's stream a small one is staged in the
// context's pinned upload buffer (src may be released on return either way)
bool h2d_async(cicy_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st);
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

// Returns a device block to the context's pool (or frees it when it is not one of the
// pool's blocks).  A system's arrays, its condensation's and its core's come from the pool,
// so a new system of the size of an earlier one allocates nothing.
void dev_release(cicy_ctx *ctx, void *p);

// dea_setup.hip: the band + remainder layout from the caller's device arrays (validated on
// the device; false with the context's error set on a malformed system), and the
// condensation of a system whose band layout is on the device
bool band_from_device(cicy_sys *s, const int64_t *row_ptr, const int64_t *col_ind,
                      const double *values, const double *diag);
cicy_cond *condensed_setup_device(cicy_sys *s, const char **why, bool *failed);
void condensed_free_device(cicy_cond *c);
// device memory for one set-up, returned to the context's pool when the scope ends
struct PoolScope {
    cicy_ctx *ctx;
    std::vector<void *> got;
    explicit PoolScope(cicy_ctx *c) : ctx(c) {}
    template <class T>
    bool get(T *&p, size_t count) {
        p = static_cast<T *>(ctx->pool.get(std::max<size_t>(count, 1) * sizeof(T)));
        if (!p) {
            set_error(ctx, "hipMalloc of %zu bytes (set-up) failed", count * sizeof(T));
            return false;
        }
        got.push_back(p);
        return true;
    }
    ~PoolScope() {
        for (void *q : got) ctx->pool.put(q);
    }
};

// dea_core.hip
cicy_core *core_setup_device(cicy_ctx *ctx, int m, const int *frp, const int *fcol,
                             const double *fval, const double *cdiag, const int *orig,
                             int64_t nblocks, const int64_t *ranges);
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
// the same from the reference's arrays in device memory (layout built on the device, no
// host copy kept)
int sys_create_device(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                      const int64_t *col_ind, const double *values, const double *diagonal,
                      cicy_sys **out);
bool sys_alloc_vectors(cicy_sys *s, size_t *total);
// sum of v[i]^2 over n device entries, in a fixed order
bool sum_squares(cicy_ctx *ctx, const double *v, size_t n, double *out);
// builds what `method` needs (idempotent); false on a device or allocation error, or when
// the method cannot run on this system (the reason in the context's error)
// core = false stops after the condensation itself (the core solver's layout is left for
// the next call, e.g. after cicy_sys_set_blocks has chosen its row groups)
bool sys_prepare(cicy_sys *s, int method, bool core = true);
// the core system of a condensed system as an ordinary cicy_sys prepared for `inner`
// (CICY_METHOD_PCG_F64 or CICY_METHOD_RESIDENT); built on first use
bool condensed_core_sys(cicy_sys *s, int inner);

}  // namespace cicy
