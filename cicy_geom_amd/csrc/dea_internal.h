// dea_internal.h -- host-side structures shared by the solver translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cicy_dea.h"
#include "dea_kernels.h"

namespace cicy {

// iterations per captured hipGraph chunk (even: the p double buffer returns to parity 0)
constexpr int kChunk = 32;
// mixed precision: reliable-update trigger r^.r^ <= kDelta^2 * (r^.r^ at the last update),
// and the number of updates in a row without halving ||b - A x|| that ends a solve
constexpr double kDelta = 0.1;
constexpr int kStallLimit = 8;

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

}  // namespace cicy

struct cicy_ctx {
    int device = 0;
    int rows_per_block = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    hipEvent_t ev_poll[2] = {nullptr, nullptr};
    cicy::PcgState *host_state = nullptr;  // pinned, 2 slots
    char err[512] = {0};
};

struct cicy_sys {
    cicy_ctx *ctx = nullptr;
    int n = 0, nnz = 0, rb = 256, nblocks = 0, nremote = 0;
    // blocked SSS, FP64
    int *rp = nullptr, *col = nullptr, *urp = nullptr, *ucol = nullptr;
    double *val = nullptr, *diag = nullptr, *uval = nullptr;
    // FP64 vectors
    double *x = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;
    double *x0 = nullptr, *r0 = nullptr, *b = nullptr;
    double *partials = nullptr, *scal = nullptr;
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
};

namespace cicy {

void set_error(cicy_ctx *ctx, const char *fmt, ...);
bool hip_ok(cicy_ctx *ctx, hipError_t e, const char *what);
bool build_blocked(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                   const int64_t *col_ind, const double *values, const double *diag, int rb,
                   BlockedHost &h);
bool run_chunks(cicy_sys *s, hipGraphExec_t exec);
bool sys_true_residual(cicy_sys *s, const double *xdev, double *rel);
bool solve_pcg64(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats);

// dea_mixed.hip
bool mixed_setup(cicy_sys *s, const BlockedHost &h);
void mixed_free(cicy_sys *s);
bool solve_mixed(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats);
void mixed_spmv32(cicy_sys *s);
void scaled_values(const BlockedHost &h, std::vector<float> &v32, std::vector<float> &u32);

}  // namespace cicy
