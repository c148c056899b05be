// dea_solver.hip -- MI355X (gfx950) Direct Electrifying solve: blocked-SSS SpMV, fused
// Jacobi-PCG, boundary current, and the C-ABI of include/cicy_dea.h.
//
// Reference path replaced: Direct_Electrifying::Solve_DEA_equations_CG_SSS and helpers
// (Direct_Electrifying.cpp:1502-1735) and the boundary current sum of
// Electrical_analysis.cpp:547-605.  Kernel layout and iteration structure are described
// in dea_kernels.h.
#include "dea_internal.h"

#include <mutex>
#include <thread>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace cicy;

// ------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------

// out = sum over rows of (a[i] - b[i])^2 and b^2 (true-residual norms); single pass,
// grid-stride, reduced with the same fence-free accumulation as the iteration kernels.
__global__ __launch_bounds__(kThreads) void k_resid_norms(int n, const double *__restrict__ bvec,
                                                          const double *__restrict__ ax,
                                                          double *__restrict__ out2,
                                                          unsigned int *__restrict__ counter,
                                                          RedBuf rbuf) {
    __shared__ double red[8];
    __shared__ int flag;
    double v[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const double d = bvec[i] - ax[i];
        v[0] += d * d;
        v[1] += bvec[i] * bvec[i];
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
__global__ __launch_bounds__(kThreads) void k_b_minus(int n, double *__restrict__ r,
                                                     const double *__restrict__ b,
                                                     const double *__restrict__ ax) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) r[i] = b[i] - ax[i];
}

// r <- r - ax (Update_residual_vector's final loop, Direct_Electrifying.cpp:1530-1532)
__global__ __launch_bounds__(kThreads) void k_sub(int n, double *__restrict__ r,
                                                  const double *__restrict__ ax) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) r[i] = r[i] - ax[i];
}

// Plain copy (16-byte vectors, grid-stride): the known-size reference for traffic counters
// and the achievable-bandwidth yardstick.
__global__ __launch_bounds__(kThreads) void k_copy(size_t n16, const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst) {
    // four 16-byte loads per thread in flight before their stores
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += 4 * stride) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i + u * stride < n16) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (i + u * stride < n16) dst[i + u * stride] = v[u];
    }
}

// Read-mostly stream: each thread sums its 16-byte vectors (four loads in flight), one word
// per work-group is written.  The yardstick of a kernel that mostly reads (an SpMV).
__global__ __launch_bounds__(kThreads) void k_read(size_t n16, const uint4 *__restrict__ src,
                                                   unsigned int *__restrict__ out) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    unsigned int acc = 0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n16; i += 4 * stride) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            v[u] = i + u * stride < n16 ? src[i + u * stride] : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int u = 0; u < 4; u++) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
    }
    __shared__ unsigned int sacc[kThreads];
    sacc[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int a = 0;
        for (int t = 0; t < kThreads; t++) a ^= sacc[t];
        out[blockIdx.x] = a;
    }
}

// Sum of squares, per work-group in a fixed order (grid-stride rows, then a tree over the
// work-group's lanes); the work-group sums are added on the host in work-group order.
constexpr int kSqGroups = 256;
__global__ __launch_bounds__(kThreads) void k_sum_squares(size_t n, const double *__restrict__ v,
                                                          double *__restrict__ out) {
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads)
        acc += v[i] * v[i];
    __shared__ double red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// FP64 boundary current I = sum_e (v[ni[e]] - v[nb[e]]) / re[e]: grid-stride over the
// elements, per-work-group sums reduced in a fixed order (reduce_sys: bit-identical from run
// to run), the total written to out[0].  An element whose node lies outside [0, nv) adds
// nothing and sets *bad.  The reduction words (RedBuf slots and counters, the counters at
// zero) are laid out by bc_red().
constexpr int kBcGroups = 1024;  // most work-groups of one launch
// reduction words of a launch: work-group sums, group sums, then kBcGroups + 2 counters
// (group arrivals, the final arrival, the bad-node flag) packed as 32-bit words
constexpr size_t kBcRedWords = 2 * (size_t)kBcGroups + ((size_t)kBcGroups + 2 + 1) / 2;
// largest element or node count of the allocation-free entry point
constexpr int64_t kBcMaxElements = (1LL << 40);

__global__ __launch_bounds__(kThreads) void k_boundary_current(
    long long ne, const long long *__restrict__ nb, const long long *__restrict__ ni,
    const double *__restrict__ re, const double *__restrict__ v, long long nv,
    double *__restrict__ out, RedBuf rb, unsigned int *__restrict__ top,
    unsigned int *__restrict__ bad) {
    __shared__ double red[8];
    __shared__ int flag;
    double s[1] = {0.0};
    for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < ne;
         e += (long long)gridDim.x * kThreads) {
        const long long a = ni[e], b = nb[e];
        if ((unsigned long long)a >= (unsigned long long)nv ||
            (unsigned long long)b >= (unsigned long long)nv) {
            atomicOr(bad, 1u);
            continue;
        }
        s[0] += (v[a] - v[b]) / re[e];
    }
    Blk k;
    k.b = blockIdx.x;
    k.sys = 0;
    k.r0 = 0;
    k.nrows = 0;
    k.first = 0;
    k.count = gridDim.x;
    double o[1];
    if (reduce_sys<1>(s, k, rb, top, red, &flag, o) && threadIdx.x == 0) out[0] = o[0];
}

// p[0 .. n) = 0 (clears counters with a kernel, so the sequence may be captured in a graph)
__global__ __launch_bounds__(kThreads) void k_zero_u32(unsigned int *__restrict__ p, int n) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) p[i] = 0u;
}

namespace {

struct BcRed {
    RedBuf rb;
    unsigned int *top, *bad;
};

// the reduction words of one launch inside `words` (kBcRedWords doubles)
BcRed bc_red(double *words) {
    unsigned int *cnt = reinterpret_cast<unsigned int *>(words + 2 * kBcGroups);
    return BcRed{RedBuf{words + kBcGroups, words, cnt}, cnt + kBcGroups, cnt + kBcGroups + 1};
}

int bc_grid(long long ne) {
    return (int)std::max<long long>(1, std::min<long long>((ne + kThreads - 1) / kThreads,
                                                           kBcGroups));
}

// clears the counters, launches the sum into out on the context's stream
void bc_launch(hipStream_t st, long long ne, const long long *nb, const long long *ni,
               const double *re, const double *v, long long nv, double *out, double *words) {
    const BcRed r = bc_red(words);
    hipLaunchKernelGGL(k_zero_u32, dim3((kBcGroups + 2 + kThreads - 1) / kThreads),
                       dim3(kThreads), 0, st, r.rb.gcount, kBcGroups + 2);
    hipLaunchKernelGGL(k_boundary_current, dim3(bc_grid(ne)), dim3(kThreads), 0, st, ne, nb, ni,
                       re, v, nv, out, r.rb, r.top, r.bad);
}

}  // namespace

// The reduction words of cicy_boundary_current_f64_noalloc: a file-scope device array, one
// copy on every GPU the library runs on, held until the process ends (kBcRedWords doubles,
// about 20 KB).  Host code never uses this name as an address: the device address of each
// GPU's copy is looked up once (scratch_address) and kept in g_scratch.
static __device__ __align__(256) double scratch[kBcRedWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// With `device` current: the device address of its copy of `scratch`, looked up once.
double *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[device] = static_cast<double *>(p);
    }
    return g_scratch[device];
}

}  // namespace

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------

namespace cicy {

void set_error(cicy_ctx *ctx, const char *fmt, ...) {
    if (!ctx) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
}

bool hip_ok(cicy_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    set_error(ctx, "%s failed: %s", what, hipGetErrorString(e));
    return false;
}

// Small read-backs go through the context's pinned buffer: a copy into pageable memory is
// staged by the runtime and waits longer.
constexpr size_t kReadbackBytes = 4 << 20;
bool d2h_sync(cicy_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (ctx->readback && bytes <= kReadbackBytes) {
        if (!hip_ok(ctx, hipMemcpyAsync(ctx->readback, src, bytes, hipMemcpyDeviceToHost,
                                        ctx->stream), "D2H") ||
            !hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync"))
            return false;
        memcpy(dst, ctx->readback, bytes);
        ctx->upload_used = 0;  // every staged upload on the stream has completed
        return true;
    }
    if (!hip_ok(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream),
                "D2H") ||
        !hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync"))
        return false;
    ctx->upload_used = 0;
    return true;
}

// Uploads staged in pinned memory are asynchronous (a pageable source makes the runtime
// stage it itself, which waits longer); the staging space is reused after the next wait on
// the context's stream (d2h_sync), and an upload that does not fit goes directly.
constexpr size_t kUploadBytes = 1 << 20;
bool h2d_async(cicy_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return true;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (st == ctx->stream && ctx->upload && ctx->upload_used + need <= kUploadBytes) {
        char *stage = ctx->upload + ctx->upload_used;
        memcpy(stage, src, bytes);
        ctx->upload_used += need;
        return hip_ok(ctx, hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, st), "H2D");
    }
    return hip_ok(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st), "H2D");
}

// Validates the reference's SSS arrays and builds the blocked layout.
bool build_blocked(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                   const int64_t *col_ind, const double *values, const double *diag, int rb,
                   BlockedHost &h) {
    if (n < 0 || nnz < 0) {
        set_error(ctx, "negative size (n=%lld nnz=%lld)", (long long)n, (long long)nnz);
        return false;
    }
    if (n >= (1LL << 31) - 1 || nnz >= (1LL << 31) - 1) {
        set_error(ctx, "system too large for 32-bit device indices (n=%lld nnz=%lld)",
                  (long long)n, (long long)nnz);
        return false;
    }
    if (n > 0 && (!row_ptr || !diag || (nnz > 0 && (!col_ind || !values)))) {
        set_error(ctx, "null array");
        return false;
    }
    if (n > 0 && (row_ptr[0] != 0 || row_ptr[n] != nnz)) {
        set_error(ctx, "row_ptr must start at 0 and end at nnz (row_ptr[n]=%lld, nnz=%lld)",
                  (long long)row_ptr[n], (long long)nnz);
        return false;
    }
    h.n = (int)n;
    h.nnz = (int)nnz;
    h.rb = rb;
    h.nblocks = n > 0 ? (int)((n + rb - 1) / rb) : 0;
    h.rp.assign((size_t)n + 1, 0);
    h.col.resize((size_t)nnz);
    // rows are checked and copied on host threads, each over a range of rows; the entries
    // whose column lies in another row block (the transposed "remote" part) are collected per
    // thread in row order and then placed by one stable counting pass, so the layout does not
    // depend on the thread count
    struct Remote {
        int c, r;
        double v;
    };
    const int nth = (int)std::max<int64_t>(
        1, std::min<int64_t>({16, (int64_t)std::max(1u, std::thread::hardware_concurrency()),
                              n / 65536 + 1}));
    std::vector<std::vector<Remote>> rem((size_t)nth);
    // first error of each thread: row, kind (1 row_ptr decreases, 2 diagonal, 3 column), entry
    std::vector<int64_t> err_row((size_t)nth, -1), err_j((size_t)nth, 0);
    std::vector<int> err_kind((size_t)nth, 0);
    auto work = [&](int t) {
        const int64_t r0 = n * t / nth, r1 = n * (t + 1) / nth;
        for (int64_t r = r0; r < r1; r++) {
            // every bound a row's entries are read with is checked by the thread that reads
            // them: row_ptr[r0] is also the end of the previous thread's last row
            if (row_ptr[r] < 0 || row_ptr[r + 1] < row_ptr[r] || row_ptr[r + 1] > nnz) {
                err_row[(size_t)t] = r;
                err_kind[(size_t)t] = 1;
                return;
            }
            if (!(diag[r] != 0.0) || !std::isfinite(diag[r])) {
                err_row[(size_t)t] = r;
                err_kind[(size_t)t] = 2;
                return;
            }
            h.rp[(size_t)r + 1] = (int)row_ptr[r + 1];
            for (int64_t j = row_ptr[r]; j < row_ptr[r + 1]; j++) {
                const int64_t c = col_ind[j];
                if (c < 0 || c >= r) {
                    err_row[(size_t)t] = r;
                    err_kind[(size_t)t] = 3;
                    err_j[(size_t)t] = j;
                    return;
                }
                h.col[(size_t)j] = (int)c;
                if (c / rb != r / rb) rem[(size_t)t].push_back({(int)c, (int)r, values[j]});
            }
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 1; t < nth; t++) th.emplace_back(work, t);
        work(0);
        for (auto &x : th) x.join();
    }
    for (int t = 0; t < nth; t++) {
        const int64_t r = err_row[(size_t)t];
        if (r < 0) continue;
        if (err_kind[(size_t)t] == 1)
            set_error(ctx, "row_ptr decreases or exceeds nnz at row %lld", (long long)r);
        else if (err_kind[(size_t)t] == 2)
            set_error(ctx, "zero or non-finite diagonal at row %lld (floating node)",
                      (long long)r);
        else
            set_error(ctx, "col_ind[%lld]=%lld is not strictly below row %lld",
                      (long long)err_j[(size_t)t], (long long)col_ind[err_j[(size_t)t]],
                      (long long)r);
        return false;
    }
    h.val.assign(values, values + nnz);
    h.diag.assign(diag, diag + n);
    h.urp.assign((size_t)n + 1, 0);
    for (const auto &v : rem)
        for (const Remote &e : v) h.urp[(size_t)e.c + 1]++;
    for (int64_t i = 0; i < n; i++) h.urp[(size_t)i + 1] += h.urp[(size_t)i];
    h.nremote = n > 0 ? h.urp[(size_t)n] : 0;
    h.ucol.resize((size_t)h.nremote);
    h.uval.resize((size_t)h.nremote);
    std::vector<int> fill(h.urp.begin(), h.urp.end() - (n > 0 ? 1 : 0));
    for (const auto &v : rem)
        for (const Remote &e : v) {
            const int at = fill[(size_t)e.c]++;
            h.ucol[(size_t)at] = e.r;
            h.uval[(size_t)at] = e.v;
        }
    return true;
}

double band_fraction(const BlockedHost &h) {
    if (h.nnz == 0) return 1.0;
    long long on = 0;
    for (int r = 0; r < h.n; r++)
        for (int j = h.rp[r]; j < h.rp[r + 1]; j++)
            if (h.col[j] == r - 1) on++;
    return (double)on / (double)h.nnz;
}

void build_band(const BlockedHost &h, BandHost &b) {
    b.band.assign((size_t)h.n + 1, 0.0);
    // the entries off the sub-diagonal, gathered on host threads in row order; the band is
    // written by the thread that owns the row
    struct Off {
        int r, c;
        double v;
    };
    const int nth = host_threads(h.n, 1 << 17);
    std::vector<std::vector<Off>> offs((size_t)nth);
    parallel_for(nth, [&](int t) {
        const int r0 = (int)((long long)h.n * t / nth), r1 = (int)((long long)h.n * (t + 1) / nth);
        for (int r = r0; r < r1; r++)
            for (int j = h.rp[r]; j < h.rp[r + 1]; j++) {
                const int c = h.col[j];
                if (c == r - 1) b.band[r] = h.val[j];
                else offs[(size_t)t].push_back({r, c, h.val[j]});
            }
    });
    std::vector<int> cnt((size_t)h.n + 1, 0);
    for (const auto &v : offs)
        for (const Off &o : v) {
            cnt[(size_t)o.r + 1]++;
            cnt[(size_t)o.c + 1]++;
        }
    for (int i = 0; i < h.n; i++) cnt[i + 1] += cnt[i];
    b.m = h.n > 0 ? cnt[h.n] : 0;
    b.xrp = cnt;
    b.xrow.resize((size_t)b.m);
    b.xnbr.resize((size_t)b.m);
    b.xval.resize((size_t)b.m);
    // each entry becomes a half-entry of its row and one of its column's row, placed in the
    // sequential row order (the layout does not depend on the thread count)
    std::vector<int> fill(cnt.begin(), cnt.end() - (h.n > 0 ? 1 : 0));
    for (const auto &v : offs)
        for (const Off &o : v) {
            int a = fill[o.r]++;
            b.xrow[a] = o.r;
            b.xnbr[a] = o.c;
            b.xval[a] = o.v;
            a = fill[o.c]++;
            b.xrow[a] = o.c;
            b.xnbr[a] = o.r;
            b.xval[a] = o.v;
        }
    b.xblk.assign((size_t)h.nblocks + 1, 0);
    for (int k = 0; k <= h.nblocks; k++) b.xblk[k] = cnt[std::min(h.n, k * h.rb)];
    // launch order: blocks with more than four times the mean number of half-entries first
    // (heaviest first), then the others in row order
    b.order.resize((size_t)h.nblocks);
    const double mean = h.nblocks > 0 ? (double)b.m / h.nblocks : 0.0;
    std::vector<std::pair<int, int>> heavy;
    int at = 0;
    for (int k = 0; k < h.nblocks; k++) {
        const int c = b.xblk[k + 1] - b.xblk[k];
        if (c > 4.0 * mean && c > 0) heavy.push_back({-c, k});
    }
    std::sort(heavy.begin(), heavy.end());
    std::vector<char> first((size_t)h.nblocks, 0);
    for (auto &e : heavy) {
        b.order[at++] = e.second;
        first[e.second] = 1;
    }
    for (int k = 0; k < h.nblocks; k++)
        if (!first[k]) b.order[at++] = k;
}

// Jacobi-scaled FP32 values of the band layout (a * s_hi * s_lo, s = 1/sqrt(d))
void scaled_band(const BlockedHost &h, const BandHost &b, std::vector<float> &band32,
                 std::vector<float> &xval32) {
    std::vector<double> sc((size_t)h.n);
    for (int i = 0; i < h.n; i++) sc[i] = 1.0 / std::sqrt(h.diag[i]);
    band32.assign((size_t)h.n + 1, 0.f);
    for (int r = 1; r < h.n; r++) band32[r] = (float)(b.band[r] * sc[r] * sc[r - 1]);
    xval32.resize((size_t)b.m);
    for (int k = 0; k < b.m; k++) {
        const int hi = std::max(b.xrow[k], b.xnbr[k]), lo = std::min(b.xrow[k], b.xnbr[k]);
        xval32[k] = (float)(b.xval[k] * sc[hi] * sc[lo]);
    }
}

// The band + remainder layout is the default for every matrix: its products sum each row
// in a fixed order (bit-identical from run to run).  The blocked SSS layout, whose in-block
// transposes are added with LDS atomics in arrival order, is used only when asked for.
int choose_format(const cicy_ctx *ctx, const BlockedHost &h) {
    (void)h;
    if (ctx->format == kFormatSss || ctx->format == kFormatBand) return ctx->format;
    return kFormatBand;
}

DevMat<double> mat64(const cicy_sys *s) {
    DevMat<double> m;
    m.format = s->format;
    m.sss = SssDev<double>{s->n, s->nblocks, s->rp, s->col, s->val, s->diag, s->urp, s->ucol,
                           s->uval};
    m.band = BandDev<double>{s->n, s->nblocks, s->band, s->diag, s->xblk, s->xrp, s->xnbr,
                             s->xval, s->xorder};
    return m;
}

DevMat<float> mat32(const cicy_sys *s) {
    DevMat<float> m;
    m.format = s->format;
    m.sss = SssDev<float>{s->n, s->nblocks, s->rp, s->col, s->val32, nullptr, s->urp, s->ucol,
                          s->uval32};
    m.band = BandDev<float>{s->n, s->nblocks, s->band32, nullptr, s->xblk, s->xrp, s->xnbr,
                            s->xval32, s->xorder};
    return m;
}

}  // namespace cicy

namespace {

template <typename T>
bool dmalloc(cicy_ctx *ctx, T **p, size_t count, size_t *total) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    *p = static_cast<T *>(ctx->pool.get(bytes));
    if (!*p) {
        set_error(ctx, "hipMalloc of %zu bytes failed", bytes);
        return false;
    }
    if (total) *total += bytes;
    return true;
}

template <typename T>
bool h2d(cicy_ctx *ctx, T *dst, const T *src, size_t count) {
    if (count == 0) return true;
    return hip_ok(ctx, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice,
                                      ctx->stream),
                  "hipMemcpyAsync H2D");
}

template <typename T>
bool d2h(cicy_ctx *ctx, T *dst, const T *src, size_t count) {
    if (count == 0) return true;
    return hip_ok(ctx, hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost,
                                      ctx->stream),
                  "hipMemcpyAsync D2H");
}

// Dispatch on the rows-per-block template parameter.
template <template <int> class F, typename... A>
bool dispatch_rb(int rb, A &&...args) {
    switch (rb) {
        case 128: F<128>::run(args...); return true;
        case 256: F<256>::run(args...); return true;
        case 512: F<512>::run(args...); return true;
        case 1024: F<1024>::run(args...); return true;
        default: return false;
    }
}

constexpr BatchDev kNoBatch = {nullptr, nullptr, nullptr, nullptr, nullptr};

template <int RB>
struct LaunchSpmv64 {
    static void run(cicy_sys *s, const double *v, double *out) {
        launch_apply<double, RB, false, false, false>(mat64(s), s->nblocks, s->ctx->stream,
                                                      kNoBatch, v, nullptr, nullptr, out,
                                                      nullptr, RedBuf{}, nullptr);
    }
};

template <int RB>
struct LaunchInit64 {
    static void run(cicy_sys *s, double tol, double stop_ref2, long long max_iter) {
        hipLaunchKernelGGL((k_init_f64<RB, false>), dim3(s->nblocks), dim3(kThreads), 0,
                           s->ctx->stream, s->n, kNoBatch, s->diag, s->r, s->p0, s->p1, s->st,
                           redbuf(s), tol, nullptr, stop_ref2, max_iter);
    }
};

template <int RB>
struct LaunchIter64 {
    // one PCG iteration; `parity` selects which p buffer is old/new
    static void run(cicy_sys *s, int parity) {
        double *pold = parity ? s->p1 : s->p0;
        double *pnew = parity ? s->p0 : s->p1;
        launch_apply<double, RB, true, false, false>(mat64(s), s->nblocks, s->ctx->stream,
                                                     kNoBatch, pold, s->r, pnew, s->q, s->st,
                                                     redbuf(s), nullptr);
        hipLaunchKernelGGL((k_update_f64<RB, false>), dim3(s->nblocks), dim3(kThreads), 0,
                           s->ctx->stream, s->n, kNoBatch, s->diag, s->x, s->r, pnew, s->q,
                           s->st, redbuf(s));
    }
};

// one chunk of kChunk iterations launched directly (no graph)
void chunk64_direct(cicy_sys *s) {
    for (int it = 0; it < kChunk; it++) dispatch_rb<LaunchIter64>(s->rb, s, it & 1);
}

bool ensure_graph64(cicy_sys *s) {
    if (s->graph_exec64 || s->ctx->no_graphs) return true;
    cicy_ctx *ctx = s->ctx;
    hipGraph_t g;
    if (!hip_ok(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture"))
        return false;
    for (int it = 0; it < kChunk; it++) dispatch_rb<LaunchIter64>(s->rb, s, it & 1);
    if (!hip_ok(ctx, hipStreamEndCapture(ctx->stream, &g), "hipStreamEndCapture")) return false;
    hipError_t e = hipGraphInstantiate(&s->graph_exec64, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return hip_ok(ctx, e, "hipGraphInstantiate");
}

}  // namespace

namespace cicy {

// Runs chunks of kChunk iterations until the device flag says done.  The host polls the
// state copied after each chunk while the next chunk is already queued.
bool run_chunks(cicy_sys *s, hipGraphExec_t exec, void (*direct)(cicy_sys *)) {
    cicy_ctx *ctx = s->ctx;
    const bool use_graph = !(ctx->no_graphs && direct);
    auto launch = [&]() -> bool {
        if (use_graph) return hip_ok(ctx, hipGraphLaunch(exec, ctx->stream), "hipGraphLaunch");
        direct(s);
        return hip_ok(ctx, hipGetLastError(), "chunk launch");
    };
    PcgState *hst = ctx->host_state;  // pinned, two slots
    const int depth = 2;
    hipEvent_t ev[2] = {ctx->ev_poll[0], ctx->ev_poll[1]};
    int slot = 0;
    for (int i = 0; i < depth; i++) {
        if (!launch()) return false;
        if (!hip_ok(ctx, hipMemcpyAsync(&hst[i], s->st, sizeof(PcgState), hipMemcpyDeviceToHost,
                                        ctx->stream), "poll copy"))
            return false;
        if (!hip_ok(ctx, hipEventRecord(ev[i], ctx->stream), "hipEventRecord")) return false;
    }
    // host-side cap on the wall time of one solve; every kernel is finite, so stopping
    // the launches ends the solve (status stays CICY_MAXITER)
    double max_s = 600.0;
    if (const char *e = getenv("CICY_MAX_SOLVE_SECONDS")) max_s = atof(e);
    const auto t_start = std::chrono::steady_clock::now();
    for (;;) {
        if (!hip_ok(ctx, hipEventSynchronize(ev[slot]), "hipEventSynchronize")) return false;
        if (hst[slot].done) break;
        if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() >
            max_s) {
            set_error(ctx, "solve stopped after %.0f s (CICY_MAX_SOLVE_SECONDS)", max_s);
            break;
        }
        if (!launch()) return false;
        if (!hip_ok(ctx, hipMemcpyAsync(&hst[slot], s->st, sizeof(PcgState),
                                        hipMemcpyDeviceToHost, ctx->stream), "poll copy"))
            return false;
        if (!hip_ok(ctx, hipEventRecord(ev[slot], ctx->stream), "hipEventRecord")) return false;
        slot ^= 1;
    }
    return hip_ok(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

void sys_spmv64(cicy_sys *s, const double *xdev, double *ydev) {
    dispatch_rb<LaunchSpmv64>(s->rb, s, xdev, ydev);
}

bool sum_squares(cicy_ctx *ctx, const double *v, size_t n, double *out) {
    double *part = nullptr;
    double h[kSqGroups];
    *out = 0.0;
    if (n == 0) return true;
    PoolScope ps(ctx);
    if (!ps.get(part, kSqGroups)) return false;
    hipLaunchKernelGGL(k_sum_squares, dim3(kSqGroups), dim3(kThreads), 0, ctx->stream, n, v, part);
    if (!hip_ok(ctx, hipGetLastError(), "k_sum_squares") ||
        !d2h_sync(ctx, h, part, sizeof(h)))
        return false;
    double t = 0.0;
    for (int g = 0; g < kSqGroups; g++) t += h[g];
    *out = t;
    return true;
}

// sums of squares of two vectors of n entries (v2 may be null: *out2 = 0), one read-back
bool sum_squares2(cicy_ctx *ctx, const double *v1, const double *v2, size_t n, double *out1,
                  double *out2) {
    double *part = nullptr;
    double h[2 * kSqGroups];
    *out1 = *out2 = 0.0;
    if (n == 0) return true;
    PoolScope ps(ctx);
    if (!ps.get(part, 2 * kSqGroups)) return false;
    hipLaunchKernelGGL(k_sum_squares, dim3(kSqGroups), dim3(kThreads), 0, ctx->stream, n, v1, part);
    if (v2)
        hipLaunchKernelGGL(k_sum_squares, dim3(kSqGroups), dim3(kThreads), 0, ctx->stream, n, v2,
                           part + kSqGroups);
    if (!hip_ok(ctx, hipGetLastError(), "k_sum_squares") ||
        !d2h_sync(ctx, h, part, (v2 ? 2 : 1) * kSqGroups * sizeof(double)))
        return false;
    double t1 = 0.0, t2 = 0.0;
    for (int g = 0; g < kSqGroups; g++) t1 += h[g];
    if (v2)
        for (int g = 0; g < kSqGroups; g++) t2 += h[kSqGroups + g];
    *out1 = t1;
    *out2 = t2;
    return true;
}

bool sys_true_residual(cicy_sys *s, const double *xdev, double *rel, double *b2) {
    cicy_ctx *ctx = s->ctx;
    if (!s->has_b) {
        *rel = -1.0;
        return true;
    }
    dispatch_rb<LaunchSpmv64>(s->rb, s, xdev, s->q);
    int nb = std::min(s->nblocks > 0 ? s->nblocks : 1, 1024);
    hipLaunchKernelGGL(k_resid_norms, dim3(nb), dim3(kThreads), 0, ctx->stream, s->n, s->b, s->q,
                       s->scal, s->counter2, redbuf(s));
    double h[2];
    if (!hip_ok(ctx, hipMemcpyAsync(h, s->scal, 2 * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream), "D2H norms") ||
        !hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync"))
        return false;
    *rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    if (b2) *b2 = h[1];
    return true;
}

// Plain FP64 Jacobi-PCG.  With CICY_REL_B the stop is guaranteed on the true residual:
// after the recursive residual meets tol*||b||, b - A x is recomputed in FP64 and, if it
// does not, the iteration continues from x with r = b - A x (at most 4 times).
// keep_x: continue from the current device x/r instead of the stored x0/r0.
bool solve_pcg64(cicy_sys *s, double tol, long long max_iter, int rel_to,
                 cicy_solve_stats *stats, bool keep_x) {
    cicy_ctx *ctx = s->ctx;
    if (s->n == 0) {
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->status = CICY_OK;
        }
        return true;
    }
    if (rel_to == CICY_REL_B && !s->has_b && s->ref2_override < 0.0) {
        set_error(ctx, "CICY_REL_B needs b (cicy_sys_set_rhs)");
        return false;
    }
    if (!ensure_graph64(s)) return false;
    size_t vb = (size_t)s->n * sizeof(double);
    hipStream_t st = ctx->stream;
    if (!keep_x && !hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "hipEventRecord")) return false;
    const double stop_ref2 = s->ref2_override >= 0.0   ? s->ref2_override
                             : rel_to == CICY_REL_B ? s->bnorm2
                                                    : -1.0;
    if (!keep_x &&
        (!hip_ok(ctx, hipMemcpyAsync(s->r, s->r0, vb, hipMemcpyDeviceToDevice, st), "copy r0") ||
         !hip_ok(ctx, hipMemcpyAsync(s->x, s->x0, vb, hipMemcpyDeviceToDevice, st), "copy x0")))
        return false;
    long long total_iters = 0;
    int restarts = 0;
    double rr_init = -1.0;
    PcgState h;
    double tr = -1.0;
    for (;;) {
        if (!hip_ok(ctx, hipMemsetAsync(s->st, 0, sizeof(PcgState), st), "memset state"))
            return false;
        dispatch_rb<LaunchInit64>(s->rb, s, tol, stop_ref2, max_iter - total_iters);
        if (!run_chunks(s, s->graph_exec64, chunk64_direct)) return false;
        if (!d2h_sync(ctx, &h, s->st, sizeof(PcgState)))
            return false;
        if (rr_init < 0.0) rr_init = h.rr_init;
        total_iters += h.iters;
        tr = -1.0;
        if (s->has_b && !sys_true_residual(s, s->x, &tr)) return false;
        if (rel_to != CICY_REL_B || h.status != CICY_OK || tr <= tol || restarts >= 4 ||
            total_iters >= max_iter)
            break;
        // q holds A x from the true-residual check: continue from r = b - A x
        hipLaunchKernelGGL(k_b_minus, dim3((s->n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                           st, s->n, s->r, s->b, s->q);
        restarts++;
    }
    if (!keep_x) {
        if (!hip_ok(ctx, hipEventRecord(ctx->ev_stop, st), "hipEventRecord") ||
            !hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "hipEventSynchronize"))
            return false;
    }
    float ms = 0.f;
    if (!keep_x) (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->iters = total_iters;
        stats->status = h.status;
        stats->refinements = restarts;
        // relative to the norm the stop was taken against
        const double rel2 = s->ref2_override >= 0.0 ? s->ref2_override : rr_init;
        stats->rel_res = rel2 > 0 ? std::sqrt(h.rr / rel2) : 0.0;
        stats->solve_ms = ms;
        stats->true_rel_res = tr;
    }
    return true;
}

}  // namespace cicy

// ------------------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------------------

extern "C" {

int cicy_create(cicy_ctx **out, int device) {
    if (!out) return 0;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return 0;
    cicy_ctx *ctx = new cicy_ctx();
    ctx->device = device;
    ctx->rows_per_block = 256;
    if (const char *e = getenv("CICY_NO_GRAPHS")) ctx->no_graphs = e[0] && e[0] != '0';
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&ctx->ev_start) != hipSuccess || hipEventCreate(&ctx->ev_stop) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_poll[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_poll[1], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&ctx->host_state, 2 * sizeof(PcgState), hipHostMallocDefault) !=
            hipSuccess ||
        hipHostMalloc(&ctx->readback, cicy::kReadbackBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&ctx->upload, cicy::kUploadBytes, hipHostMallocDefault) !=
            hipSuccess ||
        !(ctx->scratch = scratch_address(device))) {
        if (ctx->host_state) (void)hipHostFree(ctx->host_state);
        if (ctx->readback) (void)hipHostFree(ctx->readback);
        if (ctx->upload) (void)hipHostFree(ctx->upload);
        delete ctx;
        return 0;
    }
    *out = ctx;
    return 1;
}

void cicy_destroy(cicy_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->host_state) (void)hipHostFree(ctx->host_state);
    if (ctx->readback) (void)hipHostFree(ctx->readback);
    if (ctx->upload) (void)hipHostFree(ctx->upload);
    cicy::dev_release(ctx, ctx->bc_ws);
    ctx->pool.clear();
    (void)hipEventDestroy(ctx->ev_start);
    (void)hipEventDestroy(ctx->ev_stop);
    (void)hipEventDestroy(ctx->ev_poll[0]);
    (void)hipEventDestroy(ctx->ev_poll[1]);
    if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *cicy_last_error(const cicy_ctx *ctx) { return ctx ? ctx->err : "null context"; }

int cicy_set_rows_per_block(cicy_ctx *ctx, int rb) {
    if (!ctx) return 0;
    if (rb == 0) rb = 256;
    if (rb != 128 && rb != 256 && rb != 512 && rb != 1024) {
        set_error(ctx, "rows_per_block must be 128, 256, 512 or 1024");
        return 0;
    }
    ctx->rows_per_block = rb;
    return 1;
}

int cicy_set_dropin_method(cicy_ctx *ctx, int method) {
    if (!ctx) return 0;
    if (method != CICY_METHOD_PCG_F64 && method != CICY_METHOD_RESIDENT &&
        method != CICY_METHOD_CONDENSED && method != CICY_METHOD_CONDENSED_RESIDENT &&
        method != CICY_METHOD_CONDENSED_BLOCK) {
        cicy::set_error(ctx, "drop-in method must be PCG_F64, RESIDENT, CONDENSED, "
                             "CONDENSED_RESIDENT or CONDENSED_BLOCK");
        return 0;
    }
    ctx->dropin_method = method;
    return 1;
}

int cicy_set_dropin_blocks(cicy_ctx *ctx, int64_t nblocks, const int64_t *ranges) {
    if (!ctx) return 0;
    if (nblocks < 0 || (nblocks > 0 && !ranges)) {
        cicy::set_error(ctx, "cicy_set_dropin_blocks: invalid block list");
        return 0;
    }
    ctx->dropin_blocks.assign(ranges, ranges + 2 * nblocks);
    return 1;
}

int cicy_set_format(cicy_ctx *ctx, int format) {
    if (!ctx) return 0;
    if (format != 0 && format != kFormatSss && format != kFormatBand) {
        set_error(ctx, "format must be 0 (auto), 1 (blocked SSS) or 2 (band + remainder)");
        return 0;
    }
    ctx->format = format;
    return 1;
}

}  // extern "C"

namespace cicy {

// Uploads a system in the layout of the two-kernel FP64 solve and keeps a host copy for
// the layouts of the other methods, which sys_prepare builds when they are first needed.
int sys_create(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
               const int64_t *col_ind, const double *values, const double *diagonal,
               cicy_sys **out) {
    if (!ctx || !out) return 0;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return 0;
    StageClock clk;
    SysHost *keep = new SysHost();
    BlockedHost &h = keep->h;
    if (!build_blocked(ctx, n, nnz, row_ptr, col_ind, values, diagonal, ctx->rows_per_block, h)) {
        delete keep;
        return 0;
    }
    clk.mark("create: row checks + blocks");
    cicy_sys *s = new cicy_sys();
    s->host = keep;
    s->resident_rows_req = ctx->resident_rows;
    s->ctx = ctx;
    s->n = h.n;
    s->nnz = h.nnz;
    s->rb = h.rb;
    s->nblocks = h.nblocks;
    s->nremote = h.nremote;
    size_t tot = 0;
    const size_t nv = (size_t)h.n;
    s->format = choose_format(ctx, h);
    BandHost &bh = keep->bh;
    bool ok = dmalloc(ctx, &s->diag, nv, &tot);
    if (s->format == kFormatSss) {
        ok = ok && dmalloc(ctx, &s->rp, nv + 1, &tot) &&
             dmalloc(ctx, &s->col, (size_t)h.nnz, &tot) &&
             dmalloc(ctx, &s->val, (size_t)h.nnz, &tot) && dmalloc(ctx, &s->urp, nv + 1, &tot) &&
             dmalloc(ctx, &s->ucol, (size_t)h.nremote, &tot) &&
             dmalloc(ctx, &s->uval, (size_t)h.nremote, &tot) &&
             h2d(ctx, s->rp, h.rp.data(), nv + 1) && h2d(ctx, s->col, h.col.data(), h.nnz) &&
             h2d(ctx, s->val, h.val.data(), h.nnz) && h2d(ctx, s->urp, h.urp.data(), nv + 1) &&
             h2d(ctx, s->ucol, h.ucol.data(), h.nremote) &&
             h2d(ctx, s->uval, h.uval.data(), h.nremote);
    } else {
        build_band(h, bh);
        keep->have_band = true;
        clk.mark("create: band layout");
        s->nx = bh.m;
        ok = ok && dmalloc(ctx, &s->band, nv + 1, &tot) &&
             dmalloc(ctx, &s->xblk, (size_t)h.nblocks + 1, &tot) &&
             dmalloc(ctx, &s->xorder, std::max<size_t>((size_t)h.nblocks, 1), &tot) &&
             h2d(ctx, s->xorder, bh.order.data(), (size_t)h.nblocks) &&
             dmalloc(ctx, &s->xrp, nv + 1, &tot) && dmalloc(ctx, &s->xnbr, (size_t)bh.m, &tot) &&
             dmalloc(ctx, &s->xval, (size_t)bh.m, &tot) &&
             h2d(ctx, s->band, bh.band.data(), nv + 1) &&
             h2d(ctx, s->xblk, bh.xblk.data(), (size_t)h.nblocks + 1) &&
             h2d(ctx, s->xrp, bh.xrp.data(), nv + 1) &&
             h2d(ctx, s->xnbr, bh.xnbr.data(), (size_t)bh.m) &&
             h2d(ctx, s->xval, bh.xval.data(), (size_t)bh.m);
    }
    ok = ok && h2d(ctx, s->diag, h.diag.data(), nv) && sys_alloc_vectors(s, &tot);
    s->device_bytes = tot;
    clk.mark("create: allocate + upload");
    if (!ok) {
        cicy_sys_destroy(s);
        return 0;
    }
    s->res_why = s->cond_why = "not prepared";
    *out = s;
    return 1;
}

// The vectors and reduction words of a system whose matrix layout is in place.
bool sys_alloc_vectors(cicy_sys *s, size_t *total) {
    cicy_ctx *ctx = s->ctx;
    const size_t nv = (size_t)s->n;
    const size_t nparts = 2 * (size_t)std::max(s->nblocks, 1024);
    size_t tot = 0;
    bool ok = dmalloc(ctx, &s->x, nv, &tot) &&
              dmalloc(ctx, &s->r, nv, &tot) && dmalloc(ctx, &s->p0, nv, &tot) &&
              dmalloc(ctx, &s->p1, nv, &tot) && dmalloc(ctx, &s->q, nv, &tot) &&
              dmalloc(ctx, &s->x0, nv, &tot) && dmalloc(ctx, &s->r0, nv, &tot) &&
              dmalloc(ctx, &s->b, nv, &tot) && dmalloc(ctx, &s->partials, nparts, &tot) &&
              dmalloc(ctx, &s->gpart, nparts, &tot) &&
              dmalloc(ctx, &s->gcount, std::max<size_t>((size_t)s->nblocks, 1), &tot) &&
              dmalloc(ctx, &s->st, 1, &tot) && dmalloc(ctx, &s->scal, 4, &tot) &&
              dmalloc(ctx, &s->counter2, 1, &tot);
    ok = ok &&
         hip_ok(ctx, hipMemsetAsync(s->partials, 0, nparts * sizeof(double), ctx->stream),
                "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->gpart, 0, nparts * sizeof(double), ctx->stream),
                "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->x0, 0, nv * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->p0, 0, nv * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->r0, 0, nv * sizeof(double), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->counter2, 0, sizeof(unsigned), ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->gcount, 0,
                                    std::max<size_t>((size_t)s->nblocks, 1) * sizeof(unsigned),
                                    ctx->stream), "memset") &&
         hip_ok(ctx, hipMemsetAsync(s->st, 0, sizeof(PcgState), ctx->stream), "memset");
    *total += tot;
    return ok;
}

// A system from the reference's arrays resident in device memory: checked and laid out
// (band + remainder) on the device; no host copy is kept, so the methods whose layouts are
// built on the host (mixed, resident, the blocked SSS layout) are not available for it.
int sys_create_device(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                      const int64_t *col_ind, const double *values, const double *diagonal,
                      cicy_sys **out) {
    if (!ctx || !out) return 0;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return 0;
    if (n < 0 || nnz < 0) {
        set_error(ctx, "negative size (n=%lld nnz=%lld)", (long long)n, (long long)nnz);
        return 0;
    }
    if (n >= (1LL << 31) - 1 || 2 * nnz >= (1LL << 31) - 1) {
        set_error(ctx, "system too large for 32-bit device indices (n=%lld nnz=%lld)",
                  (long long)n, (long long)nnz);
        return 0;
    }
    if (n > 0 && (!row_ptr || !diagonal || (nnz > 0 && (!col_ind || !values)))) {
        set_error(ctx, "null array");
        return 0;
    }
    if (ctx->format == kFormatSss) {
        set_error(ctx, "the blocked SSS layout is built from host arrays (cicy_sys_create)");
        return 0;
    }
    StageClock clk;
    cicy_sys *s = new cicy_sys();
    s->ctx = ctx;
    s->resident_rows_req = ctx->resident_rows;
    s->n = (int)n;
    s->nnz = (int)nnz;
    s->rb = ctx->rows_per_block;
    s->nblocks = n > 0 ? (int)((n + s->rb - 1) / s->rb) : 0;
    s->format = kFormatBand;
    s->mixed_done = s->res_done = true;  // no host copy to build them from
    s->res_why = "the register-resident layout is built from host arrays (cicy_sys_create)";
    size_t tot = 0;
    bool ok = dmalloc(ctx, &s->diag, (size_t)n, &tot);
    s->device_bytes = tot;
    ok = ok && band_from_device(s, row_ptr, col_ind, values, diagonal);
    clk.mark("create (device): layout");
    tot = 0;
    ok = ok && sys_alloc_vectors(s, &tot);
    s->device_bytes += tot;
    clk.mark("create (device): vectors");
    if (!ok) {
        cicy_sys_destroy(s);
        return 0;
    }
    s->cond_why = "not prepared";
    *out = s;
    return 1;
}

// The host copy is dropped once every layout that is built from it has been attempted.
static void release_host(cicy_sys *s) {
    if (s->host && s->mixed_done && s->res_done && s->cond_done) {
        delete s->host;
        s->host = nullptr;
    }
}

static BandHost &host_band(SysHost &k) {
    if (!k.have_band) {
        build_band(k.h, k.bh);
        k.have_band = true;
    }
    return k.bh;
}

bool sys_prepare(cicy_sys *s, int method, bool core) {
    cicy_ctx *ctx = s->ctx;
    const bool need_mixed = method == CICY_METHOD_MIXED;
    const bool need_res = method == CICY_METHOD_RESIDENT;
    const bool need_cond = method == CICY_METHOD_CONDENSED ||
                           method == CICY_METHOD_CONDENSED_RESIDENT ||
                           method == CICY_METHOD_CONDENSED_BLOCK;
    if (method < CICY_METHOD_PCG_F64 || method > CICY_METHOD_CONDENSED_BLOCK) {
        set_error(ctx, "unknown method %d", method);
        return false;
    }
    if ((need_mixed || need_res) && !s->host) {
        set_error(ctx, "method %d needs the host arrays of the system (create it with "
                       "cicy_sys_create)", method);
        return false;
    }
    if (need_mixed && !s->mixed_done) {
        s->mixed_done = true;
        // on failure the buffers allocated so far are freed with the system
        if (!mixed_setup(s, s->host->h, host_band(*s->host))) return false;
    }
    if (need_res && !s->res_done) {
        // the resident layout splits the matrix into band + remainder whatever the device
        // format of the two-kernel path (remainder rows become published rows)
        s->res_done = true;
        bool failed = false;
        s->res = resident_setup(ctx, s->host->h, host_band(*s->host), s->resident_rows_req,
                                &s->res_why, &failed);
        if (failed) return false;
    }
    if (need_cond && !s->cond_done) {
        s->cond_done = true;
        bool failed = false;
        if (s->format == kFormatBand && (!ctx->setup_on_host || !s->host))
            s->cond = condensed_setup_device(s, &s->cond_why, &failed);
        else
            s->cond = condensed_setup(ctx, s->host->h, host_band(*s->host), &s->cond_why,
                                      &failed);
        if (failed) return false;
    }
    release_host(s);
    if (need_cond && s->cond && core) {
        if (method == CICY_METHOD_CONDENSED || method == CICY_METHOD_CONDENSED_RESIDENT) {
            const int inner = method == CICY_METHOD_CONDENSED ? CICY_METHOD_PCG_F64
                                                              : CICY_METHOD_RESIDENT;
            if (!condensed_core_sys(s, inner)) return false;
        }
        if (method == CICY_METHOD_CONDENSED_BLOCK && condensed_info(s->cond, 2) > 0 &&
            !condensed_info(s->cond, 9) && !condensed_set_blocks(s, 0, nullptr))
            return false;
    }
    return true;
}

}  // namespace cicy

extern "C" {

int cicy_sys_create(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                    const int64_t *col_ind, const double *values, const double *diagonal,
                    cicy_sys **out) {
    return cicy::sys_create(ctx, n, nnz, row_ptr, col_ind, values, diagonal, out);
}

int cicy_sys_create_device(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *d_row_ptr,
                           const int64_t *d_col_ind, const double *d_values,
                           const double *d_diagonal, cicy_sys **out) {
    return cicy::sys_create_device(ctx, n, nnz, d_row_ptr, d_col_ind, d_values, d_diagonal, out);
}

int cicy_sys_set_rhs_device(cicy_sys *s, const double *d_r0, const double *d_x0,
                            const double *d_b) {
    if (!s || !d_r0) return 0;
    cicy_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    const size_t nv = (size_t)s->n;
    hipStream_t st = ctx->stream;
    bool ok = hip_ok(ctx, hipMemcpyAsync(s->r0, d_r0, nv * sizeof(double),
                                         hipMemcpyDeviceToDevice, st), "copy r0");
    ok = ok && (d_x0 ? hip_ok(ctx, hipMemcpyAsync(s->x0, d_x0, nv * sizeof(double),
                                                  hipMemcpyDeviceToDevice, st), "copy x0")
                     : hip_ok(ctx, hipMemsetAsync(s->x0, 0, nv * sizeof(double), st), "memset"));
    if (d_b)
        ok = ok && hip_ok(ctx, hipMemcpyAsync(s->b, d_b, nv * sizeof(double),
                                              hipMemcpyDeviceToDevice, st), "copy b");
    double r02 = 0.0, b2 = 0.0;
    ok = ok && cicy::sum_squares2(ctx, s->r0, d_b ? s->b : nullptr, nv, &r02, &b2);
    if (!ok) return 0;
    s->r0norm2 = r02;
    s->has_b = d_b != nullptr;
    if (d_b) s->bnorm2 = b2;
    return 1;
}

int cicy_set_setup(cicy_ctx *ctx, int where) {
    if (!ctx || where < 0 || where > 1) return 0;
    ctx->setup_on_host = where == 1;
    return 1;
}

int cicy_device_alloc(cicy_ctx *ctx, int64_t bytes, void **out) {
    if (!ctx || !out || bytes < 0) return 0;
    (void)hipSetDevice(ctx->device);
    *out = nullptr;
    return hip_ok(ctx, hipMalloc(out, std::max<int64_t>(bytes, 1)), "hipMalloc") ? 1 : 0;
}

int cicy_device_free(cicy_ctx *ctx, void *p) {
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    return p == nullptr || hip_ok(ctx, hipFree(p), "hipFree") ? 1 : 0;
}

int cicy_device_upload(cicy_ctx *ctx, void *dst, const void *src, int64_t bytes) {
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return 0;
    (void)hipSetDevice(ctx->device);
    return bytes == 0 || (hip_ok(ctx, hipMemcpyAsync(dst, src, (size_t)bytes,
                                                     hipMemcpyHostToDevice, ctx->stream), "H2D") &&
                          hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync"))
               ? 1 : 0;
}

int cicy_device_download(cicy_ctx *ctx, void *dst, const void *src, int64_t bytes) {
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return 0;
    (void)hipSetDevice(ctx->device);
    return bytes == 0 || d2h_sync(ctx, dst, src, (size_t)bytes) ? 1 : 0;
}

int cicy_sys_prepare(cicy_sys *s, int method) {
    if (!s) return 0;
    (void)hipSetDevice(s->ctx->device);
    return cicy::sys_prepare(s, method) ? 1 : 0;
}

void cicy_sys_destroy(cicy_sys *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    if (s->graph_exec64) (void)hipGraphExecDestroy(s->graph_exec64);
    cicy::mixed_free(s);
    cicy::resident_free(s->res);
    cicy::condensed_free(s->cond);
    delete s->host;
    void *ptrs[] = {s->rp, s->col, s->val, s->diag, s->urp, s->ucol, s->uval, s->band,
                    s->xblk, s->xrp, s->xnbr, s->xval, s->xorder, s->x, s->r, s->p0, s->p1, s->q,
                    s->x0, s->r0, s->b, s->partials, s->gpart, s->gcount, s->st, s->scal,
                    s->counter2};
    for (void *p : ptrs) cicy::dev_release(s->ctx, p);
    delete s;
}

int64_t cicy_sys_size(const cicy_sys *s, int what) {
    if (!s) return -1;
    switch (what) {
        case 0: return s->n;
        case 1: return s->nnz;
        case 2: return s->nremote;
        case 3: return (int64_t)s->device_bytes;
        case 4: return s->format;
        case 5: return s->format == kFormatBand ? s->nx : s->nremote;
        case 6: return cicy::resident_info(s->res, 0);
        case 7: return cicy::resident_info(s->res, 1);
        case 8: return cicy::resident_info(s->res, 2);
        case 9: return cicy::condensed_info(s->cond, 0);
        case 10: return cicy::condensed_info(s->cond, 1);
        case 11: return cicy::condensed_info(s->cond, 2);
        case 12: return cicy::condensed_info(s->cond, 3);
        case 13: return cicy::condensed_info(s->cond, 4);
        case 14: return cicy::condensed_info(s->cond, 5);
        case 15: return cicy::condensed_info(s->cond, 6);
        case 16: return cicy::condensed_info(s->cond, 7);
        case 17: return cicy::condensed_info(s->cond, 8);
        default: return -1;
    }
}

int cicy_sys_set_rhs(cicy_sys *s, const double *r0, const double *x0, const double *b) {
    if (!s || !r0) return 0;
    cicy_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    size_t nv = (size_t)s->n;
    bool ok = h2d(ctx, s->r0, r0, nv);
    if (x0) ok = ok && h2d(ctx, s->x0, x0, nv);
    else ok = ok && hip_ok(ctx, hipMemsetAsync(s->x0, 0, nv * sizeof(double), ctx->stream), "memset");
    double r02 = 0.0;
    for (size_t i = 0; i < nv; i++) r02 += r0[i] * r0[i];
    s->r0norm2 = r02;
    s->has_b = b != nullptr;
    if (b) {
        ok = ok && h2d(ctx, s->b, b, nv);
        double nb2 = 0.0;
        for (size_t i = 0; i < nv; i++) nb2 += b[i] * b[i];
        s->bnorm2 = nb2;
    }
    ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return ok ? 1 : 0;
}

int cicy_sys_set_blocks(cicy_sys *s, int64_t nblocks, const int64_t *ranges) {
    if (!s) return 0;
    (void)hipSetDevice(s->ctx->device);
    if (!cicy::sys_prepare(s, CICY_METHOD_CONDENSED_BLOCK, false)) return 0;
    return cicy::condensed_set_blocks(s, nblocks, ranges);
}

int cicy_sys_solve(cicy_sys *s, int method, double tol, int64_t max_iter, int rel_to,
                   cicy_solve_stats *stats) {
    if (!s) return 0;
    (void)hipSetDevice(s->ctx->device);
    if (max_iter <= 0) max_iter = 300LL * ((long long)s->n + 2);
    if (!(tol >= 0.0)) {
        set_error(s->ctx, "tolerance must be >= 0");
        return 0;
    }
    if (!cicy::sys_prepare(s, method)) {
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->status = CICY_BAD_INPUT;
        }
        return 0;
    }
    bool ok;
    if (method == CICY_METHOD_PCG_F64)
        ok = cicy::solve_pcg64(s, tol, max_iter, rel_to, stats);
    else if (method == CICY_METHOD_MIXED)
        ok = cicy::solve_mixed(s, tol, max_iter, rel_to, stats);
    else if (method == CICY_METHOD_RESIDENT)
        ok = cicy::solve_resident(s, tol, max_iter, rel_to, stats);
    else if (method == CICY_METHOD_CONDENSED)
        ok = cicy::solve_condensed(s, CICY_METHOD_PCG_F64, tol, max_iter, rel_to, stats);
    else if (method == CICY_METHOD_CONDENSED_RESIDENT)
        ok = cicy::solve_condensed(s, CICY_METHOD_RESIDENT, tol, max_iter, rel_to, stats);
    else if (method == CICY_METHOD_CONDENSED_BLOCK)
        ok = cicy::solve_condensed(s, CICY_METHOD_CONDENSED_BLOCK, tol, max_iter, rel_to, stats);
    else {
        set_error(s->ctx, "unknown method %d", method);
        ok = false;
    }
    return ok ? 1 : 0;
}

int cicy_sys_get(cicy_sys *s, double *x, double *r) {
    if (!s) return 0;
    cicy_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    size_t nv = (size_t)s->n;
    bool ok = true;
    if (x) ok = ok && d2h(ctx, x, s->x, nv);
    if (r) ok = ok && d2h(ctx, r, s->r, nv);
    ok = ok && hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    return ok ? 1 : 0;
}

int cicy_sys_spmv_bench(cicy_sys *s, int precision, int reps, double *avg_ms) {
    if (!s || reps <= 0 || !avg_ms) return 0;
    cicy_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    if (s->n == 0) {
        *avg_ms = 0.0;
        return 1;
    }
    if (precision != 64 && precision != 32) {
        set_error(ctx, "precision must be 64 or 32");
        return 0;
    }
    if (precision == 32 && !cicy::sys_prepare(s, CICY_METHOD_MIXED)) return 0;
    // warm the caches once, then time `reps` launches between two events
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && !hip_ok(ctx, hipEventRecord(ctx->ev_start, ctx->stream), "event")) return 0;
        int nrep = pass == 0 ? 1 : reps;
        for (int i = 0; i < nrep; i++) {
            if (precision == 64) dispatch_rb<LaunchSpmv64>(s->rb, s, s->p0, s->q);
            else cicy::mixed_spmv32(s);
        }
    }
    if (!hip_ok(ctx, hipEventRecord(ctx->ev_stop, ctx->stream), "event") ||
        !hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "sync"))
        return 0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    *avg_ms = ms / reps;
    return hip_ok(ctx, hipGetLastError(), "kernel launch") ? 1 : 0;
}

int cicy_sys_spmv_bench_cold(cicy_sys *s, int precision, int reps, int64_t flush_bytes,
                             double *avg_ms) {
    if (!s || reps <= 0 || !avg_ms || flush_bytes < 0) return 0;
    cicy_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    if (s->n == 0) {
        *avg_ms = 0.0;
        return 1;
    }
    if (precision != 64 && precision != 32) {
        set_error(ctx, "precision must be 64 or 32");
        return 0;
    }
    if (precision == 32 && !cicy::sys_prepare(s, CICY_METHOD_MIXED)) return 0;
    // a copy of flush_bytes through the memory-side cache before every timed launch, so the
    // matrix and vectors come from HBM; each launch timed alone between two events
    const size_t n16 = (size_t)flush_bytes / 32;
    uint4 *a = nullptr, *b = nullptr;
    if (n16 && (!hip_ok(ctx, hipMalloc((void **)&a, n16 * 16), "hipMalloc") ||
                !hip_ok(ctx, hipMalloc((void **)&b, n16 * 16), "hipMalloc"))) {
        if (a) (void)hipFree(a);
        return 0;
    }
    bool ok = !n16 || hip_ok(ctx, hipMemsetAsync(a, 0, n16 * 16, ctx->stream), "memset");
    const int grid = (int)std::min<size_t>((n16 + kThreads - 1) / kThreads, 256 * 64);
    double total = 0.0;
    for (int i = 0; ok && i <= reps; i++) {
        if (n16) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kThreads), 0, ctx->stream, n16, a, b);
        ok = hip_ok(ctx, hipEventRecord(ctx->ev_start, ctx->stream), "event");
        if (precision == 64) dispatch_rb<LaunchSpmv64>(s->rb, s, s->p0, s->q);
        else cicy::mixed_spmv32(s);
        ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_stop, ctx->stream), "event") &&
             hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
        if (i > 0) total += ms;  // the first launch warms the code and the TLB
    }
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    *avg_ms = total / reps;
    return ok && hip_ok(ctx, hipGetLastError(), "kernel launch") ? 1 : 0;
}

int cicy_read_bench(cicy_ctx *ctx, int64_t bytes, int reps, double *avg_ms) {
    if (!ctx || bytes < 16 || reps <= 0 || !avg_ms) return 0;
    (void)hipSetDevice(ctx->device);
    const size_t n16 = (size_t)bytes / 16;
    uint4 *a = nullptr;
    unsigned int *o = nullptr;
    const int grid = (int)std::min<size_t>((n16 + kThreads - 1) / kThreads, 256 * 64);
    if (!hip_ok(ctx, hipMalloc((void **)&a, n16 * 16), "hipMalloc") ||
        !hip_ok(ctx, hipMalloc((void **)&o, (size_t)grid * sizeof(unsigned)), "hipMalloc")) {
        if (a) (void)hipFree(a);
        return 0;
    }
    bool ok = hip_ok(ctx, hipMemsetAsync(a, 1, n16 * 16, ctx->stream), "memset");
    hipLaunchKernelGGL(k_read, dim3(grid), dim3(kThreads), 0, ctx->stream, n16, a, o);
    ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_start, ctx->stream), "event");
    for (int i = 0; i < reps; i++)
        hipLaunchKernelGGL(k_read, dim3(grid), dim3(kThreads), 0, ctx->stream, n16, a, o);
    ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_stop, ctx->stream), "event") &&
         hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    *avg_ms = ms / reps;
    (void)hipFree(a);
    (void)hipFree(o);
    return ok ? 1 : 0;
}

int cicy_copy_bench(cicy_ctx *ctx, int64_t bytes, int reps, double *avg_ms) {
    if (!ctx || bytes < 16 || reps <= 0 || !avg_ms) return 0;
    (void)hipSetDevice(ctx->device);
    const size_t n16 = (size_t)bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (!hip_ok(ctx, hipMalloc((void **)&a, n16 * 16), "hipMalloc") ||
        !hip_ok(ctx, hipMalloc((void **)&b, n16 * 16), "hipMalloc")) {
        if (a) (void)hipFree(a);
        return 0;
    }
    bool ok = hip_ok(ctx, hipMemsetAsync(a, 0, n16 * 16, ctx->stream), "memset");
    const int grid = (int)std::min<size_t>((n16 + kThreads - 1) / kThreads, 256 * 64);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kThreads), 0, ctx->stream, n16, a, b);
    ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_start, ctx->stream), "event");
    for (int i = 0; i < reps; i++)
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kThreads), 0, ctx->stream, n16, a, b);
    ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_stop, ctx->stream), "event") &&
         hip_ok(ctx, hipEventSynchronize(ctx->ev_stop), "sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop);
    *avg_ms = ms / reps;
    (void)hipFree(a);
    (void)hipFree(b);
    return ok ? 1 : 0;
}

int cicy_spmv_sss_f64(cicy_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                      const double *values, const double *diagonal, int64_t nnz, const double *v,
                      double *out) {
    if (!ctx || (n > 0 && (!v || !out))) return 0;
    cicy_sys *s = nullptr;
    if (!cicy_sys_create(ctx, n, nnz, row_ptr, col_ind, values, diagonal, &s)) return 0;
    bool ok = true;
    if (n > 0) {
        ok = h2d(ctx, s->p0, v, (size_t)n);
        dispatch_rb<LaunchSpmv64>(s->rb, s, s->p0, s->q);
        ok = ok && hip_ok(ctx, hipGetLastError(), "k_apply launch") &&
             d2h(ctx, out, s->q, (size_t)n) &&
             hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    }
    cicy_sys_destroy(s);
    return ok ? 1 : 0;
}

int cicy_update_residual(cicy_ctx *ctx, int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                         const double *values, const double *diagonal, int64_t nnz,
                         const double *x0, double *R) {
    if (!ctx || (n > 0 && (!x0 || !R))) return 0;
    cicy_sys *s = nullptr;
    if (!cicy_sys_create(ctx, n, nnz, row_ptr, col_ind, values, diagonal, &s)) return 0;
    bool ok = true;
    if (n > 0) {
        ok = h2d(ctx, s->p0, x0, (size_t)n) && h2d(ctx, s->r, R, (size_t)n);
        dispatch_rb<LaunchSpmv64>(s->rb, s, s->p0, s->q);
        hipLaunchKernelGGL(k_sub, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                           ctx->stream, (int)n, s->r, s->q);
        ok = ok && hip_ok(ctx, hipGetLastError(), "launch") && d2h(ctx, R, s->r, (size_t)n) &&
             hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    }
    cicy_sys_destroy(s);
    return ok ? 1 : 0;
}

int cicy_solve_dea_cg_sss(cicy_ctx *ctx, int64_t nodes, int64_t reserved_nodes, double tolerance,
                          const int64_t *col_ind, const int64_t *row_ptr, const double *values,
                          const double *diagonal, int64_t nnz, double *R, const double *VEF,
                          double *voltages_sol, double *voltages, int64_t max_iter,
                          cicy_solve_stats *stats) {
    if (!ctx) return 0;
    auto t0 = std::chrono::steady_clock::now();
    const int64_t n = nodes - reserved_nodes;
    if (n < 0 || reserved_nodes < 0) {
        set_error(ctx, "invalid node counts (nodes=%lld reserved=%lld)", (long long)nodes,
                  (long long)reserved_nodes);
        if (stats) stats->status = CICY_BAD_INPUT;
        return 0;
    }
    if ((n > 0 && (!R || !voltages_sol)) || (reserved_nodes > 0 && !VEF)) {
        set_error(ctx, "null array");
        if (stats) stats->status = CICY_BAD_INPUT;
        return 0;
    }
    if (max_iter <= 0) max_iter = 300 * nodes;  // Direct_Electrifying.cpp:1559
    cicy_solve_stats local;
    memset(&local, 0, sizeof(local));
    local.true_rel_res = -1.0;
    if (n > 0) {
        cicy_sys *s = nullptr;
        const int method = ctx->dropin_method;
        if (!cicy::sys_create(ctx, n, nnz, row_ptr, col_ind, values, diagonal, &s)) {
            if (stats) stats->status = CICY_BAD_INPUT;
            ctx->dropin_blocks.clear();
            return 0;
        }
        // row groups handed over by cicy_set_dropin_blocks apply to this call only
        std::vector<int64_t> groups;
        groups.swap(ctx->dropin_blocks);
        if (method == CICY_METHOD_CONDENSED_BLOCK && !groups.empty() &&
            (!cicy::sys_prepare(s, method, false) ||
             (s->cond &&
              !cicy::condensed_set_blocks(s, (int64_t)groups.size() / 2, groups.data())))) {
            cicy_sys_destroy(s);
            if (stats) stats->status = CICY_BAD_INPUT;
            return 0;
        }
        bool ok = cicy_sys_set_rhs(s, R, voltages_sol, nullptr) &&
                  cicy_sys_solve(s, method, tolerance, max_iter, CICY_REL_R0, &local) &&
                  cicy_sys_get(s, voltages_sol, R);
        cicy_sys_destroy(s);
        if (!ok) {
            if (stats) {
                *stats = local;
                stats->status = CICY_DEVICE_ERROR;
            }
            return 0;
        }
    } else {
        local.status = CICY_OK;
    }
    if (voltages) {
        for (int64_t i = 0; i < reserved_nodes; i++) voltages[i] = VEF[i];
        for (int64_t i = reserved_nodes; i < nodes; i++) voltages[i] = voltages_sol[i - reserved_nodes];
    }
    local.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = local;
    return 1;
}

int cicy_boundary_current_f64(cicy_ctx *ctx, int64_t ne, const int64_t *n_bnd,
                              const int64_t *n_int, const double *re, const double *voltages,
                              int64_t nv, double *current) {
    if (!ctx || !current || ne < 0 || nv < 0) return 0;
    if (ne > 0 && (!n_bnd || !n_int || !re || !voltages)) {
        set_error(ctx, "null array");
        return 0;
    }
    for (int64_t e = 0; e < ne; e++) {
        if (n_bnd[e] < 0 || n_bnd[e] >= nv || n_int[e] < 0 || n_int[e] >= nv) {
            set_error(ctx, "boundary element %lld has a node outside [0, %lld)", (long long)e,
                      (long long)nv);
            return 0;
        }
    }
    if (ne == 0) {
        *current = 0.0;
        return 1;
    }
    (void)hipSetDevice(ctx->device);
    // inputs, result and reduction words in the context's workspace, which only grows
    const size_t words = 3 * (size_t)ne + (size_t)nv + 1 + kBcRedWords;
    if (words > ctx->bc_words) {
        cicy::dev_release(ctx, ctx->bc_ws);
        ctx->bc_ws = nullptr;
        ctx->bc_words = 0;
        if (!dmalloc(ctx, &ctx->bc_ws, words, nullptr)) return 0;
        ctx->bc_words = words;
    }
    double *red = ctx->bc_ws;
    double *dout = red + kBcRedWords;
    double *dre = dout + 1;
    double *dv = dre + ne;
    long long *dnb = reinterpret_cast<long long *>(dv + nv);
    long long *dni = dnb + ne;
    bool ok = h2d(ctx, dnb, (const long long *)n_bnd, (size_t)ne) &&
              h2d(ctx, dni, (const long long *)n_int, (size_t)ne) &&
              h2d(ctx, dre, re, (size_t)ne) && h2d(ctx, dv, voltages, (size_t)nv);
    if (ok) {
        bc_launch(ctx->stream, ne, dnb, dni, dre, dv, nv, dout, red);
        ok = hip_ok(ctx, hipGetLastError(), "launch") && d2h(ctx, current, dout, 1) &&
             hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    }
    return ok ? 1 : 0;
}

int cicy_boundary_current_f64_noalloc(cicy_ctx *ctx, int64_t ne, const int64_t *n_bnd,
                                      const int64_t *n_int, const double *re,
                                      const double *voltages, int64_t nv, double *current) {
    if (!ctx) return 0;
    if (ne < 0 || nv < 0 || ne > kBcMaxElements || nv > kBcMaxElements) {
        set_error(ctx, "cicy_boundary_current_f64_noalloc: sizes must lie in [0, %lld] "
                       "(ne=%lld nv=%lld)", (long long)kBcMaxElements, (long long)ne,
                  (long long)nv);
        return 0;
    }
    if (!current || (ne > 0 && (!n_bnd || !n_int || !re || !voltages))) {
        set_error(ctx, "cicy_boundary_current_f64_noalloc: null device array");
        return 0;
    }
    if (!ctx->scratch) {
        set_error(ctx, "cicy_boundary_current_f64_noalloc: no scratch address for device %d",
                  ctx->device);
        return 0;
    }
    if (!hip_ok(ctx, hipSetDevice(ctx->device), "hipSetDevice")) return 0;
    // every work-group's partial, the group sums, the counters and the bad-node flag live in
    // this GPU's scratch; bc_launch clears the counters and the flag with a kernel first
    bc_launch(ctx->stream, ne, (const long long *)n_bnd, (const long long *)n_int, re, voltages,
              nv, current, ctx->scratch);
    unsigned int bad = 0;
    const BcRed r = bc_red(ctx->scratch);
    if (!hip_ok(ctx, hipGetLastError(), "launch") ||
        !hip_ok(ctx, hipMemcpyAsync(&bad, r.bad, sizeof(bad), hipMemcpyDeviceToHost, ctx->stream),
                "D2H") ||
        !hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync"))
        return 0;
    if (bad) {
        set_error(ctx, "cicy_boundary_current_f64_noalloc: a boundary element has a node "
                       "outside [0, %lld)", (long long)nv);
        return 0;
    }
    return 1;
}

int cicy_host_blocked_spmv(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                           const double *values, const double *diagonal, int64_t nnz, int rb,
                           const double *v, double *out) {
    return cicy_host_layout_spmv(kFormatSss, n, row_ptr, col_ind, values, diagonal, nnz, rb, v,
                                 out);
}

int cicy_host_layout_spmv(int format, int64_t n, const int64_t *row_ptr,
                          const int64_t *col_ind, const double *values, const double *diagonal,
                          int64_t nnz, int rb, const double *v, double *out) {
    BlockedHost h;
    cicy_ctx dummy;
    if (rb != 128 && rb != 256 && rb != 512 && rb != 1024) return 0;
    if (format != kFormatSss && format != kFormatBand) return 0;
    if (!build_blocked(&dummy, n, nnz, row_ptr, col_ind, values, diagonal, rb, h)) return 0;
    std::vector<double> sy;
    if (format == kFormatBand) {
        // the band kernel, one row block at a time
        BandHost b;
        build_band(h, b);
        for (int k = 0; k < h.nblocks; k++) {
            const int r0 = k * rb, nr = std::min(rb, h.n - r0);
            sy.assign((size_t)nr, 0.0);
            for (int t = 0; t < nr; t++) {
                const int r = r0 + t;
                double y = h.diag[r] * v[r];
                if (r > 0) y += b.band[r] * v[r - 1];
                if (r + 1 < h.n) y += b.band[r + 1] * v[r + 1];
                sy[t] = y;
            }
            for (int e = b.xblk[k]; e < b.xblk[k + 1]; e++) sy[b.xrow[e] - r0] += b.xval[e] * v[b.xnbr[e]];
            for (int t = 0; t < nr; t++) out[r0 + t] = sy[t];
        }
        return 1;
    }
    // the blocked SSS kernel, one row block at a time
    for (int b = 0; b < h.nblocks; b++) {
        int r0 = b * rb, nr = std::min(rb, h.n - r0);
        sy.assign((size_t)nr, 0.0);
        for (int t = 0; t < nr; t++) sy[t] = h.diag[r0 + t] * v[r0 + t];
        for (int t = 0; t < nr; t++) {
            int row = r0 + t;
            double acc = 0.0;
            for (int j = h.rp[row]; j < h.rp[row + 1]; j++) {
                int c = h.col[j];
                if (c >= r0) {
                    acc += h.val[j] * v[c];
                    sy[c - r0] += h.val[j] * v[row];
                } else {
                    acc += h.val[j] * v[c];
                }
            }
            for (int j = h.urp[row]; j < h.urp[row + 1]; j++) acc += h.uval[j] * v[h.ucol[j]];
            sy[t] += acc;
        }
        for (int t = 0; t < nr; t++) out[r0 + t] = sy[t];
    }
    return 1;
}

}  // extern "C"
