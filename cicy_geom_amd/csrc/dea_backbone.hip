// dea_backbone.hip -- the data-parallel part of the backbone extraction on the device.
//
// The reference derives the backbone of a cluster from the voltages of its unit-resistor
// solve (Backbone_Network.cpp:10-61): the "zero current" is 1e-9 times the largest
// |dV| over every element -- tube elements, tube junctions, plate triangulation edges,
// plate and mixed junctions (Find_zero_current, :64-221) -- and the backbone of a tube is
// the run of its elements from the first whose |dV| exceeds the zero current by more than
// 1e-10 up to the next one that does not (Find_backbone_cnts, :317-409).  Both are computed
// here, from the voltages in HBM; the ordered pruning of dead plates and junctions that
// follows is host work (cicy_net_backbone).
//
// k_pair_max: grid-stride maximum of |V[a] - V[b]| over the tube element pairs and the
// other pairs; the work-group maxima are combined with an atomic maximum on the bit
// patterns, which order like the values for non-negative doubles, so the result does not
// depend on timing.  NaN compares false, as in the reference's "if (I > I_max)".
// k_tube_runs: one thread per tube scans its elements.
#include "dea_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace cicy {
namespace {

constexpr double kZeroCurrentTol = 1.0e-10;  // Zero, MathMatrix.h:18

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const double u = __shfl_down(v, o, kWave);
        v = u > v ? u : v;
    }
    return v;
}

__global__ __launch_bounds__(kThreads) void k_pair_max(
    long long nt, const long long *__restrict__ tube_off, const int *__restrict__ tube_nodes,
    long long np, const int *__restrict__ pa, const int *__restrict__ pb,
    const double *__restrict__ v, unsigned long long *__restrict__ out) {
    __shared__ double red[kThreads / kWave];
    double m = 0.0;
    const long long stride = (long long)gridDim.x * kThreads;
    const long long ne = tube_off[nt];  // element points of all tubes
    for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < ne + np;
         e += stride) {
        double c;
        if (e < ne) {
            // element (e, e + 1): the last point of every tube is stored as -1 - node, so an
            // element starts at e exactly when tube_nodes[e] >= 0, and e + 1 is then a point
            // of the same tube
            const int a = tube_nodes[e];
            if (a < 0) continue;
            const int bq = tube_nodes[e + 1];
            c = fabs(v[a] - v[bq < 0 ? -1 - bq : bq]);
        } else {
            const long long q = e - ne;
            c = fabs(v[pa[q]] - v[pb[q]]);
        }
        if (c > m) m = c;
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (int w = 0; w < kThreads / kWave; w++) b = red[w] > b ? red[w] : b;
        atomicMax(out, (unsigned long long)__double_as_longlong(b));
    }
}

__global__ __launch_bounds__(kThreads) void k_tube_runs(
    long long nt, const long long *__restrict__ tube_off, const int *__restrict__ tube_nodes,
    const double *__restrict__ v, const unsigned long long *__restrict__ imax_bits,
    int *__restrict__ seg) {
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= nt) return;
    double imax = __longlong_as_double((long long)*imax_bits);
    if (!(imax > kZeroCurrentTol)) imax = kZeroCurrentTol;
    const double zero = imax * 1e-9;
    const long long a = tube_off[t], b = tube_off[t + 1];
    int idx1 = -1, idx2 = -1;
    for (long long q = a; q + 1 < b; q++) {
        const int n0 = tube_nodes[q] < 0 ? -1 - tube_nodes[q] : tube_nodes[q];
        const int n1 = tube_nodes[q + 1] < 0 ? -1 - tube_nodes[q + 1] : tube_nodes[q + 1];
        const double c = fabs(v[n0] - v[n1]);
        if (c - zero > kZeroCurrentTol) {
            if (idx1 == -1) idx1 = (int)(q - a);
        } else if (idx1 != -1) {
            idx2 = (int)(q - a);
            break;
        }
    }
    seg[2 * t] = idx1;
    seg[2 * t + 1] = idx2;
}

}  // namespace
}  // namespace cicy

extern "C" int cicy_backbone_scan(cicy_ctx *ctx, int64_t nvolt, const double *volt,
                                  int64_t ntubes, const int64_t *tube_off,
                                  const int64_t *tube_nodes, int64_t npairs,
                                  const int64_t *pair_a, const int64_t *pair_b,
                                  double *zero_out, int32_t *seg_out) {
    using namespace cicy;
    if (!ctx || nvolt <= 0 || !volt || ntubes < 0 || !tube_off || npairs < 0 || !zero_out ||
        (ntubes > 0 && !seg_out) || (npairs > 0 && (!pair_a || !pair_b))) {
        if (ctx) set_error(ctx, "cicy_backbone_scan: invalid arguments");
        return 0;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return 0;
    // validate on the host: offsets ascending, every node inside the voltage vector and
    // below 2^31 (the device arrays hold 32-bit node ids)
    const int64_t ne = ntubes > 0 ? tube_off[ntubes] : 0;
    if (tube_off[0] != 0 || nvolt >= (1LL << 31)) {
        set_error(ctx, "cicy_backbone_scan: tube offsets must start at 0; too many nodes");
        return 0;
    }
    for (int64_t t = 0; t < ntubes; t++)
        if (tube_off[t + 1] < tube_off[t]) {
            set_error(ctx, "cicy_backbone_scan: tube offsets decrease at tube %lld",
                      (long long)t);
            return 0;
        }
    if (ne > 0 && !tube_nodes) {
        set_error(ctx, "cicy_backbone_scan: null tube nodes");
        return 0;
    }
    // the last point of every tube is stored as -1 - node, so that the element maximum
    // never pairs the last point of one tube with the first of the next
    std::vector<int> tn((size_t)std::max<int64_t>(ne, 1)), pa, pb;
    for (int64_t t = 0; t < ntubes; t++)
        for (int64_t q = tube_off[t]; q < tube_off[t + 1]; q++) {
            const int64_t n = tube_nodes[q];
            if (n < 0 || n >= nvolt) {
                set_error(ctx, "cicy_backbone_scan: node %lld of tube %lld out of range",
                          (long long)n, (long long)t);
                return 0;
            }
            tn[(size_t)q] = q + 1 == tube_off[t + 1] ? (int)(-1 - n) : (int)n;
        }
    pa.resize((size_t)std::max<int64_t>(npairs, 1));
    pb.resize(pa.size());
    for (int64_t q = 0; q < npairs; q++) {
        if (pair_a[q] < 0 || pair_a[q] >= nvolt || pair_b[q] < 0 || pair_b[q] >= nvolt) {
            set_error(ctx, "cicy_backbone_scan: pair %lld out of range", (long long)q);
            return 0;
        }
        pa[(size_t)q] = (int)pair_a[q];
        pb[(size_t)q] = (int)pair_b[q];
    }
    hipStream_t st = ctx->stream;
    double *dv = nullptr;
    long long *doff = nullptr;
    int *dtn = nullptr, *dpa = nullptr, *dpb = nullptr, *dseg = nullptr;
    unsigned long long *dmax = nullptr;
    std::vector<long long> off(tube_off, tube_off + ntubes + 1);
    bool ok = hip_ok(ctx, hipMalloc((void **)&dv, (size_t)nvolt * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&doff, off.size() * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dtn, tn.size() * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dpa, pa.size() * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dpb, pb.size() * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dseg, (size_t)std::max<int64_t>(2 * ntubes, 1) * 4),
                     "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dmax, 8), "hipMalloc");
    ok = ok &&
         hip_ok(ctx, hipMemcpyAsync(dv, volt, (size_t)nvolt * 8, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(doff, off.data(), off.size() * 8, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dtn, tn.data(), tn.size() * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dpa, pa.data(), pa.size() * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dpb, pb.data(), pb.size() * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemsetAsync(dmax, 0, 8, st), "memset");
    if (ok) {
        const long long work = ne + npairs;
        const int grid = (int)std::max<long long>(
            1, std::min<long long>((work + kThreads - 1) / kThreads, 256 * 16));
        hipLaunchKernelGGL(k_pair_max, dim3(grid), dim3(kThreads), 0, st, (long long)ntubes,
                           doff, dtn, (long long)npairs, dpa, dpb, dv, dmax);
        if (ntubes > 0)
            hipLaunchKernelGGL(k_tube_runs, dim3((unsigned)((ntubes + kThreads - 1) / kThreads)),
                               dim3(kThreads), 0, st, (long long)ntubes, doff, dtn, dv, dmax,
                               dseg);
        ok = hip_ok(ctx, hipGetLastError(), "backbone launch");
    }
    unsigned long long bits = 0;
    ok = ok && hip_ok(ctx, hipMemcpyAsync(&bits, dmax, 8, hipMemcpyDeviceToHost, st), "D2H") &&
         (ntubes == 0 ||
          hip_ok(ctx, hipMemcpyAsync(seg_out, dseg, (size_t)ntubes * 8, hipMemcpyDeviceToHost, st),
                 "D2H")) &&
         hip_ok(ctx, hipStreamSynchronize(st), "sync");
    if (ok) {
        double imax;
        std::memcpy(&imax, &bits, 8);
        if (!(imax > kZeroCurrentTol)) imax = kZeroCurrentTol;
        *zero_out = imax * 1e-9;
    }
    void *ptrs[] = {dv, doff, dtn, dpa, dpb, dseg, dmax};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    return ok ? 1 : 0;
}
