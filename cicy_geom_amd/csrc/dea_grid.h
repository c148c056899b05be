// dea_grid.h -- grid-wide reductions for cooperative (persistent) kernels on gfx950.
//
// A cooperative launch keeps every work-group resident, so work-groups can wait for each
// other.  MI355X spreads the CUs over eight XCDs whose L2 caches are not coherent with each
// other; values shared between work-groups are therefore written and read with
// device-coherent (agent-scope) atomic stores and loads, which are performed at the point
// of coherence, instead of cache-wide L2 write-back/invalidate fences.
#pragma once
#include <hip/hip_runtime.h>

#include "dea_kernels.h"

namespace cicy {

constexpr unsigned kGridSpinLimit = 1u << 24;  // polls before a grid barrier gives up

// Barrier control block; cleared by the host before every launch.  Work-group g arrives
// on counter g % 8 (eight counters on separate 128-byte lines, so arrivals do not all
// contend for one address).
struct GridCtl {
    unsigned cnt[8][32];
    int error;  // set by a work-group whose barrier wait timed out
};

__device__ __forceinline__ double ld_coh(const double *p) {
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coh(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_coh_u(const unsigned *p) {
    return __hip_atomic_load(const_cast<unsigned *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_coh_i(const int *p) {
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of NV values over a work-group of WAVES waves (valid in thread 0), lds: [NV * WAVES].
template <int NV, int WAVES>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double *lds) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) lds[i * WAVES + w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < WAVES; k++) s += lds[i * WAVES + k];
            v[i] = s;
        }
    }
}

// Grid-wide sum of NV (<= STRIDE) values and barrier number m of this launch, over G
// work-groups of WAVES waves.  Every thread first waits for its own stores (values other
// work-groups read after the barrier must be at the point of coherence before its
// work-group arrives).  Thread 0 stores the work-group's sums to its own slot of the
// barrier's parity (part: [2][G][STRIDE]) with agent-scope stores, waits for them, and arrives
// on its counter; lanes 0..7 of wave 0 poll the eight counters (bounded spin); then wave 0
// reads all G slots and sums them in a fixed order, so every work-group obtains
// bit-identical totals (the run-to-run result is deterministic too).  A slot of parity
// m & 1 is rewritten only at barrier m + 2, after every work-group has passed barrier
// m + 1 and hence finished reading it.  Returns false in every thread if the wait timed
// out (or another work-group's did); v[] holds the totals.
template <int NV, int WAVES, int STRIDE = 2>
__device__ __forceinline__ bool grid_sum(double (&v)[NV], int G, double *part, GridCtl *ctl,
                                         unsigned m, double *lds, int *flag) {
    static_assert(NV <= STRIDE, "grid_sum: more values than slot entries");
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wg_sum<NV, WAVES>(v, lds);
    double *slot = part + (size_t)(m & 1u) * (size_t)G * STRIDE;
    if (threadIdx.x < kWave) {
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) st_coh(&slot[(size_t)blockIdx.x * STRIDE + i], v[i]);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            (void)atomicAdd(&ctl->cnt[blockIdx.x & 7][0], 1u);
        }
        int err = 0;
        if (threadIdx.x < 8) {
            const unsigned share = (int)threadIdx.x < G ? (unsigned)((G - threadIdx.x + 7) / 8) : 0u;
            const unsigned target = (m + 1u) * share;
            unsigned spins = 0;
            while (ld_coh_u(&ctl->cnt[threadIdx.x][0]) < target) {
                if (++spins > kGridSpinLimit || ld_coh_i(&ctl->error)) {
                    err = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        err = __any(err) ? 1 : 0;
        double sum[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) sum[i] = 0.0;
        for (int k = threadIdx.x; k < G; k += kWave) {
#pragma unroll
            for (int i = 0; i < NV; i++) sum[i] += ld_coh(&slot[(size_t)k * STRIDE + i]);
        }
#pragma unroll
        for (int i = 0; i < NV; i++) sum[i] = wave_sum(sum[i]);
        if (threadIdx.x == 0) {
            if (err)
                __hip_atomic_store(&ctl->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int i = 0; i < NV; i++) lds[i] = sum[i];
            *flag = err;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = lds[i];
    const bool ok = *flag == 0;
    __syncthreads();  // lds is reused by the next reduction
    return ok;
}

// ---- grid sums without a counter --------------------------------------------------
// The arrival is the value itself: slots hold 64-bit words, kSlotEmpty marks a slot not
// yet written for the current barrier (a signalling-NaN pattern: arithmetic never yields
// a signalling NaN, it quiets them).  Barrier m reads buffer m % 3 of part ([3][G][STRIDE]
// words); on arriving there each work-group empties its slots of buffer (m + 1) % 3, which
// every work-group has finished reading (they were written at barrier m - 2, and this
// work-group has passed barrier m - 1).  The host fills part with kSlotEmpty before a
// launch (hipMemsetD32Async with kSlotEmptyWord).
constexpr unsigned kSlotEmptyWord = 0x7ff4deadu;
constexpr unsigned long long kSlotEmpty = 0x7ff4dead7ff4deadull;
constexpr int kSlotLanes = 4;  // slots read per lane: at most kSlotLanes * 64 work-groups

__device__ __forceinline__ void st_coh_u64(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_coh_u64(const unsigned long long *p) {
    return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide sum of NV (<= STRIDE) values and barrier number m of this launch, over G <=
// kSlotLanes * 64 work-groups of WAVES waves, as grid_sum (same guarantees: every store a
// thread issued before the call is at the point of coherence before its work-group
// arrives; totals summed in a fixed order, bit-identical in every work-group).  Thread 0
// stores the work-group's sums (its arrival); lane l of wave 0 reads the slots of
// work-groups l, l + 64, ... all at once and again only those still empty (bounded spin).
template <int NV, int WAVES, int STRIDE>
__device__ __forceinline__ bool grid_sum_slots(double (&v)[NV], int G, unsigned long long *part,
                                               GridCtl *ctl, unsigned m, double *lds,
                                               int *flag) {
    static_assert(NV <= STRIDE, "grid_sum_slots: more values than slot entries");
    const size_t gs = (size_t)G * STRIDE;
    if (threadIdx.x < (unsigned)NV)
        st_coh_u64(&part[((m + 1u) % 3u) * gs + (size_t)blockIdx.x * STRIDE + threadIdx.x],
                   kSlotEmpty);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    wg_sum<NV, WAVES>(v, lds);
    if (threadIdx.x < kWave) {
        unsigned long long *slot = part + (m % 3u) * gs;
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++)
                st_coh_u64(&slot[(size_t)blockIdx.x * STRIDE + i],
                           (unsigned long long)__double_as_longlong(v[i]));
        }
        unsigned long long w[kSlotLanes][NV];
#pragma unroll
        for (int j = 0; j < kSlotLanes; j++) {
            const int k = (int)threadIdx.x + j * kWave;
#pragma unroll
            for (int i = 0; i < NV; i++)
                w[j][i] = k < G ? ld_coh_u64(&slot[(size_t)k * STRIDE + i]) : 0ull;
        }
        int err = 0;
        unsigned spins = 0;
        while (true) {
            bool pending = false;
#pragma unroll
            for (int j = 0; j < kSlotLanes; j++)
#pragma unroll
                for (int i = 0; i < NV; i++) pending = pending || w[j][i] == kSlotEmpty;
            if (!pending) break;
            if (++spins > kGridSpinLimit || ld_coh_i(&ctl->error)) {
                err = 1;
                __hip_atomic_store(&ctl->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int j = 0; j < kSlotLanes; j++) {
                const int k = (int)threadIdx.x + j * kWave;
#pragma unroll
                for (int i = 0; i < NV; i++)
                    if (w[j][i] == kSlotEmpty) w[j][i] = ld_coh_u64(&slot[(size_t)k * STRIDE + i]);
            }
        }
        err = __any(err) ? 1 : 0;
        double sum[NV];
#pragma unroll
        for (int i = 0; i < NV; i++) {
            sum[i] = 0.0;
#pragma unroll
            for (int j = 0; j < kSlotLanes; j++) sum[i] += __longlong_as_double((long long)w[j][i]);
            sum[i] = wave_sum(sum[i]);
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < NV; i++) lds[i] = sum[i];
            *flag = err;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = lds[i];
    const bool ok = *flag == 0;
    __syncthreads();  // lds is reused by the next reduction
    return ok;
}

}  // namespace cicy
