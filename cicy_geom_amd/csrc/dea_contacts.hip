// dea_contacts.hip -- tube-tube contact search and Hoshen-Kopelman cluster labels.
//
// The reference finds contacts between tube points inside overlapping cubic subregions
// (Contact_grid.cpp:18-295): a point goes to the subregion that holds it and, when it lies
// within the overlap of a face, to the one neighbour across that face (at most one per
// axis; never diagonal neighbours).  Every pair of points of different tubes in one
// subregion, visited as (subregion, j < k), is a contact when
//   d = |P1 - P2| - r1 - r2, raised to d_vdW when below it, is <= d_tunnel
// (Hoshen_Kopelman::Label_cnts_in_window, Hoshen_Kopelman.cpp:249-371), and each contact
// runs HK76 (:2908-2989) and records the closest point of the other tube
// (point_contacts, strictly smaller distance wins, so the first in visiting order keeps
// ties).  Labels are then renumbered (Cleanup_labels, :374-425), runs of contacts between
// two tubes are compressed into one junction each (Compress_cnt_cnt_contact_segments,
// :443-555) and tube end points complete the element sets (Complete_cnt_elements, :557-573).
//
// Device, in HBM:
//   k_place  one thread per visit v (the window's tubes in order, their points in order):
//            home subregion and face neighbours, as keys (subregion << 32 | v);
//   radix sort of the keys: slot order = the reference's subregion lists, since within a
//            subregion the reference appends points in visiting order;
//   k_slots  slot-ordered arrays of coordinates, tube, radius and subregion;
//   k_pairs  one thread per slot, against the later slots of its subregion: a counting
//            pass, an exclusive scan, and a writing pass that puts every contact at its
//            place in the reference's visiting order (slot, then later slot);
//   point_contacts: a stable radix sort of the contacts by (Pa, other tube) keeps visiting
//            order inside each group; k_closest keeps the group's closest contact (the first
//            on ties); a stable sort of those entries by (tube of Pa, other tube) gives the
//            order Compress_cnt_cnt_contact_segments walks, Pa ascending inside each pair.
//   labels   union-find over the contacts' tube pairs (lock-free, larger root linked under
//            the smaller), the first contact of every component, and a sort of the
//            components by it: HK76 + Cleanup_labels' numbering (see k_uf_union below).
// Host: the segment compression over the sorted entries -- the sequential part of the
// reference, O(contacts), split over host threads at (tube, other tube) boundaries.  Labels, junctions and
// elements are the reference's bit for bit: the distance is evaluated with the same IEEE
// operations in the same order (no contraction into fused multiply-adds; the FP64 square
// root is correctly rounded), and so are the subregion coordinates and overlap flags.
#include "dea_internal.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace cicy {
namespace {

constexpr double kZeroTol = 1.0e-10;  // Zero, MathMatrix.h:18
constexpr unsigned long long kNoKey = ~0ull;

struct Grid {
    int n[3];
    double l[3], lo[3];
    double ov;  // gs_overlap_cnt
};

// Adjust_regions_if_needed and the region counts of Generate_contact_grid
// (Contact_grid.cpp:21-46, :82-102).
Grid make_grid(const double *geom) {
    Grid g;
    const double ov = std::max(geom[9], geom[10]);
    for (int k = 0; k < 3; k++) {
        g.lo[k] = geom[k];
        g.n[k] = std::max(1, (int)(geom[3 + k] / geom[6 + k]));
        g.l[k] = geom[6 + k];
        if (geom[6 + k] < 2 * ov) {
            g.l[k] = 2.0 * (ov + kZeroTol);
            g.n[k] = 1 + (int)(geom[3 + k] / g.l[k]);
        }
    }
    g.ov = geom[9];
    return g;
}

// Calculate_region_coordinates + Calculate_overlapping_flags + Assign_point_to_regions_cnts
// (Contact_grid.cpp:163-295) for one point: up to four subregions, home first.
__global__ __launch_bounds__(kThreads) void k_place(
    long long nv, const int *__restrict__ vis_pt, const double *__restrict__ xyz, Grid g,
    unsigned long long *__restrict__ keys, int *__restrict__ bad) {
#pragma clang fp contract(off)
    const long long v = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (v < nv) {
        const double *p0 = xyz + 3 * (long long)vis_pt[v];
        // a coordinate that is not finite fails the call; it is placed at the window corner
        // meanwhile, so every key stays a valid subregion
        const bool fin = isfinite(p0[0]) && isfinite(p0[1]) && isfinite(p0[2]);
        if (!fin) atomicOr(bad, 1);
        const double p[3] = {fin ? p0[0] : g.lo[0], fin ? p0[1] : g.lo[1],
                             fin ? p0[2] : g.lo[2]};
        int co[3], f[3];
        for (int k = 0; k < 3; k++) {
            const double q = (p[k] - g.lo[k]) / g.l[k];
            int a = q >= (double)g.n[k] ? g.n[k] - 1 : (q <= -1.0 ? 0 : (int)q);
            if (a < 0) a = 0;
            co[k] = a;
            const double x1 = (double)a * g.l[k] + g.lo[k];
            const double x2 = x1 + g.l[k];
            f[k] = 0;
            if (a > 0 && p[k] >= x1 && p[k] <= x1 + g.ov)
                f[k] = -1;
            else if (a < g.n[k] - 1 && p[k] >= x2 - g.ov && p[k] <= x2)
                f[k] = 1;
        }
        const long long sxy = (long long)g.n[0] * g.n[1];
        const long long t = co[0] + (long long)co[1] * g.n[0] + co[2] * sxy;
        const unsigned long long lo = (unsigned long long)v;
        keys[4 * v] = ((unsigned long long)t << 32) | lo;
        keys[4 * v + 1] = f[0] ? ((unsigned long long)(t + f[0]) << 32) | lo : kNoKey;
        keys[4 * v + 2] = f[1] ? ((unsigned long long)(t + (long long)f[1] * g.n[0]) << 32) | lo
                               : kNoKey;
        keys[4 * v + 3] = f[2] ? ((unsigned long long)(t + f[2] * sxy) << 32) | lo : kNoKey;
    }
}

// The sorted keys end with the absent neighbours (kNoKey): the slot count is the position of
// the first of them.
__global__ void k_first_absent(long long n, const unsigned long long *__restrict__ keys,
                               unsigned long long *__restrict__ out) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = lo + (hi - lo) / 2;
        if (keys[mid] == kNoKey)
            hi = mid;
        else
            lo = mid + 1;
    }
    *out = (unsigned long long)lo;
}

__global__ __launch_bounds__(kThreads) void k_slots(
    long long slots, const unsigned long long *__restrict__ keys, const int *__restrict__ vis_pt,
    const double *__restrict__ xyz, const int *__restrict__ flag,
    const double *__restrict__ radii, double *__restrict__ sx, double *__restrict__ sy,
    double *__restrict__ sz, double *__restrict__ sr, int *__restrict__ sc,
    int *__restrict__ sreg, int *__restrict__ spt) {
    const long long s = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (s >= slots) return;
    const unsigned long long key = keys[s];
    const int p = vis_pt[(long long)(key & 0xffffffffull)];
    const int c = flag[p];
    sx[s] = xyz[3 * (long long)p];
    sy[s] = xyz[3 * (long long)p + 1];
    sz[s] = xyz[3 * (long long)p + 2];
    sc[s] = c;
    sr[s] = radii[c];
    sreg[s] = (int)(key >> 32);
    spt[s] = p;
}

// One thread per slot s: tests it against the later slots of its subregion.  WRITE = false
// counts contacts (and pair tests); WRITE = true stores them from off[s] on, in k order.
template <bool WRITE>
__global__ __launch_bounds__(kThreads) void k_pairs(
    long long slots, const int *__restrict__ sreg, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ sr,
    const int *__restrict__ sc, const int *__restrict__ spt, double vdw, double tunnel,
    int *__restrict__ cnt, unsigned long long *__restrict__ tests,  // per slot
    const long long *__restrict__ off, int *__restrict__ p1, int *__restrict__ p2,
    double *__restrict__ dj) {
#pragma clang fp contract(off)
    const long long s = (long long)blockIdx.x * kThreads + threadIdx.x;
    int found = 0;
    unsigned long long tried = 0;
    if (s < slots) {
        const int t = sreg[s];
        const double x = sx[s], y = sy[s], z = sz[s], r = sr[s];
        const int c = sc[s];
        long long w = WRITE ? off[s] : 0;
        const int ps = WRITE ? spt[s] : 0;
        for (long long k = s + 1; k < slots && sreg[k] == t; k++) {
            tried++;
            if (sc[k] == c) continue;
            const double dx = x - sx[k];
            const double dy = y - sy[k];
            const double dz = z - sz[k];
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            double d = dist - r - sr[k];
            if (d < vdw) d = vdw;
            if (d <= tunnel) {
                if (WRITE) {
                    p1[w] = ps;
                    p2[w] = spt[k];
                    dj[w] = d;
                    w++;
                }
                found++;
            }
        }
        if (!WRITE) {
            cnt[s] = found;
            tests[s] = tried;
        }
    }
}

// Per contact i (visiting order): its two tubes, for the host's HK76 replay, and the sort key
// of point_contacts[Pa][b] -- Pa the point of the lower-numbered tube, b the other tube.
__global__ __launch_bounds__(kThreads) void k_contact_keys(
    long long n, const int *__restrict__ p1, const int *__restrict__ p2,
    const int *__restrict__ flag, int cb, int2 *__restrict__ tubes,
    unsigned long long *__restrict__ key, unsigned *__restrict__ val) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int a1 = p1[i], a2 = p2[i];
    const int c1 = flag[a1], c2 = flag[a2];
    tubes[i] = make_int2(c1, c2);
    const int pa = c1 < c2 ? a1 : a2;
    const int b = c1 < c2 ? c2 : c1;
    key[i] = ((unsigned long long)pa << cb) | (unsigned long long)b;
    val[i] = (unsigned)i;
}

__global__ __launch_bounds__(kThreads) void k_heads(long long n,
                                                    const unsigned long long *__restrict__ key,
                                                    int *__restrict__ head) {
    const long long j = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    head[j] = (j == 0 || key[j] != key[j - 1]) ? 1 : 0;
}

// One thread per (Pa, b) group of the stably sorted contacts: the closest contact, the first
// in visiting order on ties (point_contacts_dist keeps a value unless a strictly smaller one
// comes), as entry pos[j], with the key (tube of Pa, b) of the compression order.
__global__ __launch_bounds__(kThreads) void k_closest(
    long long n, const unsigned long long *__restrict__ key, const unsigned *__restrict__ val,
    const int *__restrict__ head, const int *__restrict__ pos, const int *__restrict__ p1,
    const int *__restrict__ p2, const double *__restrict__ dj, const int *__restrict__ flag,
    int cb, int *__restrict__ epa, int *__restrict__ epb, double *__restrict__ ed,
    unsigned long long *__restrict__ key2, unsigned *__restrict__ val2) {
    const long long j = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (j >= n || !head[j]) return;
    const unsigned long long k = key[j];
    unsigned best = val[j];
    double bd = dj[best];
    for (long long u = j + 1; u < n && key[u] == k; u++) {
        const unsigned c = val[u];
        if (dj[c] < bd) {
            bd = dj[c];
            best = c;
        }
    }
    const int pa = (int)(k >> cb);
    const unsigned long long b = k & ((1ull << cb) - 1);
    const int e = pos[j];
    epa[e] = pa;
    epb[e] = p1[best] == pa ? p2[best] : p1[best];
    ed[e] = bd;
    key2[e] = ((unsigned long long)flag[pa] << cb) | b;
    val2[e] = (unsigned)e;
}

__global__ __launch_bounds__(kThreads) void k_gather_entries(
    long long n, const unsigned *__restrict__ order, const int *__restrict__ epa,
    const int *__restrict__ epb, const double *__restrict__ ed, int *__restrict__ opa,
    int *__restrict__ opb, double *__restrict__ od) {
    const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= n) return;
    const unsigned e = order[q];
    opa[q] = epa[e];
    opb[q] = epb[e];
    od[q] = ed[e];
}

// ---- cluster labels on the device
// HK76 with smaller-root merging numbers clusters by their first contact: the first contact of
// a cluster (visiting order) always meets two unlabelled tubes, so it creates the cluster's
// smallest label, which stays its root, and Cleanup_labels (Hoshen_Kopelman.cpp:374-425)
// numbers roots in increasing order.  So: union-find over the contacts' tube pairs, the first
// contact of every component, and components ranked by it -- the reference's labels_cnt.
constexpr unsigned kNoFirst = 0xffffffffu;
constexpr int kUnionTries = 1 << 20;  // a retry bound no union reaches; exceeding it fails

__device__ __forceinline__ int ld_agent(const int *p) {
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// parent[x] <= x always (a root is only ever linked under a smaller one), so the walk ends.
// Path halving: a non-root is pointed at its grandparent on the way.  That store races only
// with other halving stores (a compare-and-swap targets roots, and a non-root never becomes
// one again), and every value written is an ancestor, so the partition is unchanged.
__device__ __forceinline__ int uf_find(int *parent, int x) {
    int p = ld_agent(parent + x);
    while (p != x) {
        const int gp = ld_agent(parent + p);
        if (gp != p)
            __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
        p = ld_agent(parent + x);
    }
    return x;
}

__global__ __launch_bounds__(kThreads) void k_uf_init(long long n, int *__restrict__ parent,
                                                      unsigned *__restrict__ first_tube,
                                                      unsigned *__restrict__ first_comp) {
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= n) return;
    parent[t] = (int)t;
    first_tube[t] = kNoFirst;
    first_comp[t] = kNoFirst;
}

// The first contact of every tube (visiting order).
__global__ __launch_bounds__(kThreads) void k_first_contact(long long n,
                                                            const int2 *__restrict__ tubes,
                                                            unsigned *first_tube) {
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int2 c = tubes[i];
    atomicMin(first_tube + c.x, (unsigned)i);
    atomicMin(first_tube + c.y, (unsigned)i);
}

// Every distinct tube pair in contact -- the heads of the entries sorted by (a, b) -- joins
// its two tubes: lock-free, the larger root linked under the smaller with a compare-and-swap,
// retried on a lost race.
__global__ __launch_bounds__(kThreads) void k_uf_union(long long n,
                                                       const unsigned long long *__restrict__ ab,
                                                       int cb, int *parent, int *fail) {
    const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= n || (q > 0 && ab[q] == ab[q - 1])) return;
    int a = (int)(ab[q] >> cb), b = (int)(ab[q] & ((1ull << cb) - 1));
    for (int it = 0; it < kUnionTries; it++) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        int expected = b;
        if (__hip_atomic_compare_exchange_strong(parent + b, &expected, a, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
    atomicOr(fail, 1);
}

// root of every tube; the first contact of each component (one atomic per wave when the
// wave's tubes share a component, the common case inside a large cluster)
__global__ __launch_bounds__(kThreads) void k_uf_roots(long long n, int *parent,
                                                       const unsigned *__restrict__ first_tube,
                                                       int *__restrict__ root,
                                                       unsigned *first_comp) {
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    const bool in = t < n;
    const int r = in ? uf_find(parent, (int)t) : -1;
    if (in) root[t] = r;
    const unsigned f = in ? first_tube[t] : kNoFirst;
    const bool valid = f != kNoFirst;
    const unsigned long long vm = __ballot(valid);
    if (!vm) return;
    const int src = __ffsll((long long)vm) - 1;
    const int r0 = __shfl(r, src, kWave);
    if (__all(!valid || r == r0)) {
        unsigned m = f;
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const unsigned u = __shfl_down(m, o, kWave);
            m = u < m ? u : m;
        }
        if ((threadIdx.x & (kWave - 1)) == 0) atomicMin(first_comp + r0, m);
    } else if (valid) {
        atomicMin(first_comp + r, f);
    }
}

__global__ __launch_bounds__(kThreads) void k_rank_keys(long long n, const int *__restrict__ root,
                                                        const unsigned *__restrict__ first_comp,
                                                        unsigned *__restrict__ key,
                                                        int *__restrict__ val) {
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= n) return;
    key[t] = root[t] == (int)t ? first_comp[t] : kNoFirst;
    val[t] = (int)t;
}

// labels: rank of the tube's component among components ordered by first contact; -1 for a
// tube without contact.  keys sorted ascending, absent ones (kNoFirst) last.
__global__ __launch_bounds__(kThreads) void k_rank(long long n, const unsigned *__restrict__ key,
                                                   const int *__restrict__ val,
                                                   int *__restrict__ rank,
                                                   unsigned long long *__restrict__ n_labels) {
    const long long q = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (q >= n) return;
    if (key[q] != kNoFirst) rank[val[q]] = (int)q;
    if (key[q] != kNoFirst && (q + 1 == n || key[q + 1] == kNoFirst)) *n_labels = q + 1;
}

__global__ __launch_bounds__(kThreads) void k_labels(long long n, const int *__restrict__ root,
                                                     const unsigned *__restrict__ first_tube,
                                                     const int *__restrict__ rank,
                                                     int *__restrict__ labels) {
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= n) return;
    labels[t] = first_tube[t] != kNoFirst ? rank[root[t]] : -1;
}

// Device buffers of one call, freed together.
struct Bufs {
    std::vector<void *> p;
    template <class T>
    bool get(cicy_ctx *ctx, T *&ptr, size_t n) {
        ptr = nullptr;
        if (!hip_ok(ctx, hipMalloc((void **)&ptr, std::max<size_t>(n, 1) * sizeof(T)),
                    "hipMalloc"))
            return false;
        p.push_back(ptr);
        return true;
    }
    ~Bufs() {
        for (void *q : p) (void)hipFree(q);
    }
};

}  // namespace
}  // namespace cicy

extern "C" int cicy_cnt_contacts(cicy_ctx *ctx, int64_t n_points, const double *xyz,
                                 int64_t n_cnts, const double *radii, const int64_t *struct_off,
                                 const int64_t *struct_pts, int64_t n_inside,
                                 const int64_t *cnts_inside, const double *geom, double vdw,
                                 double tunnel, int64_t min_points, int32_t *labels,
                                 int64_t *n_labels, int64_t junc_cap, int64_t *n_junc,
                                 int64_t *junc, double *junc_d, int64_t *elem_off,
                                 int64_t *elem_pts, double *stats) {
    using namespace cicy;
    using clk = std::chrono::steady_clock;
    const auto t_start = clk::now();
    if (!ctx) return 0;
    if (n_points < 0 || n_points >= (1LL << 31) || n_cnts < 0 || n_cnts >= (1LL << 31) ||
        (n_points > 0 && !xyz) || (n_cnts > 0 && (!radii || !labels)) || !struct_off ||
        n_inside < 0 || (n_inside > 0 && !cnts_inside) || !geom || !n_labels || !n_junc ||
        junc_cap < 0 || (junc_cap > 0 && (!junc || !junc_d)) || !elem_off ||
        !(vdw == vdw) || !(tunnel == tunnel)) {
        set_error(ctx, "cicy_cnt_contacts: invalid arguments");
        return 0;
    }
    for (int k = 0; k < 11; k++)
        if (!std::isfinite(geom[k]) || (k >= 3 && k < 9 && geom[k] <= 0.0) ||
            (k >= 9 && geom[k] < 0.0)) {
            set_error(ctx, "cicy_cnt_contacts: window, grid or overlap value %d invalid", k);
            return 0;
        }
    // host threads for the O(points) checks and copies (up to 16)
    auto run_threads = [](int nt, auto &&fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(fn, t);
        fn(0);
        for (auto &x : th) x.join();
    };
    const int64_t n_struct = struct_off[n_cnts];
    const int hth = (int)std::max<long long>(
        1, std::min<long long>({16, (long long)std::max(1u, std::thread::hardware_concurrency()),
                                (long long)(n_struct / (1 << 18)) + 1}));
    // the tube of every point (the reference's Point_3D::flag); a point may be in one tube
    std::vector<int> flag((size_t)std::max<int64_t>(n_points, 1), -1);
    if (struct_off[0] != 0) {
        set_error(ctx, "cicy_cnt_contacts: structure offsets must start at 0");
        return 0;
    }
    for (int64_t i = 0; i < n_cnts; i++)
        if (struct_off[i + 1] < struct_off[i]) {
            set_error(ctx, "cicy_cnt_contacts: structure offsets decrease at tube %lld",
                      (long long)i);
            return 0;
        }
    {
        std::vector<int64_t> bad_pt((size_t)hth, -1), bad_tube((size_t)hth, -1);
        run_threads(hth, [&](int t) {
            const int64_t i0 = n_cnts * t / hth, i1 = n_cnts * (t + 1) / hth;
            for (int64_t i = i0; i < i1; i++)
                for (int64_t q = struct_off[i]; q < struct_off[i + 1]; q++) {
                    const int64_t P = struct_pts[q];
                    int expected = -1;
                    if (P < 0 || P >= n_points ||
                        !__atomic_compare_exchange_n(&flag[(size_t)P], &expected, (int)i, false,
                                                     __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
                        bad_pt[(size_t)t] = P;
                        bad_tube[(size_t)t] = i;
                        return;
                    }
                }
        });
        for (int t = 0; t < hth; t++)
            if (bad_tube[(size_t)t] >= 0) {
                set_error(ctx, "cicy_cnt_contacts: point %lld of tube %lld is out of range "
                          "or in two tubes", (long long)bad_pt[(size_t)t],
                          (long long)bad_tube[(size_t)t]);
                return 0;
            }
    }
    const Grid g = make_grid(geom);
    const double tot_d = (double)g.n[0] * g.n[1] * g.n[2];
    if (tot_d > 1.0e8) {
        set_error(ctx, "cicy_cnt_contacts: %.0f subregions (at most 1e8)", tot_d);
        return 0;
    }
    // the visiting order of Fill_sectioned_domain_cnts (Contact_grid.cpp:104-161)
    std::vector<int64_t> voff((size_t)n_inside + 1, 0);
    for (int64_t i = 0; i < n_inside; i++) {
        const int64_t c = cnts_inside[i];
        if (c < 0 || c >= n_cnts) {
            set_error(ctx, "cicy_cnt_contacts: cnts_inside[%lld] = %lld out of range",
                      (long long)i, (long long)c);
            return 0;
        }
        voff[(size_t)i + 1] = voff[(size_t)i] + struct_off[c + 1] - struct_off[c];
    }
    std::vector<int> vis((size_t)voff[(size_t)n_inside]);
    run_threads(hth, [&](int t) {
        const int64_t i0 = n_inside * t / hth, i1 = n_inside * (t + 1) / hth;
        for (int64_t i = i0; i < i1; i++) {
            const int64_t c = cnts_inside[i];
            std::copy(struct_pts + struct_off[c], struct_pts + struct_off[c + 1],
                      vis.data() + voff[(size_t)i]);
        }
    });
    const long long nv = (long long)vis.size();
    if (4 * nv >= (1LL << 31) || nv >= (1LL << 30)) {
        set_error(ctx, "cicy_cnt_contacts: %lld visits (at most 2^29)", nv);
        return 0;
    }
    for (int &f : flag)
        if (f < 0) f = 0;  // points outside every tube are never visited
    const auto t_host = clk::now();

    // ---- device
    if (hipSetDevice(ctx->device) != hipSuccess) return 0;
    hipStream_t st = ctx->stream;
    Bufs B;
    const size_t NP = (size_t)std::max<int64_t>(n_points, 1);
    const size_t NC = (size_t)std::max<int64_t>(n_cnts, 1);
    int cb = 1, pb = 1;  // bits of a tube id and of a point id
    while ((1LL << cb) < n_cnts) cb++;
    while ((1LL << pb) < n_points) pb++;
    auto grid_of = [](long long n) {
        return dim3((unsigned)((std::max<long long>(n, 1) + kThreads - 1) / kThreads));
    };
    // hipCUB calls: size query, scratch, run
    auto cub = [&](auto &&call, const char *what) {
        size_t tmp = 0;
        if (!hip_ok(ctx, call(nullptr, tmp), what)) return false;
        char *d = nullptr;
        return B.get(ctx, d, tmp) && hip_ok(ctx, call(d, tmp), what);
    };
    double *dxyz, *drad;
    int *dflag, *dvis, *dbad;
    unsigned long long *keys, *keys2, *dctr;
    bool ok = B.get(ctx, dxyz, 3 * NP) && B.get(ctx, drad, NC) && B.get(ctx, dflag, NP) &&
              B.get(ctx, dvis, (size_t)nv) && B.get(ctx, keys, 4 * (size_t)nv) &&
              B.get(ctx, keys2, 4 * (size_t)nv) && B.get(ctx, dctr, 2) && B.get(ctx, dbad, 1);
    ok = ok &&
         (n_points == 0 ||
          hip_ok(ctx, hipMemcpyAsync(dxyz, xyz, 3 * (size_t)n_points * 8, hipMemcpyHostToDevice,
                                     st), "H2D")) &&
         (n_cnts == 0 || hip_ok(ctx, hipMemcpyAsync(drad, radii, (size_t)n_cnts * 8,
                                                    hipMemcpyHostToDevice, st), "H2D")) &&
         hip_ok(ctx, hipMemcpyAsync(dflag, flag.data(), NP * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         (nv == 0 || hip_ok(ctx, hipMemcpyAsync(dvis, vis.data(), (size_t)nv * 4,
                                                hipMemcpyHostToDevice, st), "H2D")) &&
         hip_ok(ctx, hipMemsetAsync(dctr, 0, 16, st), "memset") &&
         hip_ok(ctx, hipMemsetAsync(dbad, 0, 4, st), "memset");
    int tbits = 1;
    while ((1LL << tbits) <= (long long)tot_d) tbits++;  // kNoKey's field exceeds every t
    const int nkeys = (int)(4 * nv);
    if (ok && nv > 0) {
        hipLaunchKernelGGL(k_place, grid_of(nv), dim3(kThreads), 0, st, nv, dvis, dxyz, g, keys,
                           dbad);
        ok = cub([&](void *d, size_t &n) {
                 return hipcub::DeviceRadixSort::SortKeys(d, n, keys, keys2, nkeys, 0,
                                                          32 + tbits, st);
             }, "radix sort") ;
        if (ok)
            hipLaunchKernelGGL(k_first_absent, dim3(1), dim3(1), 0, st, (long long)nkeys, keys2,
                               dctr);
    }
    unsigned long long ctr[2] = {0, 0};
    int bad = 0;
    ok = ok && hip_ok(ctx, hipGetLastError(), "launch") &&
         hip_ok(ctx, hipMemcpyAsync(ctr, dctr, 16, hipMemcpyDeviceToHost, st), "D2H") &&
         hip_ok(ctx, hipMemcpyAsync(&bad, dbad, 4, hipMemcpyDeviceToHost, st), "D2H") &&
         hip_ok(ctx, hipStreamSynchronize(st), "sync");
    if (!ok) return 0;
    if (bad) {
        set_error(ctx, "cicy_cnt_contacts: a point of a window tube has a coordinate that is "
                  "not finite");
        return 0;
    }
    const long long slots = (long long)ctr[0];
    double *sx, *sy, *sz, *sr;
    int *sc, *sreg, *spt, *cnt;
    long long *off;
    unsigned long long *tries;
    ok = B.get(ctx, sx, (size_t)slots) && B.get(ctx, sy, (size_t)slots) &&
         B.get(ctx, sz, (size_t)slots) && B.get(ctx, sr, (size_t)slots) &&
         B.get(ctx, sc, (size_t)slots) && B.get(ctx, sreg, (size_t)slots) &&
         B.get(ctx, spt, (size_t)slots) && B.get(ctx, cnt, (size_t)slots + 1) &&
         B.get(ctx, off, (size_t)slots + 1) && B.get(ctx, tries, (size_t)slots);
    float kernel_ms = 0.f;
    long long count = 0;
    if (ok && slots > 0) {
        hipLaunchKernelGGL(k_slots, grid_of(slots), dim3(kThreads), 0, st, slots, keys2, dvis,
                           dxyz, dflag, drad, sx, sy, sz, sr, sc, sreg, spt);
        ok = hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "event");
        hipLaunchKernelGGL(k_pairs<false>, grid_of(slots), dim3(kThreads), 0, st, slots, sreg,
                           sx, sy, sz, sr, sc, spt, vdw, tunnel, cnt, tries, nullptr, nullptr,
                           nullptr, nullptr);
        ok = ok && hip_ok(ctx, hipEventRecord(ctx->ev_stop, st), "event") &&
             hip_ok(ctx, hipMemsetAsync(cnt + slots, 0, 4, st), "memset");
        // exclusive sum over slots + 1 counts (the last one 0): off[slots] = total
        ok = ok &&
             cub([&](void *d, size_t &n) {
                 return hipcub::DeviceScan::ExclusiveSum(d, n, cnt, off, (int)slots + 1, st);
             }, "scan") &&
             cub([&](void *d, size_t &n) {
                 return hipcub::DeviceReduce::Sum(d, n, tries, dctr + 1, (int)slots, st);
             }, "reduce") &&
             hip_ok(ctx, hipMemcpyAsync(&count, off + slots, 8, hipMemcpyDeviceToHost, st),
                    "D2H") &&
             hip_ok(ctx, hipMemcpyAsync(ctr, dctr, 16, hipMemcpyDeviceToHost, st), "D2H") &&
             hip_ok(ctx, hipStreamSynchronize(st), "sync") &&
             hip_ok(ctx, hipEventElapsedTime(&kernel_ms, ctx->ev_start, ctx->ev_stop), "event");
    }
    if (!ok) return 0;
    if (count >= (1LL << 31)) {
        set_error(ctx, "cicy_cnt_contacts: %lld contacts (at most 2^31)", count);
        return 0;
    }
    // contacts in visiting order; then point_contacts on the device: a stable sort by
    // (Pa, b) keeps visiting order inside each group, the group's closest contact becomes an
    // entry, and a stable sort of the entries by (tube of Pa, b) gives the compression order
    // (Pa stays ascending inside each (tube, b), as it was after the first sort)
    long long ne = 0;
    if (count == 0) {
        for (int64_t i = 0; i < n_cnts; i++) labels[i] = -1;
        *n_labels = 0;
    }
    std::vector<int> opa, opb;
    std::vector<unsigned long long> oab;  // (tube of Pa) << cb | b, the entries' sort keys
    std::vector<double> od;
    if (count > 0) {
        const int C = (int)count;
        int *dp1, *dp2, *head, *pos, *epa, *epb, *qa, *qb;
        double *ddj, *ed, *qd;
        int2 *dtubes;
        unsigned long long *k1, *k1s, *k2, *k2s;
        unsigned *v1, *v1s, *v2, *v2s;
        ok = B.get(ctx, dp1, (size_t)C) && B.get(ctx, dp2, (size_t)C) &&
             B.get(ctx, ddj, (size_t)C) && B.get(ctx, dtubes, (size_t)C) &&
             B.get(ctx, k1, (size_t)C) && B.get(ctx, k1s, (size_t)C) &&
             B.get(ctx, v1, (size_t)C) && B.get(ctx, v1s, (size_t)C) &&
             B.get(ctx, head, (size_t)C + 1) && B.get(ctx, pos, (size_t)C + 1);
        int *parent, *root, *rval, *rval_s, *rank, *dlab, *dfail;
        unsigned *first_tube, *first_comp, *rkey, *rkey_s;
        unsigned long long *dnlab;
        const long long T = n_cnts;
        ok = ok && B.get(ctx, parent, (size_t)T) && B.get(ctx, root, (size_t)T) &&
             B.get(ctx, rval, (size_t)T) && B.get(ctx, rval_s, (size_t)T) &&
             B.get(ctx, rank, (size_t)T) && B.get(ctx, dlab, (size_t)T) &&
             B.get(ctx, first_tube, (size_t)T) && B.get(ctx, first_comp, (size_t)T) &&
             B.get(ctx, rkey, (size_t)T) && B.get(ctx, rkey_s, (size_t)T) &&
             B.get(ctx, dfail, 1) && B.get(ctx, dnlab, 1) &&
             hip_ok(ctx, hipMemsetAsync(dfail, 0, 4, st), "memset") &&
             hip_ok(ctx, hipMemsetAsync(dnlab, 0, 8, st), "memset");
        if (ok) {
            hipLaunchKernelGGL(k_pairs<true>, grid_of(slots), dim3(kThreads), 0, st, slots,
                               sreg, sx, sy, sz, sr, sc, spt, vdw, tunnel, nullptr, nullptr, off,
                               dp1, dp2, ddj);
            hipLaunchKernelGGL(k_contact_keys, grid_of(C), dim3(kThreads), 0, st, (long long)C,
                               dp1, dp2, dflag, cb, dtubes, k1, v1);
            // labels_cnt, part 1: the first contact of every tube
            hipLaunchKernelGGL(k_uf_init, grid_of(T), dim3(kThreads), 0, st, T, parent,
                               first_tube, first_comp);
            hipLaunchKernelGGL(k_first_contact, grid_of(C), dim3(kThreads), 0, st,
                               (long long)C, dtubes, first_tube);
        }
        ok = ok &&
             cub([&](void *d, size_t &n) {
                 return hipcub::DeviceRadixSort::SortPairs(d, n, k1, k1s, v1, v1s, C, 0,
                                                           pb + cb, st);
             }, "radix sort") &&
             hip_ok(ctx, hipMemsetAsync(head + C, 0, 4, st), "memset");
        if (ok)
            hipLaunchKernelGGL(k_heads, grid_of(C), dim3(kThreads), 0, st, (long long)C, k1s,
                               head);
        ok = ok &&
             cub([&](void *d, size_t &n) {
                 return hipcub::DeviceScan::ExclusiveSum(d, n, head, pos, C + 1, st);
             }, "scan");
        int E = 0;
        ok = ok && hip_ok(ctx, hipMemcpyAsync(&E, pos + C, 4, hipMemcpyDeviceToHost, st), "D2H") &&
             hip_ok(ctx, hipStreamSynchronize(st), "sync");
        ok = ok && B.get(ctx, epa, (size_t)E) && B.get(ctx, epb, (size_t)E) &&
             B.get(ctx, ed, (size_t)E) && B.get(ctx, k2, (size_t)E) &&
             B.get(ctx, k2s, (size_t)E) && B.get(ctx, v2, (size_t)E) &&
             B.get(ctx, v2s, (size_t)E) && B.get(ctx, qa, (size_t)E) &&
             B.get(ctx, qb, (size_t)E) && B.get(ctx, qd, (size_t)E);
        if (ok)
            hipLaunchKernelGGL(k_closest, grid_of(C), dim3(kThreads), 0, st, (long long)C, k1s,
                               v1s, head, pos, dp1, dp2, ddj, dflag, cb, epa, epb, ed, k2, v2);
        ok = ok &&
             cub([&](void *d, size_t &n) {
                 return hipcub::DeviceRadixSort::SortPairs(d, n, k2, k2s, v2, v2s, E, 0, 2 * cb,
                                                           st);
             }, "radix sort");
        if (ok && E > 0) {
            hipLaunchKernelGGL(k_gather_entries, grid_of(E), dim3(kThreads), 0, st,
                               (long long)E, v2s, epa, epb, ed, qa, qb, qd);
            // labels_cnt, part 2: union-find over the distinct tube pairs, the first contact
            // of every component, components ranked by it
            hipLaunchKernelGGL(k_uf_union, grid_of(E), dim3(kThreads), 0, st, (long long)E,
                               k2s, cb, parent, dfail);
            hipLaunchKernelGGL(k_uf_roots, grid_of(T), dim3(kThreads), 0, st, T, parent,
                               first_tube, root, first_comp);
            hipLaunchKernelGGL(k_rank_keys, grid_of(T), dim3(kThreads), 0, st, T, root,
                               first_comp, rkey, rval);
        }
        ok = ok && cub([&](void *d, size_t &n) {
                 return hipcub::DeviceRadixSort::SortPairs(d, n, rkey, rkey_s, rval, rval_s,
                                                           (int)T, 0, 32, st);
             }, "radix sort");
        if (ok) {
            hipLaunchKernelGGL(k_rank, grid_of(T), dim3(kThreads), 0, st, T, rkey_s, rval_s,
                               rank, dnlab);
            hipLaunchKernelGGL(k_labels, grid_of(T), dim3(kThreads), 0, st, T, root,
                               first_tube, rank, dlab);
        }
        int fail = 0;
        unsigned long long nlab = 0;
        ok = ok && hip_ok(ctx, hipGetLastError(), "labels launch") &&
             hip_ok(ctx, hipMemcpyAsync(labels, dlab, (size_t)T * 4, hipMemcpyDeviceToHost, st),
                    "D2H") &&
             hip_ok(ctx, hipMemcpyAsync(&fail, dfail, 4, hipMemcpyDeviceToHost, st), "D2H") &&
             hip_ok(ctx, hipMemcpyAsync(&nlab, dnlab, 8, hipMemcpyDeviceToHost, st), "D2H");
        ne = E;
        oab.resize((size_t)E);
        opa.resize((size_t)E);
        opb.resize((size_t)E);
        od.resize((size_t)E);
        ok = ok && hip_ok(ctx, hipGetLastError(), "contacts launch") &&
             (E == 0 ||
              (hip_ok(ctx, hipMemcpyAsync(oab.data(), k2s, (size_t)E * 8, hipMemcpyDeviceToHost,
                                          st), "D2H") &&
               hip_ok(ctx, hipMemcpyAsync(opa.data(), qa, (size_t)E * 4, hipMemcpyDeviceToHost,
                                          st), "D2H") &&
               hip_ok(ctx, hipMemcpyAsync(opb.data(), qb, (size_t)E * 4, hipMemcpyDeviceToHost,
                                          st), "D2H") &&
               hip_ok(ctx, hipMemcpyAsync(od.data(), qd, (size_t)E * 8, hipMemcpyDeviceToHost,
                                          st), "D2H"))) &&
             hip_ok(ctx, hipStreamSynchronize(st), "sync");
        if (ok && fail) {
            set_error(ctx, "cicy_cnt_contacts: a union of two clusters did not settle");
            return 0;
        }
        *n_labels = (int64_t)nlab;
    }
    ok = ok && hip_ok(ctx, hipGetLastError(), "contacts launch");
    if (!ok) return 0;
    const auto t_dev = clk::now();

    // ---- host: Compress_cnt_cnt_contact_segments (:443-555) over the entries, already in
    // its order (tube a ascending, then the other tube b, then the point Pa of a ascending).
    // Groups (a, b) are independent: the entries are cut into one chunk per host thread at
    // group boundaries, each chunk compresses its groups, and the junctions are concatenated
    // in chunk order -- the reference's order.
    struct Junc {
        int64_t pi, pj;
        int i, j;
        double d;
    };
    const int nth = (int)std::max<long long>(
        1, std::min<long long>({16, (long long)std::max(1u, std::thread::hardware_concurrency()),
                                ne / 65536 + 1}));
    auto parallel = [&](auto &&fn) {
        std::vector<std::thread> th;
        for (int t = 1; t < nth; t++) th.emplace_back(fn, t);
        fn(0);
        for (auto &x : th) x.join();
    };
    std::vector<long long> cut((size_t)nth + 1, ne);
    for (int t = 0; t < nth; t++) {
        long long c = ne * t / nth;
        while (c > 0 && c < ne && oab[(size_t)c] == oab[(size_t)c - 1]) c++;
        cut[(size_t)t] = c;
    }
    for (int t = nth - 1; t > 0; t--) cut[(size_t)t - 1] = std::min(cut[(size_t)t - 1], cut[(size_t)t]);
    std::vector<std::vector<Junc>> part((size_t)nth);
    parallel([&](int t) {
        std::vector<Junc> &out = part[(size_t)t];
        for (long long q = cut[(size_t)t]; q < cut[(size_t)t + 1];) {
            const unsigned long long key = oab[(size_t)q];
            const int i = (int)(key >> cb), cj = (int)(key & ((1ull << cb) - 1));
            long long e = q + 1;
            while (e < ne && oab[(size_t)e] == key) e++;
            int64_t Pi0 = opa[(size_t)q], Pj0 = opb[(size_t)q];
            double dmin = od[(size_t)q];
            int64_t Pi_j = Pi0, Pj_j = Pj0;
            for (long long u = q + 1; u < e; u++) {
                const int64_t Pi1 = opa[(size_t)u], Pj1 = opb[(size_t)u];
                const double d = od[(size_t)u];
                if (Pi1 - Pi0 >= min_points || std::llabs(Pj1 - Pj0) >= min_points) {
                    out.push_back({Pi_j, Pj_j, i, cj, dmin});
                    dmin = d;
                    Pi_j = Pi1;
                    Pj_j = Pj1;
                } else if (d < dmin) {
                    dmin = d;
                    Pi_j = Pi1;
                    Pj_j = Pj1;
                }
                Pi0 = Pi1;
                Pj0 = Pj1;
            }
            out.push_back({Pi_j, Pj_j, i, cj, dmin});
            q = e;
        }
    });
    std::vector<int64_t> joff((size_t)nth + 1, 0);
    for (int t = 0; t < nth; t++) joff[(size_t)t + 1] = joff[(size_t)t] + (int64_t)part[(size_t)t].size();
    const int64_t nj = joff[(size_t)nth];
    *n_junc = nj;
    // Complete_cnt_elements (:557-573): each tube with a junction gets its junction points and
    // its two end points, as an ascending set.  Per-thread tube histograms, one scatter, then
    // a sort of every tube's points.
    const size_t T1 = (size_t)n_cnts + 1;
    std::vector<int64_t> hist((size_t)nth * T1, 0);
    parallel([&](int t) {
        int64_t *h = hist.data() + (size_t)t * T1;
        for (const Junc &x : part[(size_t)t]) {
            h[x.i]++;
            h[x.j]++;
        }
        // the junctions themselves, at their place in the output
        int64_t q = joff[(size_t)t];
        for (const Junc &x : part[(size_t)t]) {
            if (q < junc_cap) {
                junc[4 * q] = x.pi;
                junc[4 * q + 1] = x.i;
                junc[4 * q + 2] = x.pj;
                junc[4 * q + 3] = x.j;
                junc_d[q] = x.d;
            }
            q++;
        }
    });
    // tube offsets: junction points + 2 end points per tube that has any
    std::vector<int64_t> ecount(T1, 0);
    for (int64_t i = 0; i < n_cnts; i++) {
        int64_t c = 0;
        for (int t = 0; t < nth; t++) c += hist[(size_t)t * T1 + (size_t)i];
        ecount[(size_t)i + 1] = ecount[(size_t)i] + (c ? c + 2 : 0);
    }
    // per-thread write positions inside each tube: thread t after threads 0..t-1
    for (int64_t i = 0; i < n_cnts; i++) {
        int64_t at = ecount[(size_t)i];
        for (int t = 0; t < nth; t++) {
            const int64_t c = hist[(size_t)t * T1 + (size_t)i];
            hist[(size_t)t * T1 + (size_t)i] = at;
            at += c;
        }
    }
    std::vector<int64_t> epts((size_t)ecount[(size_t)n_cnts]);
    parallel([&](int t) {
        int64_t *h = hist.data() + (size_t)t * T1;
        for (const Junc &x : part[(size_t)t]) {
            epts[(size_t)h[x.i]++] = x.pi;
            epts[(size_t)h[x.j]++] = x.pj;
        }
    });
    std::vector<int64_t> ecnt((size_t)n_cnts, 0);
    parallel([&](int t) {
        const int64_t i0 = n_cnts * t / nth, i1 = n_cnts * (t + 1) / nth;
        for (int64_t i = i0; i < i1; i++) {
            int64_t *a0 = epts.data() + ecount[(size_t)i], *a1 = epts.data() + ecount[(size_t)i + 1];
            if (a0 == a1) continue;
            a1[-2] = struct_pts[struct_off[i]];
            a1[-1] = struct_pts[struct_off[i + 1] - 1];
            std::sort(a0, a1);
            ecnt[(size_t)i] = std::unique(a0, a1) - a0;
        }
    });
    elem_off[0] = 0;
    for (int64_t i = 0; i < n_cnts; i++) elem_off[i + 1] = elem_off[i] + ecnt[(size_t)i];
    if (elem_pts)
        parallel([&](int t) {
            const int64_t i0 = n_cnts * t / nth, i1 = n_cnts * (t + 1) / nth;
            for (int64_t i = i0; i < i1; i++)
                std::copy(epts.data() + ecount[(size_t)i],
                          epts.data() + ecount[(size_t)i] + ecnt[(size_t)i], elem_pts + elem_off[i]);
        });
    const auto t_end = clk::now();
    if (stats) {
        auto ms = [](clk::time_point a, clk::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        stats[0] = tot_d;
        stats[1] = (double)slots;
        stats[2] = (double)ctr[1];
        stats[3] = (double)count;
        stats[4] = kernel_ms;
        stats[5] = ms(t_start, t_host);
        stats[6] = ms(t_host, t_dev);
        stats[7] = ms(t_dev, t_end);
    }
    if (nj > junc_cap) {
        set_error(ctx, "cicy_cnt_contacts: %lld junctions, room for %lld", (long long)nj,
                  (long long)junc_cap);
        return 0;
    }
    return 1;
}
