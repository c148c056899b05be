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
// Here:
//   host   the subregions, in the reference's order (an O(points) pass);
//   device k_pack gathers each slot's point into slot-ordered arrays; k_pairs runs one
//          thread per slot s, testing it against the later slots of its subregion, and
//          appends every contact (s, k, d) -- the O(sum m^2) part;
//   host   sorts the contacts by (s, k), which is the reference's visiting order because
//          slots are numbered subregion by subregion in list order, and replays HK76, the
//          closest-point map and the compression over them in that order.
// Labels, junctions and elements are therefore those of the reference, bit for bit: the
// distance is evaluated with the same IEEE operations in the same order (no contraction
// into fused multiply-adds; the FP64 square root is correctly rounded).
#include "dea_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace cicy {
namespace {

constexpr double kZeroTol = 1.0e-10;  // Zero, MathMatrix.h:18

__global__ __launch_bounds__(kThreads) void k_pack(
    int64_t slots, const int *__restrict__ slot_pt, const double *__restrict__ xyz,
    const int *__restrict__ flag, const double *__restrict__ radii, double *__restrict__ sx,
    double *__restrict__ sy, double *__restrict__ sz, double *__restrict__ sr,
    int *__restrict__ sc) {
    const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= slots) return;
    const int p = slot_pt[s];
    const int c = flag[p];
    sx[s] = xyz[3 * (int64_t)p];
    sy[s] = xyz[3 * (int64_t)p + 1];
    sz[s] = xyz[3 * (int64_t)p + 2];
    sc[s] = c;
    sr[s] = radii[c];
}

// One thread per slot: tests slot s against slots s+1 .. end[s]-1 of its subregion.
__global__ __launch_bounds__(kThreads) void k_pairs(
    int64_t slots, const int *__restrict__ end, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ sr,
    const int *__restrict__ sc, double vdw, double tunnel, unsigned long long cap,
    unsigned long long *__restrict__ count, int2 *__restrict__ pairs, double *__restrict__ dj) {
#pragma clang fp contract(off)
    const int64_t s = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (s >= slots) return;
    const int e = end[s];
    const double x = sx[s], y = sy[s], z = sz[s], r = sr[s];
    const int c = sc[s];
    for (int k = (int)s + 1; k < e; k++) {
        if (sc[k] == c) continue;
        const double dx = x - sx[k];
        const double dy = y - sy[k];
        const double dz = z - sz[k];
        const double dist = sqrt(dx * dx + dy * dy + dz * dz);
        double d = dist - r - sr[k];
        if (d < vdw) d = vdw;
        if (d <= tunnel) {
            const unsigned long long i = atomicAdd(count, 1ull);
            if (i < cap) {
                pairs[i] = make_int2((int)s, k);
                dj[i] = d;
            }
        }
    }
}

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

// Calculate_region_coordinates + Calculate_overlapping_flags (Contact_grid.cpp:163-212):
// home cell co[3] and the neighbour offset f[3] in {-1, 0, 1} per axis.
void place(const Grid &g, const double *p, int co[3], int f[3]) {
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
}

int find_root(int L, const std::vector<int> &ll) {
    int r = L;
    while (ll[(size_t)r] <= 0) {
        r = -ll[(size_t)r];
        if (ll[(size_t)r] == 0) return 0;
    }
    return r;
}

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
    // the tube of every point (the reference's Point_3D::flag); a point may be in one tube
    std::vector<int> flag((size_t)std::max<int64_t>(n_points, 1), -1);
    if (struct_off[0] != 0) {
        set_error(ctx, "cicy_cnt_contacts: structure offsets must start at 0");
        return 0;
    }
    for (int64_t i = 0; i < n_cnts; i++) {
        if (struct_off[i + 1] < struct_off[i]) {
            set_error(ctx, "cicy_cnt_contacts: structure offsets decrease at tube %lld",
                      (long long)i);
            return 0;
        }
        for (int64_t q = struct_off[i]; q < struct_off[i + 1]; q++) {
            const int64_t P = struct_pts[q];
            if (P < 0 || P >= n_points || flag[(size_t)P] != -1) {
                set_error(ctx, "cicy_cnt_contacts: point %lld of tube %lld is out of range "
                          "or in two tubes", (long long)P, (long long)i);
                return 0;
            }
            flag[(size_t)P] = (int)i;
        }
    }
    for (int64_t q = 0; q < 3 * n_points; q++)
        if (!std::isfinite(xyz[q])) {
            set_error(ctx, "cicy_cnt_contacts: point %lld has a coordinate that is not finite",
                      (long long)(q / 3));
            return 0;
        }
    const Grid g = make_grid(geom);
    const double tot_d = (double)g.n[0] * g.n[1] * g.n[2];
    if (tot_d > 1.0e8) {
        set_error(ctx, "cicy_cnt_contacts: %.0f subregions (at most 1e8)", tot_d);
        return 0;
    }
    const int64_t tot = (int64_t)tot_d;
    // Fill_sectioned_domain_cnts (Contact_grid.cpp:104-161): counting pass, then placement in
    // the same visiting order, so each subregion lists its points as the reference's does
    std::vector<int64_t> roff((size_t)tot + 1, 0);
    auto visit = [&](auto &&emit) {
        for (int64_t i = 0; i < n_inside; i++) {
            const int64_t cnt = cnts_inside[i];
            for (int64_t q = struct_off[cnt]; q < struct_off[cnt + 1]; q++) {
                const int64_t P = struct_pts[q];
                int co[3], f[3];
                place(g, xyz + 3 * P, co, f);
                const int64_t sxy = (int64_t)g.n[0] * g.n[1];
                const int64_t t = co[0] + (int64_t)co[1] * g.n[0] + co[2] * sxy;
                emit(t, P);
                if (f[0]) emit(t + f[0], P);
                if (f[1]) emit(t + (int64_t)f[1] * g.n[0], P);
                if (f[2]) emit(t + f[2] * sxy, P);
            }
        }
    };
    for (int64_t i = 0; i < n_inside; i++)
        if (cnts_inside[i] < 0 || cnts_inside[i] >= n_cnts) {
            set_error(ctx, "cicy_cnt_contacts: cnts_inside[%lld] = %lld out of range",
                      (long long)i, (long long)cnts_inside[i]);
            return 0;
        }
    visit([&](int64_t t, int64_t) { roff[(size_t)t + 1]++; });
    for (int64_t t = 0; t < tot; t++) roff[(size_t)t + 1] += roff[(size_t)t];
    const int64_t slots = roff[(size_t)tot];
    if (slots >= (1LL << 31)) {
        set_error(ctx, "cicy_cnt_contacts: %lld subregion slots (at most 2^31)",
                  (long long)slots);
        return 0;
    }
    std::vector<int> slot_pt((size_t)std::max<int64_t>(slots, 1)), send(slot_pt.size());
    {
        std::vector<int64_t> fill(roff.begin(), roff.end() - 1);
        visit([&](int64_t t, int64_t P) { slot_pt[(size_t)fill[(size_t)t]++] = (int)P; });
        for (int64_t t = 0; t < tot; t++)
            for (int64_t s = roff[(size_t)t]; s < roff[(size_t)t + 1]; s++)
                send[(size_t)s] = (int)roff[(size_t)t + 1];
    }
    double pair_tests = 0.0;
    for (int64_t t = 0; t < tot; t++) {
        const double m = (double)(roff[(size_t)t + 1] - roff[(size_t)t]);
        pair_tests += 0.5 * m * (m - 1.0);
    }
    const auto t_host = clk::now();

    // ---- device: pack the slots and test every pair of each subregion
    if (hipSetDevice(ctx->device) != hipSuccess) return 0;
    hipStream_t st = ctx->stream;
    double *dxyz = nullptr, *drad = nullptr, *sx = nullptr, *sy = nullptr, *sz = nullptr,
           *sr = nullptr, *ddj = nullptr;
    int *dflag = nullptr, *dslot = nullptr, *dend = nullptr, *sc = nullptr;
    int2 *dpairs = nullptr;
    unsigned long long *dcount = nullptr;
    const size_t S = slot_pt.size();
    const size_t NP = (size_t)std::max<int64_t>(n_points, 1), NC = (size_t)std::max<int64_t>(n_cnts, 1);
    std::vector<double> rad(NC, 0.0);
    if (n_cnts > 0) std::memcpy(rad.data(), radii, (size_t)n_cnts * 8);
    std::vector<double> xyzv(3 * NP, 0.0);
    if (n_points > 0) std::memcpy(xyzv.data(), xyz, (size_t)(3 * n_points) * 8);
    for (int &f : flag)
        if (f < 0) f = 0;  // points outside every tube are never in a slot
    bool ok = hip_ok(ctx, hipMalloc((void **)&dxyz, 3 * NP * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&drad, NC * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dflag, NP * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dslot, S * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dend, S * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&sx, S * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&sy, S * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&sz, S * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&sr, S * 8), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&sc, S * 4), "hipMalloc") &&
              hip_ok(ctx, hipMalloc((void **)&dcount, 8), "hipMalloc");
    ok = ok &&
         hip_ok(ctx, hipMemcpyAsync(dxyz, xyzv.data(), 3 * NP * 8, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(drad, rad.data(), NC * 8, hipMemcpyHostToDevice, st), "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dflag, flag.data(), NP * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dslot, slot_pt.data(), S * 4, hipMemcpyHostToDevice, st),
                "H2D") &&
         hip_ok(ctx, hipMemcpyAsync(dend, send.data(), S * 4, hipMemcpyHostToDevice, st), "H2D");
    const unsigned blocks = (unsigned)((std::max<int64_t>(slots, 1) + kThreads - 1) / kThreads);
    if (ok && slots > 0)
        hipLaunchKernelGGL(k_pack, dim3(blocks), dim3(kThreads), 0, st, (int64_t)slots, dslot,
                           dxyz, dflag, drad, sx, sy, sz, sr, sc);
    // contacts: a first guess of the room, then once more with the exact count if short
    unsigned long long cap = (unsigned long long)std::max<int64_t>(4 * slots, 1024);
    unsigned long long count = 0;
    float kernel_ms = 0.f;
    for (int attempt = 0; ok && attempt < 2; attempt++) {
        if (dpairs) (void)hipFree(dpairs);
        if (ddj) (void)hipFree(ddj);
        dpairs = nullptr;
        ddj = nullptr;
        ok = hip_ok(ctx, hipMalloc((void **)&dpairs, cap * sizeof(int2)), "hipMalloc") &&
             hip_ok(ctx, hipMalloc((void **)&ddj, cap * 8), "hipMalloc") &&
             hip_ok(ctx, hipMemsetAsync(dcount, 0, 8, st), "memset") &&
             hip_ok(ctx, hipEventRecord(ctx->ev_start, st), "event");
        if (ok && slots > 0)
            hipLaunchKernelGGL(k_pairs, dim3(blocks), dim3(kThreads), 0, st, (int64_t)slots,
                               dend, sx, sy, sz, sr, sc, vdw, tunnel, cap, dcount, dpairs, ddj);
        ok = ok && hip_ok(ctx, hipGetLastError(), "k_pairs launch") &&
             hip_ok(ctx, hipEventRecord(ctx->ev_stop, st), "event") &&
             hip_ok(ctx, hipMemcpyAsync(&count, dcount, 8, hipMemcpyDeviceToHost, st), "D2H") &&
             hip_ok(ctx, hipStreamSynchronize(st), "sync") &&
             hip_ok(ctx, hipEventElapsedTime(&kernel_ms, ctx->ev_start, ctx->ev_stop), "event");
        if (!ok || count <= cap) break;
        cap = count;
    }
    std::vector<int2> pairs;
    std::vector<double> dj;
    if (ok) {
        pairs.resize((size_t)count);
        dj.resize((size_t)count);
        ok = (count == 0 ||
              (hip_ok(ctx, hipMemcpyAsync(pairs.data(), dpairs, count * sizeof(int2),
                                          hipMemcpyDeviceToHost, st), "D2H") &&
               hip_ok(ctx, hipMemcpyAsync(dj.data(), ddj, count * 8, hipMemcpyDeviceToHost, st),
                      "D2H"))) &&
             hip_ok(ctx, hipStreamSynchronize(st), "sync");
    }
    void *ptrs[] = {dxyz, drad, dflag, dslot, dend, sx, sy, sz, sr, sc, dcount, dpairs, ddj};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (!ok) return 0;
    const auto t_dev = clk::now();

    // ---- host: the reference's visiting order, then its sequential bookkeeping
    std::vector<uint32_t> order((size_t)count);
    std::iota(order.begin(), order.end(), 0u);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const int2 pa = pairs[a], pb = pairs[b];
        return pa.x != pb.x ? pa.x < pb.x : pa.y < pb.y;
    });
    std::vector<int> lab((size_t)n_cnts, -1), ll;
    int new_label = 0;
    struct Tri {
        int a, b;
        int64_t pa;
    };
    std::vector<Tri> ce;
    ce.reserve((size_t)count);
    std::unordered_map<uint64_t, std::pair<double, int64_t>> pc;
    pc.reserve((size_t)count);
    for (uint32_t id : order) {
        const int64_t P1 = slot_pt[(size_t)pairs[id].x], P2 = slot_pt[(size_t)pairs[id].y];
        const int c1 = flag[(size_t)P1], c2 = flag[(size_t)P2];
        const double d = dj[id];
        // HK76 (Hoshen_Kopelman.cpp:2908-2953)
        if (lab[(size_t)c1] == -1 && lab[(size_t)c2] == -1) {
            lab[(size_t)c1] = lab[(size_t)c2] = new_label++;
            ll.push_back(2);
        } else if (lab[(size_t)c1] == -1 || lab[(size_t)c2] == -1) {
            const int known = lab[(size_t)c1] == -1 ? c2 : c1;
            const int fresh = known == c2 ? c1 : c2;
            int L = lab[(size_t)known];
            if (ll[(size_t)L] <= 0) L = find_root(L, ll);
            lab[(size_t)fresh] = L;
            ll[(size_t)L] += 1;
        } else if (lab[(size_t)c1] != lab[(size_t)c2]) {
            const int r1 = find_root(lab[(size_t)c1], ll), r2 = find_root(lab[(size_t)c2], ll);
            if (ll[(size_t)r1] <= 0 || ll[(size_t)r2] <= 0) {
                set_error(ctx, "cicy_cnt_contacts: merging labels that are not roots");
                return 0;
            }
            if (r1 < r2) {
                ll[(size_t)r1] += ll[(size_t)r2];
                ll[(size_t)r2] = -r1;
            } else if (r2 < r1) {
                ll[(size_t)r2] += ll[(size_t)r1];
                ll[(size_t)r1] = -r2;
            }
        }
        const int a = std::min(c1, c2), b = std::max(c1, c2);
        const int64_t Pa = c1 < c2 ? P1 : P2, Pb = c2 < c1 ? P1 : P2;
        ce.push_back({a, b, Pa});
        const uint64_t key = (uint64_t)Pa * (uint64_t)n_cnts + (uint64_t)b;
        auto it = pc.find(key);
        if (it == pc.end())
            pc.emplace(key, std::make_pair(d, Pb));
        else if (it->second.first > d)
            it->second = std::make_pair(d, Pb);
    }
    // Cleanup_labels (:374-425)
    std::vector<int> label_map(ll.size(), -1);
    int proper = 0;
    for (size_t i = 0; i < ll.size(); i++)
        if (ll[i] > 0) label_map[i] = proper++;
    for (int64_t i = 0; i < n_cnts; i++) {
        int L = lab[(size_t)i];
        if (L != -1) L = label_map[(size_t)find_root(L, ll)];
        labels[i] = L;
    }
    *n_labels = proper;
    // Compress_cnt_cnt_contact_segments (:443-555): tube i ascending, then the other tube,
    // then the points of tube i in contact with it, ascending
    std::sort(ce.begin(), ce.end(), [](const Tri &x, const Tri &y) {
        return x.a != y.a ? x.a < y.a : (x.b != y.b ? x.b < y.b : x.pa < y.pa);
    });
    ce.erase(std::unique(ce.begin(), ce.end(),
                         [](const Tri &x, const Tri &y) {
                             return x.a == y.a && x.b == y.b && x.pa == y.pa;
                         }),
             ce.end());
    std::vector<std::vector<int64_t>> el((size_t)n_cnts);
    int64_t nj = 0;
    auto emit = [&](int64_t pi, int i, int64_t pj, int j, double d) {
        if (nj < junc_cap) {
            junc[4 * nj] = pi;
            junc[4 * nj + 1] = i;
            junc[4 * nj + 2] = pj;
            junc[4 * nj + 3] = j;
            junc_d[nj] = d;
        }
        nj++;
        el[(size_t)i].push_back(pi);
        el[(size_t)j].push_back(pj);
    };
    auto look = [&](int64_t pa, int b) -> const std::pair<double, int64_t> & {
        return pc.at((uint64_t)pa * (uint64_t)n_cnts + (uint64_t)b);
    };
    for (size_t q = 0; q < ce.size();) {
        size_t e = q;
        while (e < ce.size() && ce[e].a == ce[q].a && ce[e].b == ce[q].b) e++;
        const int i = ce[q].a, cj = ce[q].b;
        int64_t Pi0 = ce[q].pa;
        int64_t Pj0 = look(Pi0, cj).second;
        double dmin = look(Pi0, cj).first;
        int64_t Pi_j = Pi0, Pj_j = Pj0;
        for (size_t u = q + 1; u < e; u++) {
            const int64_t Pi1 = ce[u].pa;
            const auto &v = look(Pi1, cj);
            const int64_t Pj1 = v.second;
            const double d = v.first;
            if (Pi1 - Pi0 >= min_points || std::llabs(Pj1 - Pj0) >= min_points) {
                emit(Pi_j, i, Pj_j, cj, dmin);
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
        emit(Pi_j, i, Pj_j, cj, dmin);
        q = e;
    }
    *n_junc = nj;
    // Complete_cnt_elements (:557-573), element sets in ascending point order
    elem_off[0] = 0;
    for (int64_t i = 0; i < n_cnts; i++) {
        std::vector<int64_t> &v = el[(size_t)i];
        if (!v.empty()) {
            v.push_back(struct_pts[struct_off[i]]);
            v.push_back(struct_pts[struct_off[i + 1] - 1]);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
        }
        elem_off[i + 1] = elem_off[i] + (int64_t)v.size();
        if (elem_pts) std::copy(v.begin(), v.end(), elem_pts + elem_off[i]);
    }
    const auto t_end = clk::now();
    if (stats) {
        auto ms = [](clk::time_point a, clk::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        stats[0] = (double)tot;
        stats[1] = (double)slots;
        stats[2] = pair_tests;
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
