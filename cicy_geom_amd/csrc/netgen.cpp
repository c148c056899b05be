// netgen.cpp -- seeded straight-tube CNT networks and the DEA system assembly.
//
// Generation (the workload's own generator, not the reference's wavy-tube grower):
// tubes with uniform centres, isotropic directions and U[len_min,len_max] lengths are
// clipped to the cube, discretised into seg_len polylines, and tested pairwise (cell
// list) with the segment-segment distance of their axes.  A pair is in contact when
// d = dist - r1 - r2, clamped below at d_vdw, is <= d_tunnel, the contact rule of
// Hoshen_Kopelman::Label_cnts_in_window (Hoshen_Kopelman.cpp:281-307); one junction per
// tube pair sits at the polyline points nearest the closest approach.  Clipped ends are
// boundary points, classified with the face priority of
// Cutoff_Wins::Where_is_with_boundary (Cutoff_Wins.cpp:716-739): x_min->4, x_max->2,
// y_min->5, y_max->3, z_min->1, z_max->0.
//
// Assembly follows Direct_Electrifying.cpp:
//   Get_global_nodes            :115-137   reserved nodes from the family
//   Get_vector_of_boundaries    :238-305   boundary order -> reserved node index
//   Map_points_at_boundaries    :217-236
//   LM_matrix_for_cnts          :139-215   first/interior/last element points -> nodes
//   Fill_2d_matrices_cnts       :489-553   series elements, Three_points_in_element :555-608
//   Calculate_resistance_cnt    :611-646
//   Add_new_elements_to_2d_sparse_matrix        :648-668  (off-diagonal overwritten)
//   Add_to_existing_elements_in_2d_sparse_matrix :670-694 (off-diagonal accumulated)
//   Fill_2d_matrices_cnt_junctions :738-852   junction drop rule for R_d == 2
//   Calculate_junction_resistance  :854-901
//   From_2d_to_1d_vectors       :1329-1394 SSS arrays, b = -K_EF^T VEF, reserved diag erased
//   zero-diagonal check         :458-484
//   Initial_guess_for_CG / Initial_guess_for_cnt_nodes :1396-1470 (ramp)
//   Update_residual_vector      :1502-1535
// and the boundary element rule of Electrical_analysis.cpp:499-605 for the current sums.
// The per-row std::map of the reference is replaced by an insertion-ordered triplet list
// that is stably sorted by (row, col) and folded with the same overwrite/accumulate
// semantics, so values, diagonal and b are identical bit for bit.
#include "netgen.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kZero = 1.0e-10;

struct Junction {
    int64_t p1, p2;   // points
    int64_t n1, n2;   // tubes (n1 < n2)
    double dist;      // gap (um)
};

struct Tube {
    double a[3];      // clipped start
    double u[3];      // unit direction
    double len;       // clipped length
    double step;      // polyline spacing
};

}  // namespace

struct cicy_net {
    cicy_net_params prm;
    std::vector<double> xyz;            // points*3
    std::vector<int64_t> cnt_ptr;       // tubes+1
    std::vector<double> radius;         // tubes
    std::vector<Tube> tubes;
    std::vector<int64_t> bpts[6];       // boundary_cnt_pts
    std::vector<int64_t> bcnt[6];       // boundary_cnt
    std::vector<Junction> junctions;
    std::vector<std::vector<int64_t>> clusters;          // tube ids, ascending
    std::vector<std::vector<int64_t>> cluster_junctions; // junction ids, ascending
    std::vector<int> family;
    std::vector<int64_t> elem_ptr;      // tubes+1
    std::vector<int64_t> elem_pts;      // element points of each tube, ascending
};

struct cicy_sss_holder {
    cicy_sss_view v;
    std::vector<int64_t> row_ptr, col_ind, lmm, node_point;
    std::vector<double> values, diag, b, vef, x0, r0;
    std::vector<int64_t> e_bnd[2], e_int[2];
    std::vector<double> e_re[2];
};

namespace {

inline double u01(std::mt19937_64 &g) { return (double)(g() >> 11) * 0x1.0p-53; }

// Face of a point on the cube surface, in the priority order of
// Cutoff_Wins::Where_is_with_boundary; -1 when inside.
int face_of(const double *p, double box) {
    if (std::fabs(p[0] - 0.0) < kZero) return 4;
    if (std::fabs(p[0] - box) < kZero) return 2;
    if (std::fabs(p[1] - 0.0) < kZero) return 5;
    if (std::fabs(p[1] - box) < kZero) return 3;
    if (std::fabs(p[2] - 0.0) < kZero) return 1;
    if (std::fabs(p[2] - box) < kZero) return 0;
    return -1;
}

// Closest points of segments P0+s*d1 (s in [0,1]) and Q0+t*d2 (t in [0,1]); returns the
// squared distance.  Derived from minimising |P(s)-Q(t)|^2 over the unit square.
double seg_seg(const double *p0, const double *d1, const double *q0, const double *d2,
               double &s, double &t) {
    double r[3] = {p0[0] - q0[0], p0[1] - q0[1], p0[2] - q0[2]};
    double a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    double f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
    double c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
    double bb = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    const double eps = 1e-30;
    if (a <= eps && e <= eps) {
        s = t = 0.0;
    } else if (a <= eps) {
        s = 0.0;
        t = std::clamp(f / e, 0.0, 1.0);
    } else if (e <= eps) {
        t = 0.0;
        s = std::clamp(-c / a, 0.0, 1.0);
    } else {
        double denom = a * e - bb * bb;
        s = denom > 1e-14 * a * e ? std::clamp((bb * f - c * e) / denom, 0.0, 1.0) : 0.0;
        t = (bb * s + f) / e;
        if (t < 0.0) {
            t = 0.0;
            s = std::clamp(-c / a, 0.0, 1.0);
        } else if (t > 1.0) {
            t = 1.0;
            s = std::clamp((bb - c) / a, 0.0, 1.0);
        }
    }
    double dx = p0[0] + d1[0] * s - (q0[0] + d2[0] * t);
    double dy = p0[1] + d1[1] * s - (q0[1] + d2[1] * t);
    double dz = p0[2] + d1[2] * s - (q0[2] + d2[2] * t);
    return dx * dx + dy * dy + dz * dz;
}

// Index (within the tube) of the polyline point nearest arc length `s_len`; points sit
// at k*step, the last one at len.
int64_t nearest_point(int64_t npts, double step, double len, double s_len) {
    auto pos = [&](int64_t k) { return k == npts - 1 ? len : (double)k * step; };
    int64_t k0 = (int64_t)std::floor(s_len / step);
    k0 = std::clamp<int64_t>(k0, 0, npts - 1);
    int64_t best = k0;
    double bd = std::fabs(pos(k0) - s_len);
    for (int64_t k = k0 - 1; k <= k0 + 1; k++) {
        if (k < 0 || k >= npts) continue;
        double d = std::fabs(pos(k) - s_len);
        if (d < bd) { bd = d; best = k; }
    }
    return best;
}

int64_t find_root(std::vector<int64_t> &par, int64_t x) {
    while (par[x] != x) {
        par[x] = par[par[x]];
        x = par[x];
    }
    return x;
}

// Get_global_nodes (Direct_Electrifying.cpp:115-137)
int reserved_from_family(int family) {
    if (family == 6) return 6;
    if (family >= 3 && family <= 5) return 4;
    if (family >= 0 && family <= 2) return 2;
    return -1;
}

// Get_vector_of_boundaries (Direct_Electrifying.cpp:238-305)
std::vector<int> boundaries_of_family(int family) {
    switch (family) {
        case 0: return {2, 4};
        case 1: return {3, 5};
        case 2: return {0, 1};
        case 3: return {2, 4, 3, 5};
        case 4: return {2, 4, 0, 1};
        case 5: return {3, 5, 0, 1};
        case 6: return {2, 4, 3, 5, 0, 1};
        default: return {};
    }
}

double resistance_cnt(const cicy_net *net, int64_t p1, int64_t p2, double radius,
                      double rho) {
    // Calculate_resistance_cnt (Direct_Electrifying.cpp:611-646), R_flag == 1 branch
    const double *x = net->xyz.data();
    double length = 0;
    for (int64_t i = p1; i < p2; i++) {
        double dx = x[3 * i] - x[3 * (i + 1)];
        double dy = x[3 * i + 1] - x[3 * (i + 1) + 1];
        double dz = x[3 * i + 2] - x[3 * (i + 1) + 2];
        length = length + std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return rho * length / (kPi * radius * radius);
}

double junction_resistance(const Junction &j, double l1, double l2,
                           const cicy_elec_params *ep) {
    // Calculate_junction_resistance (Direct_Electrifying.cpp:854-901); both sides CNTs
    if (!ep->junction_exponential) return ep->junction_resistance;
    double jl1 = l1 + l1, jl2 = l2 + l2;
    double a = (jl1 - jl2 < kZero) ? jl1 * jl1 : jl2 * jl2;
    return ep->c1 * j.dist * std::exp(ep->c2 * j.dist) / a;
}

struct Trip {
    int64_t row, col;
    double val;
    int add;  // 0 overwrite (= -g), 1 accumulate (-= g)
};

}  // namespace

extern "C" {

void cicy_tunnel_constants(double h, double e, double m, double lambda, double *c1,
                           double *c2) {
    double s = std::sqrt(2.0 * m * lambda * e);
    *c1 = 10.0 * h * h / (m * m * s);
    *c2 = 4000.0 * kPi * s / h;
}

cicy_net *cicy_net_generate(const cicy_net_params *p) {
    if (!p || p->box <= 0 || p->n_cnt < 0 || p->seg_len <= 0 || p->radius <= 0 ||
        p->len_min <= 0 || p->len_max < p->len_min)
        return nullptr;
    cicy_net *net = new cicy_net();
    net->prm = *p;
    const double box = p->box, seg = p->seg_len;
    std::mt19937_64 g(p->seed);
    net->cnt_ptr.push_back(0);
    int64_t made = 0;
    while (made < p->n_cnt) {
        double c[3] = {u01(g) * box, u01(g) * box, u01(g) * box};
        double cth = 2.0 * u01(g) - 1.0;
        double phi = 2.0 * kPi * u01(g);
        double L = p->len_min + (p->len_max - p->len_min) * u01(g);
        double sth = std::sqrt(std::max(0.0, 1.0 - cth * cth));
        double u[3] = {sth * std::cos(phi), sth * std::sin(phi), cth};
        double tlo = -0.5 * L, thi = 0.5 * L;
        int axlo = -1, axhi = -1;
        for (int a = 0; a < 3; a++) {
            if (std::fabs(u[a]) < 1e-300) continue;
            double t0 = (0.0 - c[a]) / u[a], t1 = (box - c[a]) / u[a];
            if (t0 > t1) std::swap(t0, t1);
            if (t0 > tlo) { tlo = t0; axlo = a; }
            if (t1 < thi) { thi = t1; axhi = a; }
        }
        double len = thi - tlo;
        if (len < 1e-6) continue;  // degenerate clip (centre on a corner): resample
        Tube tb;
        for (int a = 0; a < 3; a++) {
            tb.a[a] = c[a] + tlo * u[a];
            tb.u[a] = u[a];
        }
        double bend[3];
        for (int a = 0; a < 3; a++) bend[a] = c[a] + thi * u[a];
        // snap clipped coordinates exactly onto their faces
        if (axlo >= 0) tb.a[axlo] = (u[axlo] > 0) ? 0.0 : box;
        if (axhi >= 0) bend[axhi] = (u[axhi] > 0) ? box : 0.0;
        tb.len = len;
        // uniform spacing as close to seg_len as the clipped length allows
        int64_t nseg = (int64_t)std::llround(len / seg);
        if (nseg < 1) nseg = 1;
        tb.step = len / (double)nseg;
        int64_t first = (int64_t)net->xyz.size() / 3;
        for (int64_t k = 0; k < nseg; k++) {
            for (int a = 0; a < 3; a++)
                net->xyz.push_back(k == 0 ? tb.a[a] : tb.a[a] + (double)k * tb.step * u[a]);
        }
        for (int a = 0; a < 3; a++) net->xyz.push_back(bend[a]);
        int64_t last = (int64_t)net->xyz.size() / 3 - 1;
        net->tubes.push_back(tb);
        net->radius.push_back(p->radius);
        net->cnt_ptr.push_back(last + 1);
        // boundary points (clipped ends); Add_cnt_point_to_boundary_vectors keeps a tube
        // once per face list (Cutoff_Wins.cpp:745-769)
        int64_t ends[2] = {first, last};
        bool clipped[2] = {axlo >= 0, axhi >= 0};
        for (int e = 0; e < 2; e++) {
            if (!clipped[e]) continue;
            int f = face_of(&net->xyz[3 * ends[e]], box);
            if (f < 0) continue;
            if (net->bcnt[f].empty() || net->bcnt[f].back() != made)
                net->bcnt[f].push_back(made);
            net->bpts[f].push_back(ends[e]);
        }
        made++;
    }
    const int64_t T = made;

    // ---- contact search: cell list over polyline segments, expanded by half the
    // contact distance, so any two axes closer than the contact distance share a cell
    const double dcontact = 2.0 * p->radius + p->d_tunnel;
    double cs = std::max(0.05, 2.0 * dcontact);
    int64_t nc = std::max<int64_t>(1, (int64_t)std::floor(box / cs));
    cs = box / (double)nc;
    auto cell_of = [&](double v) {
        int64_t i = (int64_t)std::floor(v / cs);
        return std::clamp<int64_t>(i, 0, nc - 1);
    };
    std::vector<std::pair<int64_t, int32_t>> cell_tube;  // (cell, tube)
    cell_tube.reserve((size_t)T * 64);
    std::vector<int64_t> seen;
    for (int64_t t = 0; t < T; t++) {
        seen.clear();
        for (int64_t q = net->cnt_ptr[t]; q < net->cnt_ptr[t + 1] - 1; q++) {
            const double *a = &net->xyz[3 * q], *b = &net->xyz[3 * (q + 1)];
            int64_t lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                lo[k] = cell_of(std::min(a[k], b[k]) - 0.5 * dcontact);
                hi[k] = cell_of(std::max(a[k], b[k]) + 0.5 * dcontact);
            }
            for (int64_t i = lo[0]; i <= hi[0]; i++)
                for (int64_t j = lo[1]; j <= hi[1]; j++)
                    for (int64_t k = lo[2]; k <= hi[2]; k++)
                        seen.push_back((i * nc + j) * nc + k);
        }
        std::sort(seen.begin(), seen.end());
        seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
        for (int64_t cidx : seen) cell_tube.emplace_back(cidx, (int32_t)t);
    }
    std::sort(cell_tube.begin(), cell_tube.end());
    std::vector<uint64_t> pairs;
    for (size_t s = 0; s < cell_tube.size();) {
        size_t e = s;
        while (e < cell_tube.size() && cell_tube[e].first == cell_tube[s].first) e++;
        for (size_t x = s; x < e; x++)
            for (size_t y = x + 1; y < e; y++)
                pairs.push_back(((uint64_t)cell_tube[x].second << 32) |
                                (uint64_t)cell_tube[y].second);
        s = e;
    }
    std::vector<std::pair<int64_t, int32_t>>().swap(cell_tube);
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    for (uint64_t pr : pairs) {
        int64_t i = (int64_t)(pr >> 32), j = (int64_t)(pr & 0xffffffffu);
        const Tube &ti = net->tubes[i], &tj = net->tubes[j];
        double d1[3], d2[3];
        for (int a = 0; a < 3; a++) {
            d1[a] = ti.u[a] * ti.len;
            d2[a] = tj.u[a] * tj.len;
        }
        double s, t;
        double dist = std::sqrt(seg_seg(ti.a, d1, tj.a, d2, s, t));
        double dj = dist - net->radius[i] - net->radius[j];
        if (dj < p->d_vdw) dj = p->d_vdw;
        if (dj > p->d_tunnel) continue;
        int64_t ni = net->cnt_ptr[i + 1] - net->cnt_ptr[i];
        int64_t nj = net->cnt_ptr[j + 1] - net->cnt_ptr[j];
        Junction J;
        J.p1 = net->cnt_ptr[i] + nearest_point(ni, ti.step, ti.len, s * ti.len);
        J.p2 = net->cnt_ptr[j] + nearest_point(nj, tj.step, tj.len, t * tj.len);
        J.n1 = i;
        J.n2 = j;
        J.dist = dj;
        net->junctions.push_back(J);
    }

    // ---- clusters: union-find keeping the smaller root (as Merge_labels does,
    // Hoshen_Kopelman.cpp:2977-2985); clusters need >= 2 tubes (isolated tubes are
    // never clusters)
    std::vector<int64_t> par(T);
    std::iota(par.begin(), par.end(), 0);
    std::vector<int64_t> deg(T, 0);
    for (const Junction &J : net->junctions) {
        int64_t a = find_root(par, J.n1), b = find_root(par, J.n2);
        if (a != b) {
            if (a < b) par[b] = a; else par[a] = b;
        }
        deg[J.n1]++;
        deg[J.n2]++;
    }
    std::vector<int64_t> cid(T, -1);
    for (int64_t t = 0; t < T; t++) {
        if (!deg[t]) continue;
        int64_t r = find_root(par, t);
        if (cid[r] < 0) {
            cid[r] = (int64_t)net->clusters.size();
            net->clusters.emplace_back();
        }
        cid[t] = cid[r];
        net->clusters[cid[t]].push_back(t);
    }
    net->cluster_junctions.resize(net->clusters.size());
    for (int64_t jx = 0; jx < (int64_t)net->junctions.size(); jx++)
        net->cluster_junctions[cid[find_root(par, net->junctions[jx].n1)]].push_back(jx);
    // percolation families (0 X, 1 Y, 2 Z, 3 XY, 4 XZ, 5 YZ, 6 XYZ)
    std::vector<uint8_t> touch(T, 0);
    for (int f = 0; f < 6; f++)
        for (int64_t t : net->bcnt[f]) touch[t] |= (uint8_t)(1u << f);
    for (auto &cl : net->clusters) {
        uint8_t m = 0;
        for (int64_t t : cl) m |= touch[t];
        bool px = (m & (1 << 2)) && (m & (1 << 4));
        bool py = (m & (1 << 3)) && (m & (1 << 5));
        bool pz = (m & (1 << 0)) && (m & (1 << 1));
        int fam = -1;
        if (px && py && pz) fam = 6;
        else if (px && py) fam = 3;
        else if (px && pz) fam = 4;
        else if (py && pz) fam = 5;
        else if (px) fam = 0;
        else if (py) fam = 1;
        else if (pz) fam = 2;
        net->family.push_back(fam);
    }

    // ---- element points
    std::vector<std::vector<int64_t>> extra;
    if (!p->per_point_nodes) {
        extra.resize(T);
        for (const Junction &J : net->junctions) {
            extra[J.n1].push_back(J.p1);
            extra[J.n2].push_back(J.p2);
        }
    }
    net->elem_ptr.push_back(0);
    for (int64_t t = 0; t < T; t++) {
        int64_t a = net->cnt_ptr[t], b = net->cnt_ptr[t + 1];
        if (p->per_point_nodes) {
            for (int64_t q = a; q < b; q++) net->elem_pts.push_back(q);
        } else {
            std::vector<int64_t> e = extra[t];
            e.push_back(a);
            e.push_back(b - 1);
            std::sort(e.begin(), e.end());
            e.erase(std::unique(e.begin(), e.end()), e.end());
            net->elem_pts.insert(net->elem_pts.end(), e.begin(), e.end());
        }
        net->elem_ptr.push_back((int64_t)net->elem_pts.size());
    }
    return net;
}

void cicy_net_free(cicy_net *net) { delete net; }

int64_t cicy_net_count(const cicy_net *net, int what) {
    if (!net) return -1;
    switch (what) {
        case 0: return (int64_t)net->xyz.size() / 3;
        case 1: return (int64_t)net->radius.size();
        case 2: return (int64_t)net->junctions.size();
        case 3: return (int64_t)net->clusters.size();
        default: return -1;
    }
}

const double *cicy_net_xyz(const cicy_net *net) { return net ? net->xyz.data() : nullptr; }
const int64_t *cicy_net_cnt_ptr(const cicy_net *net) {
    return net ? net->cnt_ptr.data() : nullptr;
}

int cicy_net_cluster(const cicy_net *net, int64_t k, int64_t *n_tubes, int *family) {
    if (!net || k < 0 || k >= (int64_t)net->clusters.size()) return 0;
    if (n_tubes) *n_tubes = (int64_t)net->clusters[k].size();
    if (family) *family = net->family[k];
    return 1;
}

int cicy_net_junction(const cicy_net *net, int64_t j, int64_t *p1, int64_t *p2,
                      int64_t *n1, int64_t *n2, double *dist) {
    if (!net || j < 0 || j >= (int64_t)net->junctions.size()) return 0;
    const Junction &J = net->junctions[j];
    *p1 = J.p1; *p2 = J.p2; *n1 = J.n1; *n2 = J.n2; *dist = J.dist;
    return 1;
}

int64_t cicy_net_boundary_points(const cicy_net *net, int b, const int64_t **pts) {
    if (!net || b < 0 || b > 5) return -1;
    *pts = net->bpts[b].data();
    return (int64_t)net->bpts[b].size();
}

int64_t cicy_net_boundary_tubes(const cicy_net *net, int b, const int64_t **tubes) {
    if (!net || b < 0 || b > 5) return -1;
    *tubes = net->bcnt[b].data();
    return (int64_t)net->bcnt[b].size();
}

int64_t cicy_net_cluster_tubes(const cicy_net *net, int64_t k, const int64_t **tubes) {
    if (!net || k < 0 || k >= (int64_t)net->clusters.size()) return -1;
    *tubes = net->clusters[k].data();
    return (int64_t)net->clusters[k].size();
}

int64_t cicy_net_cluster_junctions(const cicy_net *net, int64_t k, const int64_t **junc) {
    if (!net || k < 0 || k >= (int64_t)net->cluster_junctions.size()) return -1;
    *junc = net->cluster_junctions[k].data();
    return (int64_t)net->cluster_junctions[k].size();
}

int64_t cicy_net_elements(const cicy_net *net, int64_t t, const int64_t **pts) {
    if (!net || t < 0 || t + 1 >= (int64_t)net->elem_ptr.size()) return -1;
    *pts = net->elem_pts.data() + net->elem_ptr[t];
    return net->elem_ptr[t + 1] - net->elem_ptr[t];
}

cicy_sss_view *cicy_net_assemble(cicy_net *net, int64_t k, int r_flag, int family,
                                 const cicy_elec_params *ep, int direction, char *err,
                                 int64_t err_len) {
    auto fail = [&](const char *msg) -> cicy_sss_view * {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg);
        return nullptr;
    };
    if (!net || !ep || k < 0 || k >= (int64_t)net->clusters.size())
        return fail("invalid network or cluster index");
    if (r_flag != 0 && r_flag != 1) return fail("invalid R_flag (valid: 0, 1)");
    int rd = reserved_from_family(family);
    if (rd < 0) return fail("invalid family");
    const int64_t P = (int64_t)net->xyz.size() / 3;
    auto *h = new cicy_sss_holder();
    h->lmm.assign((size_t)P, -1);
    // Map_points_at_boundaries: later boundaries overwrite earlier ones
    std::vector<int> bnds = boundaries_of_family(family);
    for (size_t nidx = 0; nidx < bnds.size(); nidx++)
        for (int64_t q : net->bpts[bnds[nidx]]) h->lmm[q] = (int64_t)nidx;
    // LM_matrix_for_cnts
    int64_t gn = rd;
    const auto &cl = net->clusters[k];
    for (int64_t t : cl) {
        int64_t a = net->elem_ptr[t], b = net->elem_ptr[t + 1];
        if (b - a <= 1) {
            delete h;
            return fail("LM_matrix_for_cnts: a tube with not enough points to form an element");
        }
        int64_t pf = net->elem_pts[a];
        if (h->lmm[pf] < 0) h->lmm[pf] = gn++;
        for (int64_t q = a + 1; q < b - 1; q++) h->lmm[net->elem_pts[q]] = gn++;
        int64_t pl = net->elem_pts[b - 1];
        if (h->lmm[pl] < 0) h->lmm[pl] = gn++;
    }
    const int64_t nodes = gn, n = nodes - rd;
    std::vector<double> diag_full((size_t)nodes, 0.0);
    std::vector<Trip> trips;
    trips.reserve((size_t)(n + net->cluster_junctions[k].size() + 16));
    auto add_new = [&](int64_t n1, int64_t n2, double g) {
        diag_full[n1] += g;
        diag_full[n2] += g;
        if (n1 > n2) trips.push_back({n1, n2, -g, 0});
        else trips.push_back({n2, n1, -g, 0});
    };
    auto add_existing = [&](int64_t n1, int64_t n2, double g) {
        diag_full[n1] += g;
        diag_full[n2] += g;
        if (n1 > n2) trips.push_back({n1, n2, g, 1});
        else trips.push_back({n2, n1, g, 1});
    };
    // Fill_2d_matrices_cnts
    for (int64_t t : cl) {
        int64_t a = net->elem_ptr[t], b = net->elem_ptr[t + 1];
        const int64_t *e = &net->elem_pts[a];
        int64_t ne = b - a;
        if (ne == 3 && h->lmm[e[0]] == h->lmm[e[2]]) {
            // Three_points_in_element
            double g1 = 1.0, g2 = 1.0;
            if (r_flag == 1) {
                g1 = 1 / resistance_cnt(net, e[0], e[1], net->radius[t], ep->resistivity_cnt);
                g2 = 1 / resistance_cnt(net, e[1], e[2], net->radius[t], ep->resistivity_cnt);
            }
            int64_t n1 = h->lmm[e[0]], n2 = h->lmm[e[1]];
            diag_full[n1] += g1 + g2;
            diag_full[n2] += g1 + g2;
            trips.push_back({n2, n1, -(g1 + g2), 0});
            continue;
        }
        for (int64_t q = 1; q < ne; q++) {
            double re = (r_flag == 1)
                            ? resistance_cnt(net, e[q - 1], e[q], net->radius[t], ep->resistivity_cnt)
                            : 1.0;
            double g = (r_flag == 1) ? 1 / re : re;
            add_new(h->lmm[e[q - 1]], h->lmm[e[q]], g);
        }
    }
    // Fill_2d_matrices_cnt_junctions
    for (int64_t jx : net->cluster_junctions[k]) {
        const Junction &J = net->junctions[jx];
        int64_t n1 = h->lmm[J.p1], n2 = h->lmm[J.p2];
        if (n1 < 0 || n2 < 0) continue;
        double g = 1.0;
        if (r_flag == 1)
            g = 1 / junction_resistance(J, net->radius[J.n1], net->radius[J.n2], ep);
        if (n1 >= rd && n2 >= rd) add_new(n1, n2, g);
        else if (rd > 2) add_existing(n1, n2, g);
    }
    // From_2d_to_1d_vectors: stable sort keeps insertion order inside each (row, col)
    std::stable_sort(trips.begin(), trips.end(), [](const Trip &x, const Trip &y) {
        return x.row != y.row ? x.row < y.row : x.col < y.col;
    });
    h->vef.assign((size_t)rd, 0.0);
    double vmag = (r_flag == 1) ? ep->applied_voltage : (double)nodes;
    for (int i = 1; i < rd; i++) h->vef[i] = h->vef[i - 1] + vmag;
    h->b.assign((size_t)n, 0.0);
    h->row_ptr.assign((size_t)n + 1, 0);
    size_t s = 0;
    while (s < trips.size() && trips[s].row < rd) s++;
    for (int64_t row = rd; row < nodes; row++) {
        while (s < trips.size() && trips[s].row == row) {
            int64_t col = trips[s].col;
            double v = 0.0;
            bool have = false;
            while (s < trips.size() && trips[s].row == row && trips[s].col == col) {
                if (trips[s].add == 0) v = trips[s].val;
                else v = (have ? v : 0.0) - trips[s].val;
                have = true;
                s++;
            }
            if (col >= rd) {
                h->values.push_back(v);
                h->col_ind.push_back(col - rd);
            } else {
                h->b[row - rd] = h->b[row - rd] - (v * h->vef[col]);
            }
        }
        h->row_ptr[row - rd + 1] = (int64_t)h->values.size();
    }
    h->diag.assign(diag_full.begin() + rd, diag_full.end());
    for (int64_t i = 0; i < n; i++) {
        if (std::fabs(h->diag[i]) < kZero) {
            delete h;
            return fail("Fill_sparse_stiffness_matrix: zero element on the diagonal (floating node)");
        }
    }
    // node -> representative point, for node features
    h->node_point.assign((size_t)nodes, -1);
    for (int64_t t : cl)
        for (int64_t q = net->elem_ptr[t]; q < net->elem_ptr[t + 1]; q++) {
            int64_t nd = h->lmm[net->elem_pts[q]];
            if (nd >= rd && h->node_point[nd] < 0) h->node_point[nd] = net->elem_pts[q];
        }
    // Initial_guess_for_CG: ramp, only when R_flag && R_d == 2 (Direct_Electrifying.cpp:80)
    h->x0.assign((size_t)n, 0.0);
    h->r0 = h->b;
    int has_guess = (r_flag == 1 && rd == 2) ? 1 : 0;
    if (has_guess) {
        int fam = family;  // 0,1,2 here
        double box = net->prm.box;
        double l_win = box, max_coord = box;
        double vpc = ep->applied_voltage / l_win;
        for (int64_t t : cl)
            for (int64_t q = net->elem_ptr[t]; q < net->elem_ptr[t + 1]; q++) {
                int64_t pt = net->elem_pts[q];
                int64_t nd = h->lmm[pt];
                if (nd >= 2) {
                    double pc = net->xyz[3 * pt + (fam == 0 ? 0 : fam == 1 ? 1 : 2)];
                    h->x0[nd - rd] = vpc * (max_coord - pc);
                }
            }
        // Update_residual_vector: r0 = b - A x0 with the SSS loop order
        std::vector<double> ax((size_t)n, 1.0);
        for (int64_t r = 0; r < n; r++) {
            ax[r] = h->diag[r] * h->x0[r];
            for (int64_t j = h->row_ptr[r]; j < h->row_ptr[r + 1]; j++) {
                int64_t c = h->col_ind[j];
                ax[r] = ax[r] + h->values[j] * h->x0[c];
                ax[c] = ax[c] + h->values[j] * h->x0[r];
            }
        }
        for (int64_t i = 0; i < n; i++) h->r0[i] = h->r0[i] - ax[i];
    }
    // Boundary elements (Electrical_analysis.cpp:499-605) for b1, b2 of `direction`
    h->v.direction = -1;
    if (direction >= 0 && direction <= 2) {
        const int b1s[3] = {2, 3, 0}, b2s[3] = {4, 5, 1};
        int bb[2] = {b1s[direction], b2s[direction]};
        for (int w = 0; w < 2; w++) {
            for (int64_t t : net->bcnt[bb[w]]) {
                int64_t a = net->elem_ptr[t], b = net->elem_ptr[t + 1];
                if (b - a < 2) continue;
                const int64_t *e = &net->elem_pts[a];
                int64_t fr[2][2] = {{e[0], e[1]}, {e[b - a - 1], e[b - a - 2]}};
                for (int side = 0; side < 2; side++) {
                    int64_t p1 = fr[side][0], p2 = fr[side][1];
                    int64_t n1 = h->lmm[p1];
                    if (n1 < 0 || n1 > 1) continue;
                    int64_t n2 = h->lmm[p2];
                    if (n2 < 0) continue;
                    double re = p1 > p2 ? resistance_cnt(net, p2, p1, net->radius[t], ep->resistivity_cnt)
                                        : resistance_cnt(net, p1, p2, net->radius[t], ep->resistivity_cnt);
                    h->e_bnd[w].push_back(n1);
                    h->e_int[w].push_back(n2);
                    h->e_re[w].push_back(re);
                }
            }
        }
        h->v.direction = direction;
    }
    cicy_sss_view &v = h->v;
    v.priv = h;
    v.nodes = nodes;
    v.reserved = rd;
    v.n = n;
    v.nnz = (int64_t)h->values.size();
    v.row_ptr = h->row_ptr.data();
    v.col_ind = h->col_ind.data();
    v.values = h->values.data();
    v.diag = h->diag.data();
    v.b = h->b.data();
    v.vef = h->vef.data();
    v.x0 = h->x0.data();
    v.r0 = h->r0.data();
    v.has_guess = has_guess;
    v.n_points = P;
    v.lmm = h->lmm.data();
    v.node_point = h->node_point.data();
    for (int w = 0; w < 2; w++) {
        v.ne[w] = (int64_t)h->e_bnd[w].size();
        v.e_bnd[w] = h->e_bnd[w].data();
        v.e_int[w] = h->e_int[w].data();
        v.e_re[w] = h->e_re[w].data();
    }
    return &h->v;
}

void cicy_sss_free(cicy_sss_view *v) {
    if (!v) return;
    delete static_cast<cicy_sss_holder *>(v->priv);
}

}  // extern "C"
