// netgen.cpp -- seeded straight-tube CNT / square-plate GNP networks and the DEA system
// assembly.
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
// Nanoplatelets (GNPs): square plates (side gnp_side, thickness gnp_thick) with uniform
// centres and isotropic normals, placed after the tubes with the same random stream; a
// plate with no vertex inside the cube is drawn again.  Plates may cross each other
// (crossing plates are in contact at the van der Waals gap); the reference relocates
// GNPs to avoid interpenetration, which at 2 000 plates of 1 um in an 8 um cube jams a
// rejection sampler.  Plate points are
// created in the reference's order: one window-boundary point per face the plate
// crosses, the average of its edges' crossings with that face
// (Cutoff_Wins::Find_gnp_boundary_points, Cutoff_Wins.cpp:871-988; crossings of cube
// edges with plate faces are not added); plate-plate junction points (closest points of
// the mid-plane squares moved onto the facing surfaces, gap = distance - (t1 + t2)/2
// clamped at d_vdw); tube-plate junction points: every tube point whose gap to the plate
// (distance to the cuboid - r, clamped at d_vdw; points inside a plate are ignored) is
// within d_tunnel is a contact (Hoshen_Kopelman::Find_adjacent_labels, :1989-2142), and
// the contacts of one tube on one plate are compressed to the closest one per run of
// fewer than min_points points (Compress_mixed_contacts, :2181-2278), whose point on the
// plate becomes the plate point.  Clusters join tubes and plates (union-find, smaller root
// kept); boundary points on faces the cluster's family does not use are dropped from the
// plate's structure (Remove_gnp_points_in_non_relevant_boundaries, :2740-2892), and each
// plate of a cluster is triangulated in its mid-plane (Triangulation.cpp, gnp_geom.cpp).
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
//   LM_matrix_for_gnps          :307-349   plate points -> nodes after the tube nodes
//   Fill_2d_matrices_mixed_junctions :903-1019  tube-plate junctions (drop rule as above)
//   Fill_2d_matrices_gnp        :1021-1196 triangulation edges, scanned in reverse
//   Calculate_resistance_gnp    :1198-1232 rho = u.(rho_s, rho_s, rho_t)u in global axes
//   Fill_2d_matrices_gnp_junctions :1234-1327 plate-plate junctions
//   From_2d_to_1d_vectors       :1329-1394 SSS arrays, b = -K_EF^T VEF, reserved diag erased
//   zero-diagonal check         :458-484
//   Initial_guess_for_CG / Initial_guess_for_cnt_nodes :1396-1470 (ramp)
//   Update_residual_vector      :1502-1535
// and the boundary element rules of Electrical_analysis.cpp:499-605 (tubes) and :607-675
// (plates) for the current sums.
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

#include "geom3.h"

using cicy_geom::Plate;
using cicy_geom::seg_seg;

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

struct MixedJunction {
    int64_t p_cnt, p_gnp;  // tube point, plate point
    int64_t cnt, gnp;
    double dist;
};

struct GnpJunction {
    int64_t p1, p2;   // plate points
    int64_t g1, g2;   // plates (g1 < g2)
    double dist;
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
    // nanoplatelets
    std::vector<Plate> gnps;
    std::vector<double> gxyz;                      // plate points*3 (points_gnp)
    std::vector<std::vector<int64_t>> gstruct;     // structure_gnp
    std::vector<std::vector<std::pair<int64_t, int>>> gpts_bnd;  // (point, face) per plate
    std::vector<std::vector<int64_t>> gtri;        // triangulation edges, 2 per edge
    std::vector<int64_t> gbpts[6];                 // boundary_gnp_pts
    std::vector<int64_t> gbnd[6];                  // boundary_gnp
    std::vector<MixedJunction> mjunctions;
    std::vector<GnpJunction> gjunctions;
    std::vector<std::vector<int64_t>> cluster_gnps;    // plate ids, ascending
    std::vector<std::vector<int64_t>> cluster_mixed;   // tube-plate junction ids
    std::vector<std::vector<int64_t>> cluster_gjunc;   // plate-plate junction ids
    // flat copies of the lists of lists, filled by cicy_net_arrays_get
    struct Flat {
        std::vector<int64_t> ptr[14], ent[14];
        std::vector<double> thick;
    } flat;
};

struct cicy_sss_holder {
    cicy_sss_view v;
    std::vector<int64_t> row_ptr, col_ind, lmm, node_point, lmm_gnp;
    std::vector<double> values, diag, b, vef, x0, r0, node_xyz;
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

// Calculate_junction_resistance (Direct_Electrifying.cpp:854-901): l is a tube's radius
// (doubled to its diameter) or a plate's thickness; the junction area is the square of
// the smaller of the two.
double junction_resistance_typed(double dist, double l1, bool cnt1, double l2, bool cnt2,
                                 const cicy_elec_params *ep) {
    if (!ep->junction_exponential) return ep->junction_resistance;
    double jl1 = l1, jl2 = l2;
    if (cnt1) jl1 = jl1 + l1;
    if (cnt2) jl2 = jl2 + l2;
    double a = (jl1 - jl2 < kZero) ? jl1 * jl1 : jl2 * jl2;
    return ep->c1 * dist * std::exp(ep->c2 * dist) / a;
}

double junction_resistance(const Junction &j, double l1, double l2,
                           const cicy_elec_params *ep) {
    return junction_resistance_typed(j.dist, l1, true, l2, true, ep);
}

// Calculate_resistance_gnp (Direct_Electrifying.cpp:1198-1232): the resistivity tensor is
// diag(rho_s, rho_s, rho_t) in the global axes.  Returns false for a zero-length edge.
bool resistance_gnp(const double *p1, const double *p2, double rad1, double rad2,
                    const cicy_elec_params *ep, double *re) {
    const double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    const double len = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (len < kZero) return false;
    const double ux = (p2[0] - p1[0]) / len, uy = (p2[1] - p1[1]) / len,
                 uz = (p2[2] - p1[2]) / len;
    const double rx = ux * ep->resistivity_gnp_surf, ry = uy * ep->resistivity_gnp_surf,
                 rz = uz * ep->resistivity_gnp_t;
    const double rho = rx * ux + ry * uy + rz * uz;
    *re = rho * len / (kPi * rad1 * rad2);
    return true;
}

struct Trip {
    int64_t row, col;
    double val;
    int add;  // 0 overwrite (= -g), 1 accumulate (-= g)
};

}  // namespace

namespace {

constexpr double kSqrt2 = 1.4142135623730951;

// Uniform grid over [0,box]^3 with cells of at least `min_cell`.
struct Grid {
    int64_t nc = 1;
    double cs = 1.0;
    Grid(double box, double min_cell) {
        nc = std::max<int64_t>(1, (int64_t)std::floor(box / min_cell));
        cs = box / (double)nc;
    }
    int64_t at(double v) const {
        return std::clamp<int64_t>((int64_t)std::floor(v / cs), 0, nc - 1);
    }
    int64_t index(const double *p) const { return (at(p[0]) * nc + at(p[1])) * nc + at(p[2]); }
};

int64_t add_gnp_point(cicy_net *net, int64_t g, const double *x) {
    const int64_t id = (int64_t)net->gxyz.size() / 3;
    net->gxyz.insert(net->gxyz.end(), x, x + 3);
    net->gstruct[(size_t)g].push_back(id);
    return id;
}

// Plates with uniform centres, isotropic normals and a uniform in-plane angle, drawn again
// while no vertex lies strictly inside the cube.  Plates may cross each other: the
// generator does not relocate them (the reference's does), and crossing plates are in
// contact at the van der Waals gap.
void place_gnps(cicy_net *net, std::mt19937_64 &g) {
    const cicy_net_params &p = net->prm;
    if (p.n_gnp <= 0) return;
    const double box = p.box, l = p.gnp_side, t = p.gnp_thick;
    while ((int64_t)net->gnps.size() < p.n_gnp) {
        Plate P;
        P.l = l;
        P.t = t;
        for (int a = 0; a < 3; a++) P.c[a] = u01(g) * box;
        const double cth = 2.0 * u01(g) - 1.0, phi = 2.0 * kPi * u01(g),
                     psi = 2.0 * kPi * u01(g);
        const double sth = std::sqrt(std::max(0.0, 1.0 - cth * cth));
        const double n[3] = {sth * std::cos(phi), sth * std::sin(phi), cth};
        const double e[3] = {-std::sin(phi), std::cos(phi), 0.0};
        const double f[3] = {-cth * std::cos(phi), -cth * std::sin(phi), sth};
        for (int a = 0; a < 3; a++) {
            P.ax[0][a] = std::cos(psi) * e[a] + std::sin(psi) * f[a];
            P.ax[2][a] = n[a];
        }
        P.ax[1][0] = n[1] * P.ax[0][2] - n[2] * P.ax[0][1];
        P.ax[1][1] = n[2] * P.ax[0][0] - n[0] * P.ax[0][2];
        P.ax[1][2] = n[0] * P.ax[0][1] - n[1] * P.ax[0][0];
        bool inside = false;
        for (int v = 0; v < 8 && !inside; v++) {
            double x[3];
            cicy_geom::plate_vertex(P, v, x);
            inside = x[0] > 0.0 && x[0] < box && x[1] > 0.0 && x[1] < box && x[2] > 0.0 &&
                     x[2] < box;
        }
        if (inside) net->gnps.push_back(P);
    }
    const size_t G = net->gnps.size();
    net->gstruct.assign(G, {});
    net->gpts_bnd.assign(G, {});
    net->gtri.assign(G, {});
    // window-boundary points, face by face (Cutoff_Wins.cpp:871-988)
    for (size_t gi = 0; gi < G; gi++) {
        std::vector<double> acc[6];
        cicy_geom::plate_window_crossings(net->gnps[gi], box, acc);
        for (int f = 0; f < 6; f++) {
            if (acc[f].empty()) continue;
            const size_t m = acc[f].size() / 3;
            double avg[3] = {0.0, 0.0, 0.0};
            for (size_t q = 0; q < m; q++)
                for (int a = 0; a < 3; a++) avg[a] = avg[a] + acc[f][3 * q + a];
            for (int a = 0; a < 3; a++) avg[a] = avg[a] / (double)m;
            const int64_t id = add_gnp_point(net, (int64_t)gi, avg);
            net->gbnd[f].push_back((int64_t)gi);
            net->gbpts[f].push_back(id);
            net->gpts_bnd[gi].push_back({id, f});
        }
    }
}

// Plate-plate junctions: the closest points of the mid-plane squares, moved onto the
// facing surfaces; gap = distance - (t1 + t2)/2, clamped at d_vdw.
void gnp_gnp_junctions(cicy_net *net) {
    const cicy_net_params &p = net->prm;
    const size_t G = net->gnps.size();
    if (G < 2) return;
    Grid grid(p.box, kSqrt2 * p.gnp_side + p.gnp_thick + p.d_tunnel);
    std::vector<std::vector<int32_t>> cells((size_t)(grid.nc * grid.nc * grid.nc));
    for (size_t gi = 0; gi < G; gi++) cells[(size_t)grid.index(net->gnps[gi].c)].push_back((int32_t)gi);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    for (size_t gi = 0; gi < G; gi++) {
        const double *c = net->gnps[gi].c;
        const int64_t ci[3] = {grid.at(c[0]), grid.at(c[1]), grid.at(c[2])};
        for (int64_t i = ci[0] - 1; i <= ci[0] + 1; i++)
            for (int64_t j = ci[1] - 1; j <= ci[1] + 1; j++)
                for (int64_t m = ci[2] - 1; m <= ci[2] + 1; m++) {
                    if (i < 0 || j < 0 || m < 0 || i >= grid.nc || j >= grid.nc || m >= grid.nc)
                        continue;
                    for (int32_t q : cells[(size_t)((i * grid.nc + j) * grid.nc + m)])
                        if ((size_t)q > gi) pairs.emplace_back((int32_t)gi, q);
                }
    }
    std::sort(pairs.begin(), pairs.end());
    for (auto &pr : pairs) {
        const Plate &A = net->gnps[(size_t)pr.first], &B = net->gnps[(size_t)pr.second];
        double pa[3], pb[3];
        const double d = cicy_geom::midplane_distance(A, B, pa, pb);
        double gap = d - 0.5 * A.t - 0.5 * B.t;
        if (gap < p.d_vdw) gap = p.d_vdw;
        if (gap - p.d_tunnel > kZero) continue;
        const double ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
        const double sa = cicy_geom::dot3(A.ax[2], ab) >= 0.0 ? 1.0 : -1.0;
        const double sb = cicy_geom::dot3(B.ax[2], ab) <= 0.0 ? 1.0 : -1.0;
        double xa[3], xb[3];
        for (int a = 0; a < 3; a++) {
            xa[a] = pa[a] + sa * 0.5 * A.t * A.ax[2][a];
            xb[a] = pb[a] + sb * 0.5 * B.t * B.ax[2][a];
        }
        GnpJunction J;
        J.p1 = add_gnp_point(net, pr.first, xa);
        J.p2 = add_gnp_point(net, pr.second, xb);
        J.g1 = pr.first;
        J.g2 = pr.second;
        J.dist = gap;
        net->gjunctions.push_back(J);
    }
}

// Tube-plate junctions (Find_adjacent_labels + Compress_mixed_contacts).
void mixed_junctions(cicy_net *net) {
    const cicy_net_params &p = net->prm;
    const size_t G = net->gnps.size();
    if (G == 0) return;
    const double reach = p.radius + p.d_tunnel + 1e-9;
    Grid grid(p.box, std::min(p.box, 0.1));
    const double half_diag = 0.5 * std::sqrt(3.0) * grid.cs;
    // register every plate in the cells it can reach
    std::vector<std::pair<int64_t, int32_t>> reg;
    for (size_t gi = 0; gi < G; gi++) {
        const Plate &P = net->gnps[gi];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int v = 0; v < 8; v++) {
            double x[3];
            cicy_geom::plate_vertex(P, v, x);
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], x[a] - reach);
                hi[a] = std::max(hi[a], x[a] + reach);
            }
        }
        for (int64_t i = grid.at(lo[0]); i <= grid.at(hi[0]); i++)
            for (int64_t j = grid.at(lo[1]); j <= grid.at(hi[1]); j++)
                for (int64_t m = grid.at(lo[2]); m <= grid.at(hi[2]); m++) {
                    const double cc[3] = {(i + 0.5) * grid.cs, (j + 0.5) * grid.cs,
                                          (m + 0.5) * grid.cs};
                    double q[3], dist;
                    cicy_geom::plate_closest_point(P, cc, q, &dist);
                    if (dist <= half_diag + reach)
                        reg.emplace_back((i * grid.nc + j) * grid.nc + m, (int32_t)gi);
                }
    }
    std::sort(reg.begin(), reg.end());
    struct Contact {
        int64_t g, t, pt;
        double gap;
        double q[3];
    };
    std::vector<Contact> cont;
    const int64_t T = (int64_t)net->tubes.size();
    for (int64_t t = 0; t < T; t++)
        for (int64_t pt = net->cnt_ptr[t]; pt < net->cnt_ptr[t + 1]; pt++) {
            const double *x = &net->xyz[3 * pt];
            const int64_t ci = grid.index(x);
            auto it = std::lower_bound(reg.begin(), reg.end(), std::make_pair(ci, (int32_t)-1));
            for (; it != reg.end() && it->first == ci; ++it) {
                Contact c;
                double dist;
                if (!cicy_geom::plate_closest_point(net->gnps[(size_t)it->second], x, c.q, &dist))
                    continue;  // a tube point inside a plate makes no contact
                double gap = dist - net->radius[(size_t)t];
                if (gap < p.d_vdw) gap = p.d_vdw;
                if (gap - p.d_tunnel > kZero) continue;
                c.g = it->second;
                c.t = t;
                c.pt = pt;
                c.gap = gap;
                cont.push_back(c);
            }
        }
    std::sort(cont.begin(), cont.end(), [](const Contact &a, const Contact &b) {
        if (a.g != b.g) return a.g < b.g;
        if (a.t != b.t) return a.t < b.t;
        return a.pt < b.pt;
    });
    auto emit = [&](const Contact &c) {
        MixedJunction J;
        J.p_cnt = c.pt;
        J.p_gnp = add_gnp_point(net, c.g, c.q);
        J.cnt = c.t;
        J.gnp = c.g;
        J.dist = c.gap;
        net->mjunctions.push_back(J);
    };
    for (size_t s = 0; s < cont.size();) {
        size_t e = s;
        while (e < cont.size() && cont[e].g == cont[s].g && cont[e].t == cont[s].t) e++;
        // runs of contacts closer than min_points points to the run's first point; the
        // closest contact of each run is the junction
        int64_t p0 = cont[s].pt;
        size_t best = s;
        for (size_t q = s + 1; q < e; q++) {
            if (cont[q].pt - p0 >= p.min_points) {
                emit(cont[best]);
                best = q;
                p0 = cont[q].pt;
            } else if (cont[q].gap < cont[best].gap) {
                best = q;
            }
        }
        emit(cont[best]);
        s = e;
    }
}

// Boundary points on faces the cluster's family does not use leave the plate's structure;
// then every plate of a cluster is triangulated in its mid-plane.
void prune_and_triangulate(cicy_net *net) {
    std::vector<int> faces;
    std::vector<double> uv;
    std::vector<std::pair<int, int>> edges;
    for (size_t c = 0; c < net->cluster_gnps.size(); c++) {
        const int fam = net->family[c];
        if (fam >= 0 && fam <= 5) {
            faces = boundaries_of_family(fam);
            for (int64_t g : net->cluster_gnps[c]) {
                std::vector<int64_t> &st = net->gstruct[(size_t)g];
                for (auto &pb : net->gpts_bnd[(size_t)g]) {
                    if (std::find(faces.begin(), faces.end(), pb.second) != faces.end()) continue;
                    auto it = std::find(st.begin(), st.end(), pb.first);
                    if (it != st.end()) st.erase(it);
                }
            }
        }
        for (int64_t g : net->cluster_gnps[c]) {
            const Plate &P = net->gnps[(size_t)g];
            const std::vector<int64_t> &st = net->gstruct[(size_t)g];
            uv.clear();
            for (int64_t id : st) {
                double loc[3];
                cicy_geom::to_local(P, &net->gxyz[3 * (size_t)id], loc);
                uv.push_back(loc[0]);
                uv.push_back(loc[1]);
            }
            cicy_geom::plate_triangulation(P.l, uv, edges);
            std::vector<int64_t> &tri = net->gtri[(size_t)g];
            tri.clear();
            for (auto &e : edges) {
                tri.push_back(st[(size_t)e.first]);
                tri.push_back(st[(size_t)e.second]);
            }
        }
    }
}

// ---- backbone network (Backbone_Network.cpp) ----------------------------------------

// Length of tube polyline points P1 .. P2 times the cross-section
// (CNT_volume_between_two_points, Backbone_Network.cpp:557-574)
double cnt_volume(const cicy_net *net, int64_t p1, int64_t p2, double r) {
    double len = 0.0;
    for (int64_t i = p1; i < p2; i++) {
        const double *a = &net->xyz[3 * (size_t)i], *b = &net->xyz[3 * (size_t)(i + 1)];
        const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
        len = len + std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return len * kPi * r * r;
}

// Volume of plate g inside the cube: l^2 t times the share of a 20 x 20 grid over its
// mid-plane square that lies in [0, box]^3, when any does
// (Generate_Network::Approximate_gnp_volume_inside_sample, Generate_Network.cpp:4750-4810)
double gnp_volume(const cicy_net *net, int64_t g) {
    const Plate &P = net->gnps[(size_t)g];
    const double box = net->prm.box, h = 0.5 * P.l;
    double c[4][3];
    for (int k = 0; k < 3; k++) {
        c[0][k] = P.c[k] - h * P.ax[0][k] - h * P.ax[1][k];  // corner
        c[1][k] = P.c[k] + h * P.ax[0][k] - h * P.ax[1][k];  // right
        c[2][k] = P.c[k] - h * P.ax[0][k] + h * P.ax[1][k];  // up
        c[3][k] = P.c[k] + h * P.ax[0][k] + h * P.ax[1][k];  // up right
    }
    const int np = 20;
    int in = 0;
    for (int i = 0; i < np; i++) {
        const double lr = (double)i / (np - 1);
        for (int j = 0; j < np; j++) {
            const double lu = (double)j / (np - 1);
            bool inside = true;
            for (int k = 0; k < 3; k++) {
                const double st = c[0][k] + (c[1][k] - c[0][k]) * lr;
                const double en = c[2][k] + (c[3][k] - c[2][k]) * lr;
                const double q = st + (en - st) * lu;
                if (q < 0.0 || q > box) inside = false;
            }
            if (inside) in++;
        }
    }
    double v = P.l * P.l * P.t;
    if (in != 0) v = v * ((double)in / (np * np));
    return v;
}

void remove_point(std::vector<int64_t> &v, int64_t p) {
    auto it = std::find(v.begin(), v.end(), p);
    if (it != v.end()) v.erase(it);
}

}  // namespace

extern "C" {

int cicy_net_backbone(cicy_net *net, int64_t k, const double *volt, int64_t nvolt,
                      const int64_t *lmm, int64_t n_points, const int64_t *lmm_gnp,
                      int64_t n_gnp_points, int reserved, int avoid_resistance,
                      const double *zero_in, const int32_t *seg_in, double *out, char *err,
                      int64_t err_len) {
    auto fail = [&](const char *msg) {
        if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", msg);
        return 0;
    };
    if (!net || !volt || !lmm || !out || k < 0 || k >= (int64_t)net->clusters.size())
        return fail("invalid network, cluster or array");
    if (n_points != (int64_t)net->xyz.size() / 3 ||
        (lmm_gnp == nullptr ? 0 : n_gnp_points) != (int64_t)net->gxyz.size() / 3)
        return fail("LMM arrays do not match the network's points");
    auto node_c = [&](int64_t p) -> int64_t {
        const int64_t n = lmm[p];
        return (n >= 0 && n < nvolt) ? n : -1;
    };
    auto node_g = [&](int64_t p) -> int64_t {
        const int64_t n = lmm_gnp[p];
        return (n >= 0 && n < nvolt) ? n : -1;
    };
    bool bad = false;
    auto I = [&](int64_t n1, int64_t n2) {
        if (n1 < 0 || n2 < 0) {
            bad = true;
            return 0.0;
        }
        return std::fabs(volt[n1] - volt[n2]);
    };
    std::vector<int64_t> &cl = net->clusters[(size_t)k];
    std::vector<int64_t> &cj = net->cluster_junctions[(size_t)k];
    std::vector<int64_t> &cg = net->cluster_gnps[(size_t)k];
    std::vector<int64_t> &cm = net->cluster_mixed[(size_t)k];
    std::vector<int64_t> &cgj = net->cluster_gjunc[(size_t)k];
    // the tubes' element points (HoKo->elements_cnt), edited below
    const size_t T = net->radius.size();
    std::vector<std::vector<int64_t>> el(T);
    for (size_t t = 0; t < T; t++)
        el[t].assign(net->elem_pts.begin() + net->elem_ptr[t],
                     net->elem_pts.begin() + net->elem_ptr[t + 1]);
    // Find_zero_current (:64-221): the largest |dV| over tube elements, tube junctions,
    // plate triangulation edges, plate junctions and mixed junctions, times 1e-9
    double imax = kZero;
    for (int64_t t : cl)
        for (size_t q = 1; q < el[(size_t)t].size(); q++) {
            const double c = I(node_c(el[(size_t)t][q - 1]), node_c(el[(size_t)t][q]));
            if (c > imax) imax = c;
        }
    for (int64_t j : cj) {
        const Junction &J = net->junctions[(size_t)j];
        const double c = I(node_c(J.p1), node_c(J.p2));
        if (c > imax) imax = c;
    }
    for (int64_t g : cg) {
        const std::vector<int64_t> &tri = net->gtri[(size_t)g];
        for (size_t e = 0; e + 1 < tri.size(); e += 2) {
            const double c = I(node_g(tri[e]), node_g(tri[e + 1]));
            if (c > imax) imax = c;
        }
    }
    for (int64_t j : cgj) {
        const GnpJunction &J = net->gjunctions[(size_t)j];
        const double c = I(node_g(J.p1), node_g(J.p2));
        if (c > imax) imax = c;
    }
    for (int64_t j : cm) {
        const MixedJunction &J = net->mjunctions[(size_t)j];
        const double c = I(node_c(J.p_cnt), node_g(J.p_gnp));
        if (c > imax) imax = c;
    }
    if (bad) return fail("a point of the cluster has no node in the given LMM");
    const double zero = zero_in ? *zero_in : imax * 1e-9;
    double vol_bb = 0.0, vol_dead = 0.0, gvol_bb = 0.0, gvol_dead = 0.0;
    // Find_backbone_cnts (:317-409) with Calculate_cnt_volumes (:412-555): the backbone of
    // a tube is the run of elements from the first carrying more than the zero current to
    // the next one that does not; tubes are visited last to first
    if (!cl.empty()) {
        for (int64_t i = (int64_t)cl.size() - 1; i >= 0; i--) {
            const int64_t t = cl[(size_t)i];
            std::vector<int64_t> &e = el[(size_t)t];
            int idx1 = -1, idx2 = -1;
            if (seg_in) {
                idx1 = seg_in[2 * i];
                idx2 = seg_in[2 * i + 1];
            } else {
                for (size_t q = 0; q + 1 < e.size(); q++) {
                    const double c = I(node_c(e[q]), node_c(e[q + 1]));
                    if (c - zero > kZero) {
                        if (idx1 == -1) idx1 = (int)q;
                    } else if (idx1 != -1 && idx2 == -1) {
                        idx2 = (int)q;
                        break;
                    }
                }
            }
            if (e.empty()) return fail("a tube of the cluster has no element points");
            if (idx1 >= (int)e.size() || idx2 >= (int)e.size() || (idx2 >= 0 && idx2 <= idx1))
                return fail("backbone segment out of the tube's elements");
            const double r = net->radius[(size_t)t];
            const int64_t ps = e.front(), pe = e.back();
            if (idx1 == -1) {
                vol_dead = vol_dead + cnt_volume(net, ps, pe, r);
                if (!avoid_resistance) e.clear();
            } else {
                const int64_t p1 = e[(size_t)idx1];
                vol_dead = vol_dead + cnt_volume(net, ps, p1, r);
                double vb;
                if (idx2 == -1) {
                    vb = cnt_volume(net, p1, pe, r);
                } else {
                    const int64_t p2 = e[(size_t)idx2];
                    vb = cnt_volume(net, p1, p2, r);
                    vol_dead = vol_dead + cnt_volume(net, p2, pe, r);
                    if (!avoid_resistance) e.erase(e.begin() + idx2 + 1, e.end());
                }
                vol_bb = vol_bb + vb;
                if (!avoid_resistance && idx1 > 0) e.erase(e.begin(), e.begin() + idx1);
            }
            if (!avoid_resistance && e.empty()) cl.erase(cl.begin() + i);
        }
    }
    if (!cg.empty()) {
        auto in_el = [&](int64_t t, int64_t p) {
            const std::vector<int64_t> &e = el[(size_t)t];
            return std::binary_search(e.begin(), e.end(), p);
        };
        // Find_and_remove_gnp_and_mixed_junctions_below_zero_current (:663-796)
        for (int64_t i = (int64_t)cm.size() - 1; i >= 0; i--) {
            const MixedJunction &J = net->mjunctions[(size_t)cm[(size_t)i]];
            std::vector<int64_t> &e = el[(size_t)J.cnt];
            const bool present = in_el(J.cnt, J.p_cnt);
            const double c = I(node_c(J.p_cnt), node_g(J.p_gnp));
            if (c - zero < kZero || e.empty() || !present) {
                remove_point(net->gstruct[(size_t)J.gnp], J.p_gnp);
                cm.erase(cm.begin() + i);
                if (present) {
                    e.erase(std::lower_bound(e.begin(), e.end(), J.p_cnt));
                    if (e.size() < 2) e.clear();
                }
            }
        }
        for (int64_t i = (int64_t)cgj.size() - 1; i >= 0; i--) {
            const GnpJunction &J = net->gjunctions[(size_t)cgj[(size_t)i]];
            const double c = I(node_g(J.p1), node_g(J.p2));
            if (c - zero < kZero) {
                remove_point(net->gstruct[(size_t)J.g1], J.p1);
                remove_point(net->gstruct[(size_t)J.g2], J.p2);
                cgj.erase(cgj.begin() + i);
            }
        }
        // a plate's relevant boundary points: those of its structure on the reserved nodes
        // (points on the family's other faces were removed before the solve)
        std::vector<int64_t> nbnd(cg.size(), 0);
        for (size_t i = 0; i < cg.size(); i++)
            for (int64_t q : net->gstruct[(size_t)cg[i]])
                if (lmm_gnp[q] >= 0 && lmm_gnp[q] < reserved) nbnd[i]++;
        // Clear_structure_of_dead_gnps (:797-855)
        auto clear_dead = [&](int &to_remove) -> bool {
            for (size_t i = 0; i < cg.size(); i++) {
                std::vector<int64_t> &st = net->gstruct[(size_t)cg[i]];
                const int64_t nt = (int64_t)st.size(), nb = nbnd[i], nj = nt - nb;
                if (nt == 0 || (nb == 0 && nj == 1) || nb == nt) {
                    if (!st.empty()) {
                        st.clear();
                        to_remove++;
                    }
                    if (nj == 1) to_remove++;
                } else if (st.size() < 2) {
                    return false;
                }
            }
            return true;
        };
        int to_remove = 0;
        if (!clear_dead(to_remove)) return fail("a live plate with fewer than two points");
        while (to_remove) {
            to_remove = 0;
            // Remove_junctions_with_dead_particles_gnps (:999-1048)
            for (int64_t i = (int64_t)cgj.size() - 1; i >= 0; i--) {
                const GnpJunction &J = net->gjunctions[(size_t)cgj[(size_t)i]];
                std::vector<int64_t> &s1 = net->gstruct[(size_t)J.g1];
                std::vector<int64_t> &s2 = net->gstruct[(size_t)J.g2];
                if (s1.empty() || s2.empty()) {
                    cgj.erase(cgj.begin() + i);
                    if (!s1.empty()) remove_point(s1, J.p1);
                    if (!s2.empty()) remove_point(s2, J.p2);
                }
            }
            // Remove_junctions_with_dead_particles_mixed (:1061-1127)
            for (int64_t i = (int64_t)cm.size() - 1; i >= 0; i--) {
                const MixedJunction &J = net->mjunctions[(size_t)cm[(size_t)i]];
                std::vector<int64_t> &st = net->gstruct[(size_t)J.gnp];
                if (el[(size_t)J.cnt].empty() || st.empty()) {
                    cm.erase(cm.begin() + i);
                    if (!st.empty()) remove_point(st, J.p_gnp);
                } else if (!in_el(J.cnt, J.p_cnt)) {
                    cm.erase(cm.begin() + i);
                    remove_point(st, J.p_gnp);
                }
            }
            if (!clear_dead(to_remove)) return fail("a live plate with fewer than two points");
        }
        // Remove_additional_cnts_from_cluster (:857-880)
        for (int64_t i = (int64_t)cl.size() - 1; i >= 0; i--)
            if (el[(size_t)cl[(size_t)i]].empty()) cl.erase(cl.begin() + i);
        // Calculate_gnp_backbone_volume (:882-951)
        for (int64_t g : cg) {
            if (!net->gstruct[(size_t)g].empty()) gvol_bb = gvol_bb + gnp_volume(net, g);
            else gvol_dead = gvol_dead + gnp_volume(net, g);
        }
    }
    // Remove_junctions_with_dead_particles_cnts (:953-997)
    if (!avoid_resistance) {
        auto in_el = [&](int64_t t, int64_t p) {
            const std::vector<int64_t> &e = el[(size_t)t];
            return std::binary_search(e.begin(), e.end(), p);
        };
        for (int64_t i = (int64_t)cj.size() - 1; i >= 0; i--) {
            const Junction &J = net->junctions[(size_t)cj[(size_t)i]];
            if (el[(size_t)J.n1].empty() || el[(size_t)J.n2].empty() || !in_el(J.n1, J.p1) ||
                !in_el(J.n2, J.p2))
                cj.erase(cj.begin() + i);
        }
    }
    // write the edited element lists back; plates are triangulated again from their edited
    // structures, as the reference's next assembly does (Direct_Electrifying.cpp:1037-1040)
    net->elem_ptr.assign(T + 1, 0);
    net->elem_pts.clear();
    for (size_t t = 0; t < T; t++) {
        net->elem_pts.insert(net->elem_pts.end(), el[t].begin(), el[t].end());
        net->elem_ptr[t + 1] = (int64_t)net->elem_pts.size();
    }
    std::vector<double> uv;
    std::vector<std::pair<int, int>> edges;
    for (int64_t g : cg) {
        const Plate &P = net->gnps[(size_t)g];
        const std::vector<int64_t> &st = net->gstruct[(size_t)g];
        uv.clear();
        for (int64_t id : st) {
            double loc[3];
            cicy_geom::to_local(P, &net->gxyz[3 * (size_t)id], loc);
            uv.push_back(loc[0]);
            uv.push_back(loc[1]);
        }
        std::vector<int64_t> &tri = net->gtri[(size_t)g];
        tri.clear();
        if (st.empty()) continue;
        cicy_geom::plate_triangulation(P.l, uv, edges);
        for (auto &e : edges) {
            tri.push_back(st[(size_t)e.first]);
            tri.push_back(st[(size_t)e.second]);
        }
    }
    out[0] = zero;
    out[1] = vol_bb;
    out[2] = vol_dead;
    out[3] = gvol_bb;
    out[4] = gvol_dead;
    return 1;
}

void cicy_tunnel_constants(double h, double e, double m, double lambda, double *c1,
                           double *c2) {
    double s = std::sqrt(2.0 * m * lambda * e);
    *c1 = 10.0 * h * h / (m * m * s);
    *c2 = 4000.0 * kPi * s / h;
}

cicy_net *cicy_net_generate(const cicy_net_params *p) {
    if (!p || p->box <= 0 || p->n_cnt < 0 || p->seg_len <= 0 || p->radius <= 0 ||
        p->len_min <= 0 || p->len_max < p->len_min || p->n_gnp < 0 ||
        (p->n_gnp > 0 && (p->gnp_side <= 0 || p->gnp_thick <= 0 || p->min_points < 1)))
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
    place_gnps(net, g);
    const int64_t G = (int64_t)net->gnps.size();

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

    gnp_gnp_junctions(net);
    mixed_junctions(net);

    // ---- clusters: union-find over tubes (0..T-1) and plates (T..T+G-1) keeping the
    // smaller root (as Merge_labels does, Hoshen_Kopelman.cpp:2977-2985); clusters need
    // >= 2 particles (isolated particles are never clusters)
    std::vector<int64_t> par(T + G);
    std::iota(par.begin(), par.end(), 0);
    std::vector<int64_t> deg(T + G, 0);
    auto unite = [&](int64_t x, int64_t y) {
        int64_t a = find_root(par, x), b = find_root(par, y);
        if (a != b) {
            if (a < b) par[b] = a; else par[a] = b;
        }
        deg[x]++;
        deg[y]++;
    };
    for (const Junction &J : net->junctions) unite(J.n1, J.n2);
    for (const GnpJunction &J : net->gjunctions) unite(T + J.g1, T + J.g2);
    for (const MixedJunction &J : net->mjunctions) unite(J.cnt, T + J.gnp);
    std::vector<int64_t> cid(T + G, -1);
    for (int64_t x = 0; x < T + G; x++) {
        if (!deg[x]) continue;
        int64_t r = find_root(par, x);
        if (cid[r] < 0) {
            cid[r] = (int64_t)net->clusters.size();
            net->clusters.emplace_back();
            net->cluster_gnps.emplace_back();
        }
        cid[x] = cid[r];
        if (x < T) net->clusters[cid[x]].push_back(x);
        else net->cluster_gnps[cid[x]].push_back(x - T);
    }
    const size_t NC = net->clusters.size();
    net->cluster_junctions.resize(NC);
    net->cluster_mixed.resize(NC);
    net->cluster_gjunc.resize(NC);
    for (int64_t jx = 0; jx < (int64_t)net->junctions.size(); jx++)
        net->cluster_junctions[cid[find_root(par, net->junctions[jx].n1)]].push_back(jx);
    for (int64_t jx = 0; jx < (int64_t)net->mjunctions.size(); jx++)
        net->cluster_mixed[cid[find_root(par, net->mjunctions[jx].cnt)]].push_back(jx);
    for (int64_t jx = 0; jx < (int64_t)net->gjunctions.size(); jx++)
        net->cluster_gjunc[cid[find_root(par, T + net->gjunctions[jx].g1)]].push_back(jx);
    // percolation families (0 X, 1 Y, 2 Z, 3 XY, 4 XZ, 5 YZ, 6 XYZ) from the tubes and
    // plates at each face (boundary_cnt, boundary_gnp)
    std::vector<uint8_t> touch(T + G, 0);
    for (int f = 0; f < 6; f++) {
        for (int64_t t : net->bcnt[f]) touch[t] |= (uint8_t)(1u << f);
        for (int64_t gp : net->gbnd[f]) touch[T + gp] |= (uint8_t)(1u << f);
    }
    for (size_t c = 0; c < NC; c++) {
        uint8_t m = 0;
        for (int64_t t : net->clusters[c]) m |= touch[t];
        for (int64_t gp : net->cluster_gnps[c]) m |= touch[T + gp];
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
        for (const MixedJunction &J : net->mjunctions) extra[J.cnt].push_back(J.p_cnt);
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
    prune_and_triangulate(net);
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
        case 4: return (int64_t)net->gnps.size();
        case 5: return (int64_t)net->gxyz.size() / 3;
        case 6: return (int64_t)net->mjunctions.size();
        case 7: return (int64_t)net->gjunctions.size();
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

const double *cicy_net_gnp_xyz(const cicy_net *net) { return net ? net->gxyz.data() : nullptr; }

int cicy_net_gnp(const cicy_net *net, int64_t g, double *centre, double *axes, double *side,
                 double *thick) {
    if (!net || g < 0 || g >= (int64_t)net->gnps.size()) return 0;
    const Plate &P = net->gnps[(size_t)g];
    for (int a = 0; a < 3; a++) {
        centre[a] = P.c[a];
        for (int b = 0; b < 3; b++) axes[3 * a + b] = P.ax[a][b];
    }
    *side = P.l;
    *thick = P.t;
    return 1;
}

int64_t cicy_net_gnp_structure(const cicy_net *net, int64_t g, const int64_t **pts) {
    if (!net || g < 0 || g >= (int64_t)net->gstruct.size()) return -1;
    *pts = net->gstruct[(size_t)g].data();
    return (int64_t)net->gstruct[(size_t)g].size();
}

int64_t cicy_net_gnp_triangulation(const cicy_net *net, int64_t g, const int64_t **edges) {
    if (!net || g < 0 || g >= (int64_t)net->gtri.size()) return -1;
    *edges = net->gtri[(size_t)g].data();
    return (int64_t)net->gtri[(size_t)g].size() / 2;
}

int64_t cicy_net_gnp_boundary_points(const cicy_net *net, int b, const int64_t **pts) {
    if (!net || b < 0 || b > 5) return -1;
    *pts = net->gbpts[b].data();
    return (int64_t)net->gbpts[b].size();
}

int64_t cicy_net_boundary_gnps(const cicy_net *net, int b, const int64_t **gnps) {
    if (!net || b < 0 || b > 5) return -1;
    *gnps = net->gbnd[b].data();
    return (int64_t)net->gbnd[b].size();
}

int cicy_net_mixed_junction(const cicy_net *net, int64_t j, int64_t *p_cnt, int64_t *p_gnp,
                            int64_t *cnt, int64_t *gnp, double *dist) {
    if (!net || j < 0 || j >= (int64_t)net->mjunctions.size()) return 0;
    const MixedJunction &J = net->mjunctions[(size_t)j];
    *p_cnt = J.p_cnt; *p_gnp = J.p_gnp; *cnt = J.cnt; *gnp = J.gnp; *dist = J.dist;
    return 1;
}

int cicy_net_gnp_junction(const cicy_net *net, int64_t j, int64_t *p1, int64_t *p2,
                          int64_t *g1, int64_t *g2, double *dist) {
    if (!net || j < 0 || j >= (int64_t)net->gjunctions.size()) return 0;
    const GnpJunction &J = net->gjunctions[(size_t)j];
    *p1 = J.p1; *p2 = J.p2; *g1 = J.g1; *g2 = J.g2; *dist = J.dist;
    return 1;
}

int64_t cicy_net_cluster_gnps(const cicy_net *net, int64_t k, const int64_t **gnps) {
    if (!net || k < 0 || k >= (int64_t)net->cluster_gnps.size()) return -1;
    *gnps = net->cluster_gnps[(size_t)k].data();
    return (int64_t)net->cluster_gnps[(size_t)k].size();
}

int64_t cicy_net_cluster_mixed_junctions(const cicy_net *net, int64_t k, const int64_t **junc) {
    if (!net || k < 0 || k >= (int64_t)net->cluster_mixed.size()) return -1;
    *junc = net->cluster_mixed[(size_t)k].data();
    return (int64_t)net->cluster_mixed[(size_t)k].size();
}

int64_t cicy_net_cluster_gnp_junctions(const cicy_net *net, int64_t k, const int64_t **junc) {
    if (!net || k < 0 || k >= (int64_t)net->cluster_gjunc.size()) return -1;
    *junc = net->cluster_gjunc[(size_t)k].data();
    return (int64_t)net->cluster_gjunc[(size_t)k].size();
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
    // LM_matrix_for_gnps (Direct_Electrifying.cpp:307-349): boundary plate points of the
    // family's faces take the reserved nodes, then the cluster's plates number their points
    const int64_t GP = (int64_t)net->gxyz.size() / 3;
    h->lmm_gnp.assign((size_t)GP, -1);
    std::vector<int64_t> &lmm_g = h->lmm_gnp;
    for (size_t nidx = 0; nidx < bnds.size(); nidx++)
        for (int64_t q : net->gbpts[bnds[nidx]]) lmm_g[q] = (int64_t)nidx;
    const auto &clg = net->cluster_gnps[k];
    const int64_t first_gnp_node = gn;
    for (int64_t gp : clg)
        for (int64_t q : net->gstruct[gp])
            if (lmm_g[q] < 0) lmm_g[q] = gn++;
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
    // Fill_2d_matrices_mixed_junctions (:903-1019); points_cnt_rad records the tube radius
    // at the plate point of each junction (negative: absent)
    std::vector<double> cnt_rad((size_t)GP, -1.0);
    for (int64_t jx : net->cluster_mixed[k]) {
        const MixedJunction &J = net->mjunctions[jx];
        int64_t n1 = h->lmm[J.p_cnt], n2 = lmm_g[J.p_gnp];
        if (n1 < 0 || n2 < 0) continue;
        double g = 1.0;
        if (r_flag == 1)
            g = 1 / junction_resistance_typed(J.dist, net->radius[J.cnt], true,
                                              net->gnps[J.gnp].t, false, ep);
        if (n1 >= rd && n2 >= rd) add_new(n1, n2, g);
        else if (rd > 2) add_existing(n1, n2, g);
        cnt_rad[J.p_gnp] = net->radius[J.cnt];
    }
    // Fill_2d_matrices_gnp (:1021-1196): triangulation edges from the last to the first; an
    // end without a tube junction uses half the plate thickness as its radius
    const double *gx = net->gxyz.data();
    for (int64_t gp : clg) {
        const std::vector<int64_t> &tri = net->gtri[gp];
        const double half_t = net->gnps[gp].t / 2;
        for (int64_t j = (int64_t)tri.size() / 2 - 1; j >= 0; j--) {
            const int64_t v1 = tri[2 * j], v2 = tri[2 * j + 1];
            const int64_t n1 = lmm_g[v1], n2 = lmm_g[v2];
            if (n1 < 0 || n2 < 0) {
                delete h;
                return fail("Fill_2d_matrices_gnp: triangulation vertex without a node");
            }
            double g = 1.0;
            if (r_flag == 1) {
                const double rad1 = cnt_rad[v1] >= 0 ? cnt_rad[v1] : half_t;
                const double rad2 = cnt_rad[v2] >= 0 ? cnt_rad[v2] : half_t;
                double re;
                if (!resistance_gnp(gx + 3 * v1, gx + 3 * v2, rad1, rad2, ep, &re)) {
                    delete h;
                    return fail("Calculate_resistance_gnp: triangulation edge of zero length");
                }
                g = 1 / re;
            }
            add_new(n1, n2, g);
        }
    }
    // Fill_2d_matrices_gnp_junctions (:1234-1327): no reserved-node rule here
    for (int64_t jx : net->cluster_gjunc[k]) {
        const GnpJunction &J = net->gjunctions[jx];
        int64_t n1 = lmm_g[J.p1], n2 = lmm_g[J.p2];
        if (n1 < 0 || n2 < 0) continue;
        double g = 1.0;
        if (r_flag == 1)
            g = 1 / junction_resistance_typed(J.dist, net->gnps[J.g1].t, false,
                                              net->gnps[J.g2].t, false, ep);
        add_new(n1, n2, g);
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
    // node -> representative point (tube nodes), and every node's position
    h->node_point.assign((size_t)nodes, -1);
    h->node_xyz.assign((size_t)nodes * 3, NAN);
    for (int64_t t : cl)
        for (int64_t q = net->elem_ptr[t]; q < net->elem_ptr[t + 1]; q++) {
            int64_t nd = h->lmm[net->elem_pts[q]];
            if (nd >= rd && h->node_point[nd] < 0) {
                h->node_point[nd] = net->elem_pts[q];
                for (int a = 0; a < 3; a++) h->node_xyz[3 * nd + a] = net->xyz[3 * net->elem_pts[q] + a];
            }
        }
    for (int64_t gp : clg)
        for (int64_t q : net->gstruct[gp]) {
            int64_t nd = lmm_g[q];
            if (nd >= first_gnp_node)
                for (int a = 0; a < 3; a++) h->node_xyz[3 * nd + a] = gx[3 * q + a];
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
            // Currents_through_boundary_gnps (Electrical_analysis.cpp:607-675): the
            // triangulation edges of the boundary's plates that touch node 0 or 1, with
            // node = 0 for b1 and 1 for b2 (plates outside the cluster have no nodes here)
            const int64_t node = w;
            for (int64_t gp : net->gbnd[bb[w]]) {
                if (!std::binary_search(clg.begin(), clg.end(), gp)) continue;
                const std::vector<int64_t> &tri = net->gtri[gp];
                const double half_t = net->gnps[gp].t / 2;
                for (size_t j = 0; j < tri.size() / 2; j++) {
                    const int64_t v1 = tri[2 * j], v2 = tri[2 * j + 1];
                    const int64_t na = lmm_g[v1], nb = lmm_g[v2];
                    if (na > 1 && nb > 1) continue;
                    if (na == nb) {
                        delete h;
                        return fail("Currents_through_boundary_gnps: both nodes at the same boundary");
                    }
                    const int64_t n1 = (na == node) ? na : nb, n2 = (na == node) ? nb : na;
                    const double rad1 = cnt_rad[v1] >= 0 ? cnt_rad[v1] : half_t;
                    const double rad2 = cnt_rad[v2] >= 0 ? cnt_rad[v2] : half_t;
                    double re;
                    if (!resistance_gnp(gx + 3 * v1, gx + 3 * v2, rad1, rad2, ep, &re)) {
                        delete h;
                        return fail("Calculate_resistance_gnp: boundary edge of zero length");
                    }
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
    v.node_xyz = h->node_xyz.data();
    v.gnp_nodes = nodes - first_gnp_node;
    v.n_gnp_points = GP;
    v.lmm_gnp = h->lmm_gnp.data();
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

static_assert(sizeof(Junction) == sizeof(cicy_junction_rec) &&
                  sizeof(MixedJunction) == sizeof(cicy_junction_rec) &&
                  sizeof(GnpJunction) == sizeof(cicy_junction_rec),
              "junction records share one layout");

int cicy_net_arrays_get(cicy_net *net, cicy_net_arrays *out) {
    if (!net || !out) return 0;
    auto &f = net->flat;
    int slot = 0;
    // a list of lists as offsets + entries, kept in the network until the next call
    // (`count` lists: missing ones are empty)
    auto flatten = [&](const std::vector<std::vector<int64_t>> &ll, size_t count,
                       const int64_t **ptr, const int64_t **ent) {
        std::vector<int64_t> &p = f.ptr[slot], &e = f.ent[slot];
        slot++;
        p.assign(1, 0);
        e.clear();
        for (size_t i = 0; i < count; i++) {
            if (i < ll.size()) e.insert(e.end(), ll[i].begin(), ll[i].end());
            p.push_back((int64_t)e.size());
        }
        *ptr = p.data();
        *ent = e.data();
    };
    auto faces = [&](const std::vector<int64_t> *six, const int64_t **ptr, const int64_t **ent) {
        std::vector<std::vector<int64_t>> ll(six, six + 6);
        flatten(ll, 6, ptr, ent);
    };
    cicy_net_arrays &a = *out;
    a = cicy_net_arrays{};
    a.n_points = (int64_t)net->xyz.size() / 3;
    a.xyz = net->xyz.data();
    a.n_tubes = (int64_t)net->radius.size();
    a.radius = net->radius.data();
    a.elem_ptr = net->elem_ptr.data();
    a.elem_pts = net->elem_pts.data();
    a.n_junctions = (int64_t)net->junctions.size();
    a.junctions = reinterpret_cast<const cicy_junction_rec *>(net->junctions.data());
    a.n_gnp_points = (int64_t)net->gxyz.size() / 3;
    a.gxyz = net->gxyz.data();
    a.n_gnps = (int64_t)net->gnps.size();
    f.thick.resize(net->gnps.size());
    for (size_t g = 0; g < net->gnps.size(); g++) f.thick[g] = net->gnps[g].t;
    a.gnp_thickness = f.thick.data();
    flatten(net->gstruct, net->gnps.size(), &a.gstruct_ptr, &a.gstruct);
    flatten(net->gtri, net->gnps.size(), &a.gtri_ptr, &a.gtri);
    a.n_mixed = (int64_t)net->mjunctions.size();
    a.mixed = reinterpret_cast<const cicy_junction_rec *>(net->mjunctions.data());
    a.n_gjunctions = (int64_t)net->gjunctions.size();
    a.gjunctions = reinterpret_cast<const cicy_junction_rec *>(net->gjunctions.data());
    faces(net->bpts, &a.bpts_ptr, &a.bpts);
    faces(net->bcnt, &a.bcnt_ptr, &a.bcnt);
    faces(net->gbpts, &a.gbpts_ptr, &a.gbpts);
    faces(net->gbnd, &a.gbnd_ptr, &a.gbnd);
    a.n_clusters = (int64_t)net->clusters.size();
    flatten(net->clusters, net->clusters.size(), &a.cl_tubes_ptr, &a.cl_tubes);
    flatten(net->cluster_junctions, net->clusters.size(), &a.cl_junc_ptr, &a.cl_junc);
    flatten(net->cluster_gnps, net->clusters.size(), &a.cl_gnps_ptr, &a.cl_gnps);
    flatten(net->cluster_mixed, net->clusters.size(), &a.cl_mixed_ptr, &a.cl_mixed);
    flatten(net->cluster_gjunc, net->clusters.size(), &a.cl_gjunc_ptr, &a.cl_gjunc);
    for (int d = 0; d < 3; d++) {
        a.window_lo[d] = 0.0;
        a.window_hi[d] = net->prm.box;
    }
    return 1;
}

}  // extern "C"
