// netgen.h -- synthetic CNT/GNP networks and the DEA system assembly that feeds the solver.
//
// Host-side C++ (no GPU code).  It produces the same inputs the reference's
// Direct_Electrifying::Compute_voltage_field builds from Hoshen_Kopelman and Cutoff_Wins
// output (Direct_Electrifying.cpp:11-113): the SSS lower triangle, the diagonal, the
// right-hand side b = -K_EF^T VEF, the prescribed voltages VEF, the point->node map
// (LMM_cnts) and the ramp warm start, plus the boundary element lists that
// Electrical_analysis::Calculate_parallel_resistor sums currents over.
#pragma once
#include <stdint.h>

#include "cicy_assembly.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cicy_net_params {
    double box;            // cube side of the RVE, micrometres; the RVE is [0,box]^3
    int64_t n_cnt;         // number of straight tubes
    double radius;         // tube radius (um); 0.005 for 10 nm diameter
    double len_min;        // tube length ~ U[len_min, len_max] (um)
    double len_max;
    double seg_len;        // polyline segment length (um), 0.02
    double d_tunnel;       // tunnelling cutoff (um), 0.0014
    double d_vdw;          // van der Waals clamp (um), 0.00034
    uint64_t seed;         // mt19937_64 seed
    int per_point_nodes;   // 1: every polyline point is an element point (node);
                           // 0: element points are tube ends + junction points (reference rule)
    int64_t n_gnp;         // number of nanoplatelets (square plates) to place
    double gnp_side;       // plate side (um)
    double gnp_thick;      // plate thickness (um)
    int min_points;        // contact points of one tube on one plate closer than this many
                           // points apart form one junction (Cutoff_dist::min_points)
} cicy_net_params;

// cicy_elec_params (the element laws' parameters) is declared in cicy_assembly.h.

typedef struct cicy_net cicy_net;

// Assembled DEA system for one (cluster, R_flag, family).  All arrays are owned by the
// view and freed with cicy_sss_free.
typedef struct cicy_sss_view {
    int64_t nodes;          // global nodes (reserved + free)
    int64_t reserved;       // R_d
    int64_t n;              // free nodes = nodes - reserved
    int64_t nnz;            // stored strictly-lower entries L
    int64_t *row_ptr;       // [n+1]
    int64_t *col_ind;       // [nnz]
    double *values;         // [nnz]  (-G)
    double *diag;           // [n]
    double *b;              // [n]    -K_EF^T VEF (zero-guess residual)
    double *vef;            // [reserved]
    double *x0;             // [n]    ramp guess (zeros when the reference uses none)
    double *r0;             // [n]    b - A x0 (equals b when x0 == 0)
    int has_guess;          // 1 when R_flag==1 && R_d==2 (Direct_Electrifying.cpp:80)
    int64_t n_points;
    int64_t *lmm;           // [n_points] point -> global node or -1
    int64_t *node_point;    // [nodes]   global node -> representative point (-1 reserved)
    int64_t ne[2];          // boundary element counts for b1, b2 of `direction`
    int64_t *e_bnd[2];      // boundary-side global node of each element
    int64_t *e_int[2];      // other-end global node
    double *e_re[2];        // element resistance (Ohm)
    int direction;          // current direction used for the element lists (-1: none)
    double *node_xyz;       // [nodes*3] position of each node's point (NaN for reserved)
    int64_t gnp_nodes;      // nodes that are nanoplatelet points
    int64_t n_gnp_points;
    int64_t *lmm_gnp;       // [n_gnp_points] plate point -> global node or -1 (LMM_gnps)
    void *priv;             // owner of the arrays (internal)
} cicy_sss_view;

cicy_net *cicy_net_generate(const cicy_net_params *p);
void cicy_net_free(cicy_net *net);

// what: 0 points, 1 tubes, 2 junctions, 3 clusters, 4 plates, 5 plate points,
// 6 tube-plate junctions, 7 plate-plate junctions
int64_t cicy_net_count(const cicy_net *net, int what);
const double *cicy_net_xyz(const cicy_net *net);          // [points*3]
const int64_t *cicy_net_cnt_ptr(const cicy_net *net);     // [tubes+1]
// Cluster k: number of tubes and percolation family (-1 when it does not percolate).
int cicy_net_cluster(const cicy_net *net, int64_t k, int64_t *n_tubes, int *family);
// Junction j: P1, P2 (points), N1, N2 (tubes), gap distance (um)
int cicy_net_junction(const cicy_net *net, int64_t j, int64_t *p1, int64_t *p2,
                      int64_t *n1, int64_t *n2, double *dist);
int64_t cicy_net_boundary_points(const cicy_net *net, int b, const int64_t **pts);
// Tubes touching face b (the reference's boundary_cnt[b], Cutoff_Wins.h:21-35).
int64_t cicy_net_boundary_tubes(const cicy_net *net, int b, const int64_t **tubes);
// Tubes of cluster k (ascending) and its junction indices (ascending).
int64_t cicy_net_cluster_tubes(const cicy_net *net, int64_t k, const int64_t **tubes);
int64_t cicy_net_cluster_junctions(const cicy_net *net, int64_t k, const int64_t **junc);
// Element points of tube t (the reference's elements_cnt[t], ascending).
int64_t cicy_net_elements(const cicy_net *net, int64_t t, const int64_t **pts);

// ---- nanoplatelets (GNPs).  Plate points form their own numbering (the reference's
// points_gnp); structure_gnp[g] lists plate g's points in creation order: window-boundary
// points (Cutoff_Wins::Find_gnp_boundary_points, Cutoff_Wins.cpp:871-988), plate-plate
// junction points, tube-plate junction points (Hoshen_Kopelman::Add_mixed_junction,
// :2280-2307), minus the boundary points on faces the cluster's family does not use
// (Remove_gnp_points_in_non_relevant_boundaries, :2740-2892).
const double *cicy_net_gnp_xyz(const cicy_net *net);       // [plate points*3]
// Plate g: centre[3], axes[9] (row k = axis k; axis 2 is the normal), side, thickness.
int cicy_net_gnp(const cicy_net *net, int64_t g, double *centre, double *axes, double *side,
                 double *thick);
int64_t cicy_net_gnp_structure(const cicy_net *net, int64_t g, const int64_t **pts);
// Triangulation edges of plate g as [2*count] plate point pairs, in the reference's order
// (empty for plates outside every cluster).
int64_t cicy_net_gnp_triangulation(const cicy_net *net, int64_t g, const int64_t **edges);
int64_t cicy_net_gnp_boundary_points(const cicy_net *net, int b, const int64_t **pts);
int64_t cicy_net_boundary_gnps(const cicy_net *net, int b, const int64_t **gnps);
// Tube-plate junction j: tube point, plate point, tube, plate, gap (um).
int cicy_net_mixed_junction(const cicy_net *net, int64_t j, int64_t *p_cnt, int64_t *p_gnp,
                            int64_t *cnt, int64_t *gnp, double *dist);
// Plate-plate junction j: points on plate 1 and 2, plates, gap (um).
int cicy_net_gnp_junction(const cicy_net *net, int64_t j, int64_t *p1, int64_t *p2,
                          int64_t *g1, int64_t *g2, double *dist);
int64_t cicy_net_cluster_gnps(const cicy_net *net, int64_t k, const int64_t **gnps);
int64_t cicy_net_cluster_mixed_junctions(const cicy_net *net, int64_t k, const int64_t **junc);
int64_t cicy_net_cluster_gnp_junctions(const cicy_net *net, int64_t k, const int64_t **junc);

// Assemble the reduced system of cluster `k` with resistor flag `r_flag` (0 unit, 1 real)
// for `family` (0..6).  `direction` (0..2 or -1) selects which boundary element lists to
// build.  Returns NULL on a reference error condition (zero diagonal, invalid family,
// element with < 2 points) and writes a message into err (may be NULL).
cicy_sss_view *cicy_net_assemble(cicy_net *net, int64_t k, int r_flag, int family,
                                 const cicy_elec_params *ep, int direction, char *err,
                                 int64_t err_len);
void cicy_sss_free(cicy_sss_view *v);

// The network as flat arrays (cicy_assembly.h), for cicy_dnet_create.  The pointers stay
// valid until the next call, a change of the network (cicy_net_backbone) or cicy_net_free.
int cicy_net_arrays_get(cicy_net *net, cicy_net_arrays *out);

// Backbone of cluster k (Backbone_Network::Determine_backbone_network,
// Backbone_Network.cpp:10-61), from the voltages [nodes] of its unit-resistor solve (R_flag
// 0: VEF first) and that assembly's point -> node maps (lmm over the tube points, lmm_gnp
// over the plate points; reserved = its reserved nodes).  Edits the cluster as the
// reference does: dead tube ends and tubes, junctions with dead particles, plate junction
// points; the plates are triangulated again, so the next cicy_net_assemble gives the
// backbone's system.  avoid_resistance: volumes only, the tubes' elements and tube
// junctions stay.  zero_in (may be NULL): the zero current when computed elsewhere;
// seg_in (may be NULL): the backbone run (first, end element, -1 when none) of each
// cluster tube, in cluster order, when computed elsewhere.  out[5]: zero current, backbone
// and dead tube volume, backbone and dead plate volume.  0 on an error (message in err).
int cicy_net_backbone(cicy_net *net, int64_t k, const double *volt, int64_t nvolt,
                      const int64_t *lmm, int64_t n_points, const int64_t *lmm_gnp,
                      int64_t n_gnp_points, int reserved, int avoid_resistance,
                      const double *zero_in, const int32_t *seg_in, double *out, char *err,
                      int64_t err_len);

// Tunnelling constants from the mantissas the reference's input file gives
// (Input_Reader.cpp:1450-1456).
void cicy_tunnel_constants(double h, double e, double m, double lambda, double *c1,
                           double *c2);

#ifdef __cplusplus
}
#endif
