// dea_assembly.hip -- a cluster's DEA system assembled on the device from a network in HBM.
//
// The reference builds a cluster's system in Direct_Electrifying::Compute_voltage_field
// (Direct_Electrifying.cpp:11-113): point -> node maps (LM_matrix_for_cnts :139-215,
// Map_points_at_boundaries :217-247, LM_matrix_for_gnps :307-349), one std::map per row
// filled resistor by resistor (Fill_2d_matrices_* :489-1327; a new resistor overwrites the
// entry, one added to a reserved node's entry subtracts), the maps walked in key order into
// the SSS arrays and b (From_2d_to_1d_vectors :1329-1394), the ramp guess and its residual
// (:1396-1535).  Here, one network copy in HBM serves every (cluster, family, direction):
//
//   1. node numbers: reserved faces marked in the family's order, then per-tube counts and a
//      prefix scan number the tubes' element points in cluster order, and per-plate counts
//      and a scan number the plates' points;
//   2. every resistor becomes a record at its slot in the reference's insertion order (tube
//      elements, tube junctions, tube-plate junctions, plate triangulation edges from the
//      last, plate junctions): key (row, col), value, rule; and two diagonal contributions;
//   3. a stable radix sort of the records on (row, col) gives each row's map in key order
//      with the insertion order kept inside a key, so one thread per row folds each key's
//      records with the reference's rule, writes the row's entries and sums its b terms in
//      column order; a stable radix sort of the diagonal contributions by node gives each
//      node its insertion-order sum;
//   4. the ramp guess (one thread per tube) and r0 = b - A x0 with the reference's loop order
//      for each node: diagonal, its own row, then the rows that hold it as a column in
//      ascending order (a stable sort of the entries by column);
//   5. the boundary element lists of the direction (counts, a scan, writes in list order).
//
// Every floating-point step is the host assembler's (cicy_net_assemble) in the same order
// with contraction off, so the results equal the host's bit for bit except where the
// tunnelling law's exp() enters: the device library's exp, like the host's, is within an ulp
// of the exact value but not always the same double.
#include "dea_internal.h"
#include "cicy_assembly.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <vector>

#pragma clang fp contract(off)

using namespace cicy;

namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kZero = 1.0e-10;
constexpr int kAsmThreads = 256;
constexpr long long kNone = LLONG_MAX;
constexpr int kWave = 64;
constexpr int kTubesPerBlock = kAsmThreads / kWave;  // one wave per tube

inline unsigned grid_for(long long n) {
    return (unsigned)std::max<long long>(1, (n + kAsmThreads - 1) / kAsmThreads);
}

// work-groups for one wave per tube
inline unsigned grid_tubes(long long T) {
    return (unsigned)std::max<long long>(1, (T + kTubesPerBlock - 1) / kTubesPerBlock);
}

// the network's arrays on the device
struct NetDev {
    const double *xyz, *radius, *gxyz, *thick;
    const int64_t *elem_ptr, *elem_pts, *gstruct_ptr, *gstruct, *gtri_ptr, *gtri;
    const cicy_junction_rec *junc, *mixed, *gjunc;
};

// error words: the first offending item of each kind, in the host assembler's order
struct AsmErr {
    long long bad_tube;  // cluster tube index of a tube with fewer than two element points
    long long gnp;       // 2 * slot + (0: vertex without a node, 1: zero-length edge)
    long long bnd;       // (w << 61 | plate << 30 | edge << 1) + (0: both ends at one
                         // boundary, 1: zero-length edge)
    long long floating;  // free node with a zero diagonal
};

struct Vef {
    double v[8];
};

// Calculate_resistance_cnt (Direct_Electrifying.cpp:611-646): the polyline length between
// two element points of one tube, over the points between them
__device__ double resistance_cnt(const double *x, int64_t p1, int64_t p2, double radius,
                                 double rho) {
    double length = 0;
    for (int64_t i = p1; i < p2; i++) {
        double dx = x[3 * i] - x[3 * (i + 1)];
        double dy = x[3 * i + 1] - x[3 * (i + 1) + 1];
        double dz = x[3 * i + 2] - x[3 * (i + 1) + 2];
        length = length + sqrt(dx * dx + dy * dy + dz * dz);
    }
    return rho * length / (kPi * radius * radius);
}

// Calculate_junction_resistance (Direct_Electrifying.cpp:854-901); l is a tube's radius
// (doubled) or a plate's thickness
__device__ double junction_resistance(double dist, double l1, bool cnt1, double l2, bool cnt2,
                                      const cicy_elec_params &ep) {
    if (!ep.junction_exponential) return ep.junction_resistance;
    double jl1 = l1, jl2 = l2;
    if (cnt1) jl1 = jl1 + l1;
    if (cnt2) jl2 = jl2 + l2;
    double a = (jl1 - jl2 < kZero) ? jl1 * jl1 : jl2 * jl2;
    return ep.c1 * dist * exp(ep.c2 * dist) / a;
}

// Calculate_resistance_gnp (Direct_Electrifying.cpp:1198-1232); false for a zero-length edge
__device__ bool resistance_gnp(const double *p1, const double *p2, double rad1, double rad2,
                               const cicy_elec_params &ep, double *re) {
    const double dx = p1[0] - p2[0], dy = p1[1] - p2[1], dz = p1[2] - p2[2];
    const double len = sqrt(dx * dx + dy * dy + dz * dz);
    if (len < kZero) return false;
    const double ux = (p2[0] - p1[0]) / len, uy = (p2[1] - p1[1]) / len,
                 uz = (p2[2] - p1[2]) / len;
    const double rx = ux * ep.resistivity_gnp_surf, ry = uy * ep.resistivity_gnp_surf,
                 rz = uz * ep.resistivity_gnp_t;
    const double rho = rx * ux + ry * uy + rz * uz;
    *re = rho * len / (kPi * rad1 * rad2);
    return true;
}

__global__ void k_fill64(int64_t *p, long long n, int64_t v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill32(int *p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_iota(int *p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

__global__ void k_err_init(AsmErr *e) {
    if (threadIdx.x == 0) {
        e->bad_tube = kNone;
        e->gnp = kNone;
        e->bnd = kNone;
        e->floating = kNone;
    }
}

// Map_points_at_boundaries: the face's points take reserved node `value`
__global__ void k_mark(int64_t *lmm, const int64_t *pts, long long count, int64_t value) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) lmm[pts[i]] = value;
}

// per cluster tube: new nodes (the end points only when no face reserved them) and
// resistors (one for the three-point element whose ends share a reserved node); entry T = 0.
// ends[i]: bit 0 the first end is new, bit 1 the last end is new.
__global__ void k_tube_count(NetDev d, const int64_t *ct, long long T, const int64_t *lmm,
                             int64_t *nodes, int64_t *edges, unsigned char *ends, AsmErr *err) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > T) return;
    if (i == T) {
        nodes[i] = 0;
        edges[i] = 0;
        return;
    }
    const int64_t t = ct[i];
    const int64_t a = d.elem_ptr[t], b = d.elem_ptr[t + 1];
    const int64_t ne = b - a;
    if (ne <= 1) {
        atomicMin(&err->bad_tube, i);
        nodes[i] = 0;
        edges[i] = 0;
        ends[i] = 0;
        return;
    }
    const int64_t l0 = lmm[d.elem_pts[a]], l1 = lmm[d.elem_pts[b - 1]];
    nodes[i] = (ne - 2) + (l0 < 0 ? 1 : 0) + (l1 < 0 ? 1 : 0);
    edges[i] = (ne == 3 && l0 >= 0 && l0 == l1) ? 1 : ne - 1;
    ends[i] = (unsigned char)((l0 < 0 ? 1 : 0) | (l1 < 0 ? 2 : 0));
}

// LM_matrix_for_cnts: the tube's points from its first node on, in element order (one wave
// per tube, a lane per element point)
__global__ void __launch_bounds__(kAsmThreads)
k_tube_number(NetDev d, const int64_t *ct, long long T, int64_t *lmm, const int64_t *base,
              const unsigned char *ends, int64_t first) {
    const long long i = (long long)blockIdx.x * kTubesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (i >= T) return;
    const int64_t t = ct[i];
    const int64_t a = d.elem_ptr[t], b = d.elem_ptr[t + 1];
    const int64_t ne = b - a;
    if (ne <= 1) return;
    const int64_t g0 = first + base[i];
    const int e = ends[i];
    const int64_t off = (e & 1) ? 1 : 0;
    for (int64_t q = lane; q < ne; q += kWave) {
        const int64_t p = d.elem_pts[a + q];
        if (q == 0) {
            if (e & 1) lmm[p] = g0;
        } else if (q == ne - 1) {
            if (e & 2) lmm[p] = g0 + off + (ne - 2);
        } else {
            lmm[p] = g0 + off + (q - 1);
        }
    }
}

// per cluster plate (one work-group each): its points that no face reserved, and its
// triangulation edges; entry G = 0 (written by the last work-group)
__global__ void __launch_bounds__(kAsmThreads)
k_gnp_count(NetDev d, const int64_t *cg, long long G, const int64_t *lmm_g, int64_t *nodes,
            int64_t *tris) {
    typedef hipcub::BlockReduce<int, kAsmThreads> Reduce;
    __shared__ typename Reduce::TempStorage tmp;
    const long long i = blockIdx.x;
    if (i == G) {
        if (threadIdx.x == 0) {
            nodes[i] = 0;
            tris[i] = 0;
        }
        return;
    }
    const int64_t gp = cg[i];
    int c = 0;
    for (int64_t q = d.gstruct_ptr[gp] + threadIdx.x; q < d.gstruct_ptr[gp + 1]; q += blockDim.x)
        c += lmm_g[d.gstruct[q]] < 0 ? 1 : 0;
    c = Reduce(tmp).Sum(c);
    if (threadIdx.x == 0) {
        nodes[i] = c;
        tris[i] = (d.gtri_ptr[gp + 1] - d.gtri_ptr[gp]) / 2;
    }
}

// LM_matrix_for_gnps: the plate's unreserved points numbered in structure order (a block
// scan per chunk of points)
__global__ void __launch_bounds__(kAsmThreads)
k_gnp_number(NetDev d, const int64_t *cg, long long G, int64_t *lmm_g, const int64_t *base,
             int64_t first) {
    typedef hipcub::BlockScan<int, kAsmThreads> Scan;
    __shared__ typename Scan::TempStorage tmp;
    const long long i = blockIdx.x;
    if (i >= G) return;
    const int64_t gp = cg[i];
    int64_t g = first + base[i];
    for (int64_t q0 = d.gstruct_ptr[gp]; q0 < d.gstruct_ptr[gp + 1]; q0 += blockDim.x) {
        const int64_t q = q0 + threadIdx.x;
        const bool in = q < d.gstruct_ptr[gp + 1];
        const int64_t p = in ? d.gstruct[q] : 0;
        const int f = (in && lmm_g[p] < 0) ? 1 : 0;
        int off, total;
        Scan(tmp).ExclusiveSum(f, off, total);
        if (f) lmm_g[p] = g + off;
        g += total;
        __syncthreads();
    }
}

// points_cnt_rad: the last (in cluster order) kept tube-plate junction at a plate point
__global__ void k_mixed_owner(NetDev d, const int64_t *cm, long long M, const int64_t *lmm,
                              const int64_t *lmm_g, int *owner) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const cicy_junction_rec J = d.mixed[cm[i]];
    if (lmm[J.p1] < 0 || lmm_g[J.p2] < 0) return;
    atomicMax(&owner[J.p2], (int)i);
}

struct Recs {
    unsigned long long *key;  // row << cb | col; dropped: nodes << cb
    double *val;
    unsigned char *rule;      // 0 overwrite (-g), 1 subtract g
    int *dnode;               // two diagonal contributions per slot (dropped: nodes)
    double *dval;
    int cb;
    int64_t nodes;
};

__device__ inline void put_rec(const Recs &r, long long slot, int64_t row, int64_t col, double v,
                               int rule) {
    r.key[slot] = ((unsigned long long)row << r.cb) | (unsigned long long)col;
    r.val[slot] = v;
    r.rule[slot] = (unsigned char)rule;
}

__device__ inline void put_diag(const Recs &r, long long slot, int64_t n1, int64_t n2, double g) {
    r.dnode[2 * slot] = (int)n1;
    r.dval[2 * slot] = g;
    r.dnode[2 * slot + 1] = (int)n2;
    r.dval[2 * slot + 1] = g;
}

__device__ inline void drop_rec(const Recs &r, long long slot) {
    r.key[slot] = (unsigned long long)r.nodes << r.cb;
    r.val[slot] = 0.0;
    r.rule[slot] = 0;
    put_diag(r, slot, r.nodes, r.nodes, 0.0);
}

// Add_new_elements_to_2d_sparse_matrix (:648-669) and Add_to_existing_elements_in_2d_
// sparse_matrix (:671-694): the entry sits in the larger node's row
__device__ inline void add_edge(const Recs &r, long long slot, int64_t n1, int64_t n2, double g,
                                bool existing) {
    put_diag(r, slot, n1, n2, g);
    const int64_t hi = n1 > n2 ? n1 : n2, lo = n1 > n2 ? n2 : n1;
    if (existing) put_rec(r, slot, hi, lo, g, 1);
    else put_rec(r, slot, hi, lo, -g, 0);
}

// Fill_2d_matrices_cnts (:489-553) with Three_points_in_element (:555-609): one wave per
// tube, a lane per element (consecutive slots, so the records are written coalesced)
__global__ void __launch_bounds__(kAsmThreads)
k_edges_tubes(NetDev d, const int64_t *ct, long long T, const int64_t *lmm, const int64_t *base,
              int r_flag, cicy_elec_params ep, Recs r) {
    const long long i = (long long)blockIdx.x * kTubesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (i >= T) return;
    const int64_t t = ct[i];
    const int64_t a = d.elem_ptr[t], b = d.elem_ptr[t + 1];
    const int64_t *e = d.elem_pts + a;
    const int64_t ne = b - a;
    const double rad = d.radius[t];
    const long long slot0 = base[i];
    // the three-point element (its ends share a reserved node) took one slot in the count;
    // deciding from the slots keeps the writes inside them whatever the network holds
    if (ne == 3 && base[i + 1] - base[i] == 1) {
        if (lane != 0) return;
        double g1 = 1.0, g2 = 1.0;
        if (r_flag == 1) {
            g1 = 1 / resistance_cnt(d.xyz, e[0], e[1], rad, ep.resistivity_cnt);
            g2 = 1 / resistance_cnt(d.xyz, e[1], e[2], rad, ep.resistivity_cnt);
        }
        const int64_t n1 = lmm[e[0]], n2 = lmm[e[1]];
        const double s = g1 + g2;
        r.dnode[2 * slot0] = (int)n1;
        r.dval[2 * slot0] = s;
        r.dnode[2 * slot0 + 1] = (int)n2;
        r.dval[2 * slot0 + 1] = s;
        put_rec(r, slot0, n2, n1, -(g1 + g2), 0);
        return;
    }
    for (int64_t q = 1 + lane; q < ne; q += kWave) {
        const long long slot = slot0 + q - 1;
        double re = (r_flag == 1) ? resistance_cnt(d.xyz, e[q - 1], e[q], rad, ep.resistivity_cnt)
                                  : 1.0;
        double g = (r_flag == 1) ? 1 / re : re;
        add_edge(r, slot, lmm[e[q - 1]], lmm[e[q]], g, false);
    }
}

// Fill_2d_matrices_cnt_junctions (:737-852): junctions with a point outside the cluster are
// skipped; one at a reserved node adds to the existing entry when R_d > 2, else is skipped
__global__ void k_edges_junc(NetDev d, const int64_t *cj, long long J, const int64_t *lmm,
                             long long slot0, int64_t rd, int r_flag, cicy_elec_params ep,
                             Recs r) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= J) return;
    const cicy_junction_rec jr = d.junc[cj[i]];
    const long long slot = slot0 + i;
    const int64_t n1 = lmm[jr.p1], n2 = lmm[jr.p2];
    if (n1 < 0 || n2 < 0) {
        drop_rec(r, slot);
        return;
    }
    double g = 1.0;
    if (r_flag == 1)
        g = 1 / junction_resistance(jr.dist, d.radius[jr.e1], true, d.radius[jr.e2], true, ep);
    if (n1 >= rd && n2 >= rd) add_edge(r, slot, n1, n2, g, false);
    else if (rd > 2) add_edge(r, slot, n1, n2, g, true);
    else drop_rec(r, slot);
}

// Fill_2d_matrices_mixed_junctions (:903-1019)
__global__ void k_edges_mixed(NetDev d, const int64_t *cm, long long M, const int64_t *lmm,
                              const int64_t *lmm_g, long long slot0, int64_t rd, int r_flag,
                              cicy_elec_params ep, Recs r) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const cicy_junction_rec jr = d.mixed[cm[i]];
    const long long slot = slot0 + i;
    const int64_t n1 = lmm[jr.p1], n2 = lmm_g[jr.p2];
    if (n1 < 0 || n2 < 0) {
        drop_rec(r, slot);
        return;
    }
    double g = 1.0;
    if (r_flag == 1)
        g = 1 / junction_resistance(jr.dist, d.radius[jr.e1], true, d.thick[jr.e2], false, ep);
    if (n1 >= rd && n2 >= rd) add_edge(r, slot, n1, n2, g, false);
    else if (rd > 2) add_edge(r, slot, n1, n2, g, true);
    else drop_rec(r, slot);
}

// Fill_2d_matrices_gnp (:1021-1196): each plate's triangulation edges from the last to the
// first (one work-group per plate, edge j at slot base + ntri - 1 - j); an end without a tube
// junction takes half the plate thickness as its radius
__global__ void __launch_bounds__(kAsmThreads)
k_edges_gnp(NetDev d, const int64_t *cg, long long G, const int64_t *lmm_g, const int *owner,
            const int64_t *cm, const int64_t *base, long long slot0, int r_flag,
            cicy_elec_params ep, Recs r, AsmErr *err) {
    const long long i = blockIdx.x;
    if (i >= G) return;
    const int64_t gp = cg[i];
    const int64_t *tri = d.gtri + d.gtri_ptr[gp];
    const int64_t ntri = (d.gtri_ptr[gp + 1] - d.gtri_ptr[gp]) / 2;
    const double half_t = d.thick[gp] / 2;
    for (int64_t j = threadIdx.x; j < ntri; j += blockDim.x) {
        const long long slot = slot0 + base[i] + (ntri - 1 - j);
        const int64_t v1 = tri[2 * j], v2 = tri[2 * j + 1];
        const int64_t n1 = lmm_g[v1], n2 = lmm_g[v2];
        if (n1 < 0 || n2 < 0) {
            atomicMin(&err->gnp, 2 * slot);
            drop_rec(r, slot);
            continue;
        }
        double g = 1.0;
        if (r_flag == 1) {
            const int o1 = owner[v1], o2 = owner[v2];
            const double rad1 = o1 >= 0 ? d.radius[d.mixed[cm[o1]].e1] : half_t;
            const double rad2 = o2 >= 0 ? d.radius[d.mixed[cm[o2]].e1] : half_t;
            double re;
            if (!resistance_gnp(d.gxyz + 3 * v1, d.gxyz + 3 * v2, rad1, rad2, ep, &re)) {
                atomicMin(&err->gnp, 2 * slot + 1);
                drop_rec(r, slot);
                continue;
            }
            g = 1 / re;
        }
        add_edge(r, slot, n1, n2, g, false);
    }
}

// Fill_2d_matrices_gnp_junctions (:1234-1327): no reserved-node rule
__global__ void k_edges_gjunc(NetDev d, const int64_t *cgj, long long GJ, const int64_t *lmm_g,
                              long long slot0, int r_flag, cicy_elec_params ep, Recs r) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= GJ) return;
    const cicy_junction_rec jr = d.gjunc[cgj[i]];
    const long long slot = slot0 + i;
    const int64_t n1 = lmm_g[jr.p1], n2 = lmm_g[jr.p2];
    if (n1 < 0 || n2 < 0) {
        drop_rec(r, slot);
        return;
    }
    double g = 1.0;
    if (r_flag == 1)
        g = 1 / junction_resistance(jr.dist, d.thick[jr.e1], false, d.thick[jr.e2], false, ep);
    add_edge(r, slot, n1, n2, g, false);
}

// first sorted record of each row
__global__ void k_row_heads(const unsigned long long *skey, long long S, int cb, int64_t nodes,
                            int *row_start) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    const long long row = (long long)(skey[i] >> cb);
    if (row >= nodes) return;
    if (i == 0 || (long long)(skey[i - 1] >> cb) != row) row_start[row] = (int)i;
}

// From_2d_to_1d_vectors (:1329-1394): one thread per free row walks its keys in column
// order; each key's records fold with the reference's rule; columns of reserved nodes go
// into b.  WRITE = false counts the row's entries (entry n = 0), WRITE = true writes them.
template <bool WRITE>
__global__ void k_fold(const unsigned long long *skey, const int *sslot, const double *val,
                       const unsigned char *rule, long long S, const int *row_start, int cb,
                       int64_t rd, int64_t nodes, Vef vef, int64_t *cnt, const int64_t *row_ptr,
                       int64_t *col_ind, double *values, int *erow, int *col32, double *b) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = nodes - rd;
    if (t > n) return;
    if (t == n) {
        if (!WRITE) cnt[n] = 0;
        return;
    }
    const long long row = rd + t;
    const unsigned long long cmask = (1ull << cb) - 1;
    long long i = row_start[row];
    int64_t pos = WRITE ? row_ptr[t] : 0;
    int64_t c = 0;
    double bb = 0.0;
    if (i >= 0) {
        while (i < S && (long long)(skey[i] >> cb) == row) {
            const unsigned long long key = skey[i];
            const int64_t col = (int64_t)(key & cmask);
            double v = 0.0;
            bool have = false;
            while (i < S && skey[i] == key) {
                const int s = sslot[i];
                if (rule[s] == 0) v = val[s];
                else v = (have ? v : 0.0) - val[s];
                have = true;
                i++;
            }
            if (col >= rd) {
                if (WRITE) {
                    col_ind[pos] = col - rd;
                    values[pos] = v;
                    if (erow) {
                        erow[pos] = (int)t;
                        col32[pos] = (int)(col - rd);
                    }
                }
                pos++;
                c++;
            } else {
                bb = bb - (v * vef.v[col]);
            }
        }
    }
    if (WRITE) b[t] = bb;
    else cnt[t] = c;
}

// first index of each key below `limit` in a sorted int array
__global__ void k_heads(const int *keys, long long count, long long limit, int *start) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int k = keys[i];
    if (k < limit && (i == 0 || keys[i - 1] != k)) start[k] = (int)i;
}

// each free node's diagonal: its contributions in insertion order from 0
__global__ void k_diag(const int *sdnode, const double *sdval, const int *dstart, long long S2,
                       int64_t rd, int64_t nodes, double *diag, AsmErr *err) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nodes - rd) return;
    const int node = (int)(rd + t);
    double s = 0.0;
    const int lo = dstart[node];
    if (lo >= 0)
        for (long long i = lo; i < S2 && sdnode[i] == node; i++) s += sdval[i];
    diag[t] = s;
    if (fabs(s) < kZero) atomicMin(&err->floating, t);
}

// Initial_guess_for_cnt_nodes (:1434-1470): V_app / L * (max - coordinate) at each free tube
// node (one wave per tube, a lane per element point)
__global__ void __launch_bounds__(kAsmThreads)
k_ramp(NetDev d, const int64_t *ct, long long T, const int64_t *lmm, int64_t rd, int axis,
       double vpc, double max_coord, double *x0) {
    const long long i = (long long)blockIdx.x * kTubesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (i >= T) return;
    const int64_t t = ct[i];
    for (int64_t q = d.elem_ptr[t] + lane; q < d.elem_ptr[t + 1]; q += kWave) {
        const int64_t pt = d.elem_pts[q];
        const int64_t nd = lmm[pt];
        if (nd >= 2) {
            const double pc = d.xyz[3 * pt + axis];
            x0[nd - rd] = vpc * (max_coord - pc);
        }
    }
}

// Update_residual_vector (:1502-1535): r0 = b - A x0, each node's product summed as the
// reference's row loop reaches it: diagonal, own row, then the later rows holding it
__global__ void k_r0(const int64_t *row_ptr, const int64_t *col_ind, const double *values,
                     const double *diag, const double *x0, const double *b, const int *scol,
                     const int *sent, const int *cstart, const int *erow, long long nnz,
                     int64_t n, double *r0) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double ax = diag[i] * x0[i];
    for (int64_t j = row_ptr[i]; j < row_ptr[i + 1]; j++) ax = ax + values[j] * x0[col_ind[j]];
    const int lo = cstart[i];
    if (lo >= 0)
        for (long long e = lo; e < nnz && scol[e] == i; e++) {
            const int j = sent[e];
            ax = ax + values[j] * x0[erow[j]];
        }
    r0[i] = b[i] - ax;
}

struct BndDev {
    const int64_t *tubes;  // the face's tubes (boundary_cnt)
    long long nt;
    const int64_t *gnps;   // the face's plates (boundary_gnp)
    long long ng;
    const int64_t *cg;     // the cluster's plates, ascending
    long long G;
    int w;                 // 0: b1 (node 0), 1: b2 (node 1)
};

__device__ bool in_cluster(const BndDev &bd, int64_t gp) {
    long long lo = 0, hi = bd.G;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (bd.cg[mid] < gp) lo = mid + 1;
        else hi = mid;
    }
    return lo < bd.G && bd.cg[lo] == gp;
}

// Currents_through_boundary_cnts (Electrical_analysis.cpp:499-545): each face tube's first
// then last element when its end node is 0 or 1 (one thread per tube).  WRITE = false
// counts per tube (entry nt + ng = 0, written here), WRITE = true writes at the scanned
// positions.
template <bool WRITE>
__global__ void k_bnd_tubes(NetDev d, BndDev bd, const int64_t *lmm, cicy_elec_params ep,
                            int64_t *cnt, const int64_t *pos0, int64_t *e_bnd, int64_t *e_int,
                            double *e_re) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > bd.nt) return;
    if (i == bd.nt) {
        if (!WRITE) cnt[bd.nt + bd.ng] = 0;
        return;
    }
    int64_t c = 0;
    int64_t pos = WRITE ? pos0[i] : 0;
    const int64_t t = bd.tubes[i];
    const int64_t a = d.elem_ptr[t], b = d.elem_ptr[t + 1];
    if (b - a >= 2) {
        const int64_t *e = d.elem_pts + a;
        const int64_t fr[2][2] = {{e[0], e[1]}, {e[b - a - 1], e[b - a - 2]}};
        for (int side = 0; side < 2; side++) {
            const int64_t p1 = fr[side][0], p2 = fr[side][1];
            const int64_t n1 = lmm[p1];
            if (n1 < 0 || n1 > 1) continue;
            const int64_t n2 = lmm[p2];
            if (n2 < 0) continue;
            if (WRITE) {
                const double rad = d.radius[t];
                const double re = p1 > p2 ? resistance_cnt(d.xyz, p2, p1, rad, ep.resistivity_cnt)
                                          : resistance_cnt(d.xyz, p1, p2, rad, ep.resistivity_cnt);
                e_bnd[pos] = n1;
                e_int[pos] = n2;
                e_re[pos] = re;
                pos++;
            }
            c++;
        }
    }
    if (!WRITE) cnt[i] = c;
}

// Currents_through_boundary_gnps (Electrical_analysis.cpp:607-675): the triangulation edges
// of the face's plates in the cluster that touch node 0 or 1, oriented from node w (one
// work-group per face plate; positions inside a plate by a block scan per chunk of edges).
// The count pass records the first error in the host's order.
template <bool WRITE>
__global__ void __launch_bounds__(kAsmThreads)
k_bnd_plates(NetDev d, BndDev bd, const int64_t *lmm_g, const int *owner, const int64_t *cm,
             cicy_elec_params ep, int64_t *cnt, const int64_t *pos0, int64_t *e_bnd,
             int64_t *e_int, double *e_re, AsmErr *err) {
    typedef hipcub::BlockScan<int, kAsmThreads> Scan;
    __shared__ typename Scan::TempStorage tmp;
    const long long ip = blockIdx.x;
    if (ip >= bd.ng) return;
    const int64_t gp = bd.gnps[ip];
    int64_t pos = WRITE ? pos0[bd.nt + ip] : 0;
    if (!in_cluster(bd, gp)) {
        if (!WRITE && threadIdx.x == 0) cnt[bd.nt + ip] = 0;
        return;
    }
    const int64_t node = bd.w;
    const int64_t *tri = d.gtri + d.gtri_ptr[gp];
    const int64_t ntri = (d.gtri_ptr[gp + 1] - d.gtri_ptr[gp]) / 2;
    const double half_t = d.thick[gp] / 2;
    for (int64_t j0 = 0; j0 < ntri; j0 += blockDim.x) {
        const int64_t j = j0 + threadIdx.x;
        int f = 0;
        int64_t n1 = 0, n2 = 0;
        double re = 0.0;
        if (j < ntri) {
            const int64_t v1 = tri[2 * j], v2 = tri[2 * j + 1];
            const int64_t na = lmm_g[v1], nb = lmm_g[v2];
            if (!(na > 1 && nb > 1)) {
                const long long ekey = ((long long)bd.w << 61) | ((long long)ip << 30) |
                                       ((long long)j << 1);
                if (na == nb) {
                    if (!WRITE) atomicMin(&err->bnd, ekey);
                } else {
                    n1 = (na == node) ? na : nb;
                    n2 = (na == node) ? nb : na;
                    const int o1 = owner[v1], o2 = owner[v2];
                    const double rad1 = o1 >= 0 ? d.radius[d.mixed[cm[o1]].e1] : half_t;
                    const double rad2 = o2 >= 0 ? d.radius[d.mixed[cm[o2]].e1] : half_t;
                    if (resistance_gnp(d.gxyz + 3 * v1, d.gxyz + 3 * v2, rad1, rad2, ep, &re))
                        f = 1;
                    else if (!WRITE)
                        atomicMin(&err->bnd, ekey | 1);
                }
            }
        }
        int off, total;
        Scan(tmp).ExclusiveSum(f, off, total);
        if (WRITE && f) {
            e_bnd[pos + off] = n1;
            e_int[pos + off] = n2;
            e_re[pos + off] = re;
        }
        pos += total;
        __syncthreads();
    }
    if (!WRITE && threadIdx.x == 0) cnt[bd.nt + ip] = pos;
}

// up to 8 int64 entries of device arrays, gathered for one read-back
struct Pick {
    const int64_t *p[8];
    long long at[8];
};

__global__ void k_pick(Pick pk, int k, int64_t *out) {
    const int i = threadIdx.x;
    if (i < k) out[i] = pk.p[i][pk.at[i]];
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

struct cicy_dnet {
    cicy_ctx *ctx = nullptr;
    int64_t n_points = 0, n_tubes = 0, n_junc = 0, n_gpts = 0, n_gnps = 0, n_mixed = 0,
            n_gjunc = 0, n_clusters = 0;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    NetDev d{};
    // face and cluster lists: entries on the device, offsets on the host
    const int64_t *faces[4] = {nullptr, nullptr, nullptr, nullptr};  // bpts bcnt gbpts gbnd
    std::vector<int64_t> face_ptr[4];
    const int64_t *lists[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<int64_t> list_ptr[5];  // tubes junctions plates mixed plate-junctions
    std::vector<void *> owned;
};

struct cicy_asm {
    cicy_ctx *ctx = nullptr;
    cicy_asm_view v{};
    std::vector<int64_t> blocks;
    std::vector<void *> owned;
};

namespace {

// Get_vector_of_boundaries (Direct_Electrifying.cpp:249-305) and Get_global_nodes (:115-137)
std::vector<int> family_faces(int family) {
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

// host checks of a network's arrays before they are copied
struct NetCheck {
    cicy_ctx *ctx;
    bool ok = true;
    bool fail(const char *what) {
        if (ok) set_error(ctx, "cicy_dnet_create: %s", what);
        ok = false;
        return false;
    }
    bool range(const char *name, const int64_t *a, int64_t count, int64_t hi) {
        if (!ok) return false;
        if (count > 0 && !a) return fail(name);
        for (int64_t i = 0; i < count; i++)
            if (a[i] < 0 || a[i] >= hi) {
                char m[160];
                snprintf(m, sizeof(m), "%s[%lld] = %lld is outside [0, %lld)", name, (long long)i,
                         (long long)a[i], (long long)hi);
                return fail(m);
            }
        return true;
    }
    // offsets [count + 1] from 0, non-decreasing; returns the entry count (-1 on failure)
    int64_t offsets(const char *name, const int64_t *p, int64_t count) {
        if (!ok) return -1;
        if (!p) {
            fail(name);
            return -1;
        }
        if (p[0] != 0) {
            char m[160];
            snprintf(m, sizeof(m), "%s must start at 0", name);
            fail(m);
            return -1;
        }
        for (int64_t i = 0; i < count; i++)
            if (p[i + 1] < p[i]) {
                char m[160];
                snprintf(m, sizeof(m), "%s decreases at %lld", name, (long long)i);
                fail(m);
                return -1;
            }
        return p[count];
    }
    bool juncs(const char *name, const cicy_junction_rec *j, int64_t count, int64_t p1hi,
               int64_t p2hi, int64_t e1hi, int64_t e2hi) {
        if (!ok) return false;
        if (count > 0 && !j) return fail(name);
        for (int64_t i = 0; i < count; i++)
            if (j[i].p1 < 0 || j[i].p1 >= p1hi || j[i].p2 < 0 || j[i].p2 >= p2hi || j[i].e1 < 0 ||
                j[i].e1 >= e1hi || j[i].e2 < 0 || j[i].e2 >= e2hi) {
                char m[160];
                snprintf(m, sizeof(m), "%s[%lld] has an index out of range", name, (long long)i);
                return fail(m);
            }
        return true;
    }
};

template <class T>
bool upload(cicy_dnet *dn, const T *src, int64_t count, const T **dst) {
    cicy_ctx *ctx = dn->ctx;
    void *p = ctx->pool.get((size_t)std::max<int64_t>(count, 1) * sizeof(T));
    if (!p) {
        set_error(ctx, "hipMalloc of %lld bytes (network) failed",
                  (long long)(count * (int64_t)sizeof(T)));
        return false;
    }
    dn->owned.push_back(p);
    *dst = static_cast<const T *>(p);
    if (count == 0) return true;
    return hip_ok(ctx, hipMemcpyAsync(p, src, (size_t)count * sizeof(T), hipMemcpyHostToDevice,
                                      ctx->stream),
                  "H2D (network)");
}

// device helpers of one assembly: output arrays owned by the assembly, temporaries by
// the scope
struct AsmRun {
    cicy_ctx *ctx;
    cicy_asm *a;
    PoolScope &ps;
    template <class T>
    bool out(T *&p, long long count) {
        p = static_cast<T *>(ctx->pool.get((size_t)std::max<long long>(count, 1) * sizeof(T)));
        if (!p) {
            set_error(ctx, "hipMalloc of %lld bytes (assembly) failed",
                      (long long)(count * (long long)sizeof(T)));
            return false;
        }
        a->owned.push_back(p);
        return true;
    }
    bool launched(const char *what) { return hip_ok(ctx, hipGetLastError(), what); }
    bool scan(const int64_t *in, int64_t *outp, long long count) {
        size_t bytes = 0;
        if (!hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, outp, (int)count,
                                                          ctx->stream),
                    "scan"))
            return false;
        unsigned char *t = nullptr;
        if (!ps.get(t, bytes)) return false;
        return hip_ok(ctx, hipcub::DeviceScan::ExclusiveSum(t, bytes, in, outp, (int)count,
                                                            ctx->stream),
                      "scan");
    }
    template <class K, class V>
    bool sort(const K *kin, K *kout, const V *vin, V *vout, long long count, int bits) {
        if (count == 0) return true;
        size_t bytes = 0;
        if (!hip_ok(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout,
                                                            (int)count, 0, bits, ctx->stream),
                    "radix sort"))
            return false;
        unsigned char *t = nullptr;
        if (!ps.get(t, bytes)) return false;
        return hip_ok(ctx, hipcub::DeviceRadixSort::SortPairs(t, bytes, kin, kout, vin, vout,
                                                              (int)count, 0, bits, ctx->stream),
                      "radix sort");
    }
    // entries of device arrays, one wait
    bool pick(int64_t *buf, std::initializer_list<std::pair<const int64_t *, long long>> at,
              int64_t *host) {
        Pick pk{};
        int k = 0;
        for (const auto &x : at) {
            pk.p[k] = x.first;
            pk.at[k] = x.second;
            k++;
        }
        hipLaunchKernelGGL(k_pick, dim3(1), dim3(64), 0, ctx->stream, pk, k, buf);
        return launched("k_pick") && d2h_sync(ctx, host, buf, (size_t)k * sizeof(int64_t));
    }
};

bool timing_on() {
    static const bool on = std::getenv("CICY_ASM_TIMING") != nullptr;
    return on;
}

}  // namespace

extern "C" {

int cicy_dnet_create(cicy_ctx *ctx, const cicy_net_arrays *a, cicy_dnet **out) {
    if (!ctx || !a || !out) return 0;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        set_error(ctx, "hipSetDevice failed");
        return 0;
    }
    NetCheck c{ctx};
    if (a->n_points < 0 || a->n_tubes < 0 || a->n_junctions < 0 || a->n_gnp_points < 0 ||
        a->n_gnps < 0 || a->n_mixed < 0 || a->n_gjunctions < 0 || a->n_clusters < 0)
        return c.fail("negative count"), 0;
    if (a->n_points > 0 && !a->xyz) return c.fail("xyz"), 0;
    if (a->n_tubes > 0 && !a->radius) return c.fail("radius"), 0;
    if (a->n_gnp_points > 0 && !a->gxyz) return c.fail("gxyz"), 0;
    if (a->n_gnps > 0 && !a->gnp_thickness) return c.fail("gnp_thickness"), 0;
    const int64_t n_elem = c.offsets("elem_ptr", a->elem_ptr, a->n_tubes);
    c.range("elem_pts", a->elem_pts, n_elem, a->n_points);
    for (int64_t t = 0; c.ok && t < a->n_tubes; t++)
        for (int64_t q = a->elem_ptr[t] + 1; q < a->elem_ptr[t + 1]; q++)
            if (a->elem_pts[q] <= a->elem_pts[q - 1]) {
                char m[160];
                snprintf(m, sizeof(m), "element points of tube %lld are not ascending",
                         (long long)t);
                c.fail(m);
                break;
            }
    // a point belongs to one tube and a plate point to one plate, once (the numbering runs
    // per tube and per plate in parallel)
    if (c.ok) {
        std::vector<unsigned char> seen((size_t)a->n_points, 0);
        for (int64_t q = 0; q < n_elem; q++) {
            if (seen[(size_t)a->elem_pts[q]]) {
                char m[160];
                snprintf(m, sizeof(m), "point %lld is an element point of two tubes",
                         (long long)a->elem_pts[q]);
                c.fail(m);
                break;
            }
            seen[(size_t)a->elem_pts[q]] = 1;
        }
    }
    c.juncs("junctions", a->junctions, a->n_junctions, a->n_points, a->n_points, a->n_tubes,
            a->n_tubes);
    const int64_t n_gs = c.offsets("gstruct_ptr", a->gstruct_ptr, a->n_gnps);
    c.range("gstruct", a->gstruct, n_gs, a->n_gnp_points);
    if (c.ok) {
        std::vector<int64_t> plate((size_t)a->n_gnp_points, -1);
        for (int64_t g = 0; c.ok && g < a->n_gnps; g++)
            for (int64_t q = a->gstruct_ptr[g]; q < a->gstruct_ptr[g + 1]; q++) {
                int64_t &o = plate[(size_t)a->gstruct[q]];
                if (o >= 0) {
                    char m[160];
                    snprintf(m, sizeof(m), "plate point %lld appears twice in the plates' "
                             "structures", (long long)a->gstruct[q]);
                    c.fail(m);
                    break;
                }
                o = g;
            }
    }
    const int64_t n_gt = c.offsets("gtri_ptr", a->gtri_ptr, a->n_gnps);
    for (int64_t g = 0; c.ok && g < a->n_gnps; g++)
        if ((a->gtri_ptr[g + 1] - a->gtri_ptr[g]) % 2) c.fail("gtri_ptr: odd edge list");
    c.range("gtri", a->gtri, n_gt, a->n_gnp_points);
    c.juncs("mixed", a->mixed, a->n_mixed, a->n_points, a->n_gnp_points, a->n_tubes, a->n_gnps);
    c.juncs("gjunctions", a->gjunctions, a->n_gjunctions, a->n_gnp_points, a->n_gnp_points,
            a->n_gnps, a->n_gnps);
    const int64_t *fptr[4] = {a->bpts_ptr, a->bcnt_ptr, a->gbpts_ptr, a->gbnd_ptr};
    const int64_t *fent[4] = {a->bpts, a->bcnt, a->gbpts, a->gbnd};
    const char *fname[4] = {"bpts", "bcnt", "gbpts", "gbnd"};
    const int64_t fhi[4] = {a->n_points, a->n_tubes, a->n_gnp_points, a->n_gnps};
    int64_t fcount[4] = {0, 0, 0, 0};
    for (int f = 0; f < 4; f++) {
        fcount[f] = c.offsets(fname[f], fptr[f], 6);
        c.range(fname[f], fent[f], fcount[f], fhi[f]);
    }
    const int64_t *lptr[5] = {a->cl_tubes_ptr, a->cl_junc_ptr, a->cl_gnps_ptr, a->cl_mixed_ptr,
                              a->cl_gjunc_ptr};
    const int64_t *lent[5] = {a->cl_tubes, a->cl_junc, a->cl_gnps, a->cl_mixed, a->cl_gjunc};
    const char *lname[5] = {"cl_tubes", "cl_junc", "cl_gnps", "cl_mixed", "cl_gjunc"};
    const int64_t lhi[5] = {a->n_tubes, a->n_junctions, a->n_gnps, a->n_mixed, a->n_gjunctions};
    int64_t lcount[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < 5; l++) {
        lcount[l] = c.offsets(lname[l], lptr[l], a->n_clusters);
        c.range(lname[l], lent[l], lcount[l], lhi[l]);
    }
    // a cluster's plates are searched by bisection
    for (int64_t k = 0; c.ok && k < a->n_clusters; k++)
        for (int64_t q = a->cl_gnps_ptr[k] + 1; q < a->cl_gnps_ptr[k + 1]; q++)
            if (a->cl_gnps[q] <= a->cl_gnps[q - 1]) {
                c.fail("cl_gnps: a cluster's plates are not ascending");
                break;
            }
    if (a->n_mixed >= INT_MAX) c.fail("too many tube-plate junctions");
    if (!c.ok) return 0;

    auto *dn = new cicy_dnet();
    dn->ctx = ctx;
    dn->n_points = a->n_points;
    dn->n_tubes = a->n_tubes;
    dn->n_junc = a->n_junctions;
    dn->n_gpts = a->n_gnp_points;
    dn->n_gnps = a->n_gnps;
    dn->n_mixed = a->n_mixed;
    dn->n_gjunc = a->n_gjunctions;
    dn->n_clusters = a->n_clusters;
    for (int k = 0; k < 3; k++) {
        dn->lo[k] = a->window_lo[k];
        dn->hi[k] = a->window_hi[k];
    }
    bool ok = upload(dn, a->xyz, 3 * a->n_points, &dn->d.xyz) &&
              upload(dn, a->radius, a->n_tubes, &dn->d.radius) &&
              upload(dn, a->gxyz, 3 * a->n_gnp_points, &dn->d.gxyz) &&
              upload(dn, a->gnp_thickness, a->n_gnps, &dn->d.thick) &&
              upload(dn, a->elem_ptr, a->n_tubes + 1, &dn->d.elem_ptr) &&
              upload(dn, a->elem_pts, n_elem, &dn->d.elem_pts) &&
              upload(dn, a->gstruct_ptr, a->n_gnps + 1, &dn->d.gstruct_ptr) &&
              upload(dn, a->gstruct, n_gs, &dn->d.gstruct) &&
              upload(dn, a->gtri_ptr, a->n_gnps + 1, &dn->d.gtri_ptr) &&
              upload(dn, a->gtri, n_gt, &dn->d.gtri) &&
              upload(dn, a->junctions, a->n_junctions, &dn->d.junc) &&
              upload(dn, a->mixed, a->n_mixed, &dn->d.mixed) &&
              upload(dn, a->gjunctions, a->n_gjunctions, &dn->d.gjunc);
    for (int f = 0; ok && f < 4; f++) {
        dn->face_ptr[f].assign(fptr[f], fptr[f] + 7);
        ok = upload(dn, fent[f], fcount[f], &dn->faces[f]);
    }
    for (int l = 0; ok && l < 5; l++) {
        dn->list_ptr[l].assign(lptr[l], lptr[l] + a->n_clusters + 1);
        ok = upload(dn, lent[l], lcount[l], &dn->lists[l]);
    }
    if (ok) ok = hip_ok(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (!ok) {
        cicy_dnet_destroy(dn);
        return 0;
    }
    *out = dn;
    return 1;
}

void cicy_dnet_destroy(cicy_dnet *dn) {
    if (!dn) return;
    (void)hipStreamSynchronize(dn->ctx->stream);
    for (void *p : dn->owned) dev_release(dn->ctx, p);
    delete dn;
}

int cicy_dnet_assemble(cicy_dnet *dn, int64_t k, int r_flag, int family,
                       const cicy_elec_params *ep_in, int direction, cicy_asm **out) {
    if (!dn || !ep_in || !out) return 0;
    *out = nullptr;
    cicy_ctx *ctx = dn->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) {
        set_error(ctx, "hipSetDevice failed");
        return 0;
    }
    if (k < 0 || k >= dn->n_clusters) {
        set_error(ctx, "invalid network or cluster index");
        return 0;
    }
    if (r_flag != 0 && r_flag != 1) {
        set_error(ctx, "invalid R_flag (valid: 0, 1)");
        return 0;
    }
    const std::vector<int> faces = family_faces(family);
    if (faces.empty()) {
        set_error(ctx, "invalid family");
        return 0;
    }
    const cicy_elec_params ep = *ep_in;
    const int64_t rd = (int64_t)faces.size();
    hipStream_t st = ctx->stream;
    const NetDev &d = dn->d;
    auto list = [&](int l, long long *count) {
        *count = dn->list_ptr[l][k + 1] - dn->list_ptr[l][k];
        return dn->lists[l] + dn->list_ptr[l][k];
    };
    long long T, J, M, G, GJ;
    const int64_t *ct = list(0, &T), *cj = list(1, &J), *cg = list(2, &G), *cm = list(3, &M),
                  *cgj = list(4, &GJ);
    const bool timing = timing_on();
    double t0 = timing ? now_ms() : 0.0, tl = t0;
    auto mark = [&](const char *what) {
        if (!timing) return;
        (void)hipStreamSynchronize(st);
        const double t = now_ms();
        fprintf(stderr, "[assembly] %-22s %8.3f ms\n", what, t - tl);
        tl = t;
    };

    auto *as = new cicy_asm();
    as->ctx = ctx;
    cicy_asm_view &v = as->v;
    PoolScope ps(ctx);
    AsmRun run{ctx, as, ps};
    auto fail = [&]() {
        cicy_asm_free(as);
        return 0;
    };

    int64_t *lmm = nullptr, *lmm_g = nullptr, *pickbuf = nullptr;
    AsmErr *err = nullptr;
    if (!run.out(lmm, dn->n_points) || !run.out(lmm_g, dn->n_gpts) || !ps.get(pickbuf, 8) ||
        !ps.get(err, 1))
        return fail();
    hipLaunchKernelGGL(k_err_init, dim3(1), dim3(64), 0, st, err);
    hipLaunchKernelGGL(k_fill64, dim3(grid_for(dn->n_points)), dim3(kAsmThreads), 0, st, lmm,
                       (long long)dn->n_points, (int64_t)-1);
    hipLaunchKernelGGL(k_fill64, dim3(grid_for(dn->n_gpts)), dim3(kAsmThreads), 0, st, lmm_g,
                       (long long)dn->n_gpts, (int64_t)-1);
    // Map_points_at_boundaries: later faces overwrite earlier ones, so one launch per face
    for (size_t f = 0; f < faces.size(); f++) {
        const int b = faces[f];
        const long long nb = dn->face_ptr[0][b + 1] - dn->face_ptr[0][b];
        if (nb > 0)
            hipLaunchKernelGGL(k_mark, dim3(grid_for(nb)), dim3(kAsmThreads), 0, st, lmm,
                               dn->faces[0] + dn->face_ptr[0][b], nb, (int64_t)f);
        const long long ng = dn->face_ptr[2][b + 1] - dn->face_ptr[2][b];
        if (ng > 0)
            hipLaunchKernelGGL(k_mark, dim3(grid_for(ng)), dim3(kAsmThreads), 0, st, lmm_g,
                               dn->faces[2] + dn->face_ptr[2][b], ng, (int64_t)f);
    }
    // node and resistor counts, scans
    int64_t *tn = nullptr, *te = nullptr, *tnb = nullptr, *teb = nullptr, *gn = nullptr,
            *gt = nullptr, *gnb = nullptr, *gtb = nullptr;
    unsigned char *ends = nullptr;
    if (!ps.get(ends, T) || !ps.get(tn, T + 1) || !ps.get(te, T + 1) || !ps.get(tnb, T + 1) ||
        !ps.get(teb, T + 1) ||
        !ps.get(gn, G + 1) || !ps.get(gt, G + 1) || !ps.get(gnb, G + 1) || !ps.get(gtb, G + 1))
        return fail();
    hipLaunchKernelGGL(k_tube_count, dim3(grid_for(T + 1)), dim3(kAsmThreads), 0, st, d, ct, T, lmm,
                       tn, te, ends, err);
    hipLaunchKernelGGL(k_gnp_count, dim3((unsigned)(G + 1)), dim3(kAsmThreads), 0, st, d, cg, G,
                       lmm_g, gn, gt);
    if (!run.launched("k_tube_count") || !run.scan(tn, tnb, T + 1) || !run.scan(te, teb, T + 1) ||
        !run.scan(gn, gnb, G + 1) || !run.scan(gt, gtb, G + 1))
        return fail();
    // the error words as int64 entries for k_pick
    auto word = [](long long *p) { return reinterpret_cast<const int64_t *>(p); };
    int64_t h[8];
    if (!run.pick(pickbuf,
                  {{tnb, T}, {teb, T}, {gnb, G}, {gtb, G}, {word(&err->bad_tube), 0}}, h))
        return fail();
    if (h[4] != kNone) {
        set_error(ctx, "LM_matrix_for_cnts: a tube with not enough points to form an element");
        return fail();
    }
    const int64_t first_gnp = rd + h[0];
    const int64_t nodes = first_gnp + h[2], n = nodes - rd;
    const long long Et = h[1], Eg = h[3];
    const long long S = Et + J + M + Eg + GJ;
    int cb = 1;
    while (cb < 40 && (nodes >> cb) != 0) cb++;
    if (2 * cb > 62 || nodes >= INT_MAX || S >= (1ll << 30) || n <= 0) {
        set_error(ctx, "cicy_dnet_assemble: system too large for the device assembly "
                       "(nodes %lld, resistors %lld) or empty", (long long)nodes, S);
        return fail();
    }
    mark("numbering counts");
    hipLaunchKernelGGL(k_tube_number, dim3(grid_tubes(T)), dim3(kAsmThreads), 0, st, d, ct, T,
                       lmm, tnb, ends, rd);
    if (G > 0)
        hipLaunchKernelGGL(k_gnp_number, dim3((unsigned)G), dim3(kAsmThreads), 0, st, d, cg, G,
                           lmm_g, gnb, first_gnp);
    int *owner = nullptr;
    if (!ps.get(owner, dn->n_gpts)) return fail();
    hipLaunchKernelGGL(k_fill32, dim3(grid_for(dn->n_gpts)), dim3(kAsmThreads), 0, st, owner,
                       (long long)dn->n_gpts, -1);
    if (M > 0)
        hipLaunchKernelGGL(k_mixed_owner, dim3(grid_for(M)), dim3(kAsmThreads), 0, st, d, cm, M, lmm,
                           lmm_g, owner);
    if (!run.launched("numbering")) return fail();
    mark("numbering");

    // resistor records in the reference's insertion order
    Recs r{};
    r.cb = cb;
    r.nodes = nodes;
    int *iota = nullptr, *sslot = nullptr, *sdnode = nullptr;
    unsigned long long *skey = nullptr;
    double *sdval = nullptr;
    if (!ps.get(r.key, S) || !ps.get(r.val, S) || !ps.get(r.rule, S) || !ps.get(r.dnode, 2 * S) ||
        !ps.get(r.dval, 2 * S) || !ps.get(iota, S) || !ps.get(skey, S) || !ps.get(sslot, S) ||
        !ps.get(sdnode, 2 * S) || !ps.get(sdval, 2 * S))
        return fail();
    long long s0 = 0;
    if (T > 0)
        hipLaunchKernelGGL(k_edges_tubes, dim3(grid_tubes(T)), dim3(kAsmThreads), 0, st, d, ct, T, lmm,
                           teb, r_flag, ep, r);
    s0 += Et;
    if (J > 0)
        hipLaunchKernelGGL(k_edges_junc, dim3(grid_for(J)), dim3(kAsmThreads), 0, st, d, cj, J, lmm,
                           s0, rd, r_flag, ep, r);
    s0 += J;
    if (M > 0)
        hipLaunchKernelGGL(k_edges_mixed, dim3(grid_for(M)), dim3(kAsmThreads), 0, st, d, cm, M, lmm,
                           lmm_g, s0, rd, r_flag, ep, r);
    s0 += M;
    if (G > 0)
        hipLaunchKernelGGL(k_edges_gnp, dim3((unsigned)G), dim3(kAsmThreads), 0, st, d, cg, G,
                           lmm_g, owner, cm, gtb, s0, r_flag, ep, r, err);
    s0 += Eg;
    if (GJ > 0)
        hipLaunchKernelGGL(k_edges_gjunc, dim3(grid_for(GJ)), dim3(kAsmThreads), 0, st, d, cgj, GJ,
                           lmm_g, s0, r_flag, ep, r);
    hipLaunchKernelGGL(k_iota, dim3(grid_for(S)), dim3(kAsmThreads), 0, st, iota, S);
    if (!run.launched("resistor records")) return fail();
    mark("resistor records");
    if (!run.sort(r.key, skey, iota, sslot, S, 2 * cb) ||
        !run.sort(r.dnode, sdnode, r.dval, sdval, 2 * S, cb))
        return fail();
    mark("sorts");

    // rows
    Vef vef{};
    const double vmag = (r_flag == 1) ? ep.applied_voltage : (double)nodes;
    for (int i = 1; i < rd; i++) vef.v[i] = vef.v[i - 1] + vmag;
    const bool has_guess = (r_flag == 1 && rd == 2);
    int *row_start = nullptr;
    int64_t *cnt = nullptr, *row_ptr = nullptr;
    if (!ps.get(row_start, nodes) || !ps.get(cnt, n + 1) || !run.out(row_ptr, n + 1))
        return fail();
    hipLaunchKernelGGL(k_fill32, dim3(grid_for(nodes)), dim3(kAsmThreads), 0, st, row_start,
                       (long long)nodes, -1);
    hipLaunchKernelGGL(k_row_heads, dim3(grid_for(S)), dim3(kAsmThreads), 0, st, skey, S, cb, nodes,
                       row_start);
    hipLaunchKernelGGL(k_fold<false>, dim3(grid_for(n + 1)), dim3(kAsmThreads), 0, st, skey, sslot,
                       r.val, r.rule, S, row_start, cb, rd, nodes, vef, cnt, nullptr, nullptr,
                       nullptr, nullptr, nullptr, nullptr);
    if (!run.launched("k_fold") || !run.scan(cnt, row_ptr, n + 1) ||
        !run.pick(pickbuf, {{row_ptr, n}, {word(&err->gnp), 0}}, h))
        return fail();
    if (h[1] != kNone) {
        set_error(ctx, (h[1] & 1) ? "Calculate_resistance_gnp: triangulation edge of zero length"
                                  : "Fill_2d_matrices_gnp: triangulation vertex without a node");
        return fail();
    }
    const long long nnz = h[0];
    int64_t *col_ind = nullptr;
    double *values = nullptr, *b = nullptr, *diag = nullptr, *x0 = nullptr, *r0 = nullptr;
    int *erow = nullptr, *col32 = nullptr, *dstart = nullptr;
    if (!run.out(col_ind, nnz) || !run.out(values, nnz) || !run.out(b, n) || !run.out(diag, n) ||
        !run.out(x0, n) || !run.out(r0, n) || !ps.get(dstart, nodes) ||
        (has_guess && (!ps.get(erow, nnz) || !ps.get(col32, nnz))))
        return fail();
    hipLaunchKernelGGL(k_fold<true>, dim3(grid_for(n + 1)), dim3(kAsmThreads), 0, st, skey, sslot,
                       r.val, r.rule, S, row_start, cb, rd, nodes, vef, nullptr, row_ptr, col_ind,
                       values, erow, col32, b);
    hipLaunchKernelGGL(k_fill32, dim3(grid_for(nodes)), dim3(kAsmThreads), 0, st, dstart,
                       (long long)nodes, -1);
    hipLaunchKernelGGL(k_heads, dim3(grid_for(2 * S)), dim3(kAsmThreads), 0, st, sdnode, 2 * S,
                       (long long)nodes, dstart);
    hipLaunchKernelGGL(k_diag, dim3(grid_for(n)), dim3(kAsmThreads), 0, st, sdnode, sdval, dstart,
                       2 * S, rd, nodes, diag, err);
    if (!run.launched("rows")) return fail();
    mark("rows and diagonal");

    // ramp guess and its residual
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(kAsmThreads), 0, st, x0, (long long)n,
                       0.0);
    if (has_guess) {
        const int axis = family == 0 ? 0 : family == 1 ? 1 : 2;
        const double l_win = dn->hi[axis] - dn->lo[axis], max_coord = dn->hi[axis];
        const double vpc = ep.applied_voltage / l_win;
        if (T > 0)
            hipLaunchKernelGGL(k_ramp, dim3(grid_tubes(T)), dim3(kAsmThreads), 0, st, d, ct, T, lmm, rd,
                               axis, vpc, max_coord, x0);
        int *eiota = nullptr, *sent = nullptr, *scol = nullptr, *cstart = nullptr;
        if (!ps.get(eiota, nnz) || !ps.get(sent, nnz) || !ps.get(scol, nnz) ||
            !ps.get(cstart, n))
            return fail();
        hipLaunchKernelGGL(k_iota, dim3(grid_for(nnz)), dim3(kAsmThreads), 0, st, eiota, nnz);
        hipLaunchKernelGGL(k_fill32, dim3(grid_for(n)), dim3(kAsmThreads), 0, st, cstart,
                           (long long)n, -1);
        if (!run.launched("k_ramp") || !run.sort(col32, scol, eiota, sent, nnz, cb))
            return fail();
        hipLaunchKernelGGL(k_heads, dim3(grid_for(nnz)), dim3(kAsmThreads), 0, st, scol, nnz,
                           (long long)n, cstart);
        hipLaunchKernelGGL(k_r0, dim3(grid_for(n)), dim3(kAsmThreads), 0, st, row_ptr, col_ind,
                           values, diag, x0, b, scol, sent, cstart, erow, nnz, n, r0);
    } else if (!hip_ok(ctx, hipMemcpyAsync(r0, b, (size_t)n * sizeof(double),
                                           hipMemcpyDeviceToDevice, st),
                       "D2D")) {
        return fail();
    }
    if (!run.launched("k_r0")) return fail();
    mark("ramp and residual");

    // boundary element lists
    long long bnd_tot[2] = {0, 0};
    int64_t *bcnt[2] = {nullptr, nullptr}, *bpos[2] = {nullptr, nullptr};
    BndDev bd[2] = {};
    v.direction = -1;
    if (direction >= 0 && direction <= 2) {
        const int b1s[3] = {2, 3, 0}, b2s[3] = {4, 5, 1};
        const int bb[2] = {b1s[direction], b2s[direction]};
        for (int w = 0; w < 2; w++) {
            BndDev &q = bd[w];
            q.tubes = dn->faces[1] + dn->face_ptr[1][bb[w]];
            q.nt = dn->face_ptr[1][bb[w] + 1] - dn->face_ptr[1][bb[w]];
            q.gnps = dn->faces[3] + dn->face_ptr[3][bb[w]];
            q.ng = dn->face_ptr[3][bb[w] + 1] - dn->face_ptr[3][bb[w]];
            q.cg = cg;
            q.G = G;
            q.w = w;
            bnd_tot[w] = q.nt + q.ng;
            if (!ps.get(bcnt[w], bnd_tot[w] + 1) || !ps.get(bpos[w], bnd_tot[w] + 1)) return fail();
            hipLaunchKernelGGL(k_bnd_tubes<false>, dim3(grid_for(q.nt + 1)), dim3(kAsmThreads),
                               0, st, d, q, lmm, ep, bcnt[w], nullptr, nullptr, nullptr,
                               nullptr);
            if (q.ng > 0)
                hipLaunchKernelGGL(k_bnd_plates<false>, dim3((unsigned)q.ng), dim3(kAsmThreads),
                                   0, st, d, q, lmm_g, owner, cm, ep, bcnt[w], nullptr, nullptr,
                                   nullptr, nullptr, err);
            if (!run.launched("k_bnd") || !run.scan(bcnt[w], bpos[w], bnd_tot[w] + 1))
                return fail();
        }
    }
    if (!run.pick(pickbuf,
                  {{word(&err->floating), 0}, {word(&err->bnd), 0}, {bpos[0] ? bpos[0] : row_ptr,
                                                        bpos[0] ? bnd_tot[0] : 0},
                   {bpos[1] ? bpos[1] : row_ptr, bpos[1] ? bnd_tot[1] : 0}},
                  h))
        return fail();
    if (h[0] != kNone) {
        set_error(ctx, "Fill_sparse_stiffness_matrix: zero element on the diagonal (floating node)");
        return fail();
    }
    if (h[1] != kNone) {
        set_error(ctx, (h[1] & 1) ? "Calculate_resistance_gnp: boundary edge of zero length"
                                  : "Currents_through_boundary_gnps: both nodes at the same boundary");
        return fail();
    }
    if (direction >= 0 && direction <= 2) {
        for (int w = 0; w < 2; w++) {
            v.ne[w] = h[2 + w];
            if (!run.out(v.d_e_bnd[w], v.ne[w]) || !run.out(v.d_e_int[w], v.ne[w]) ||
                !run.out(v.d_e_re[w], v.ne[w]))
                return fail();
            hipLaunchKernelGGL(k_bnd_tubes<true>, dim3(grid_for(bd[w].nt + 1)),
                               dim3(kAsmThreads), 0, st, d, bd[w], lmm, ep, nullptr, bpos[w],
                               v.d_e_bnd[w], v.d_e_int[w], v.d_e_re[w]);
            if (bd[w].ng > 0)
                hipLaunchKernelGGL(k_bnd_plates<true>, dim3((unsigned)bd[w].ng),
                                   dim3(kAsmThreads), 0, st, d, bd[w], lmm_g, owner, cm, ep,
                                   nullptr, bpos[w], v.d_e_bnd[w], v.d_e_int[w], v.d_e_re[w],
                                   err);
            if (!run.launched("k_bnd")) return fail();
        }
        v.direction = direction;
    }
    // each plate's free rows (its new nodes are consecutive)
    std::vector<int64_t> hgn((size_t)G + 1), hgnb((size_t)G + 1);
    if (G > 0 && (!d2h_sync(ctx, hgn.data(), gn, (size_t)(G + 1) * sizeof(int64_t)) ||
                  !d2h_sync(ctx, hgnb.data(), gnb, (size_t)(G + 1) * sizeof(int64_t))))
        return fail();
    for (long long i = 0; i < G; i++)
        if (hgn[(size_t)i] > 0) {
            as->blocks.push_back(first_gnp - rd + hgnb[(size_t)i]);
            as->blocks.push_back(first_gnp - rd + hgnb[(size_t)i] + hgn[(size_t)i]);
        }
    if (!hip_ok(ctx, hipStreamSynchronize(st), "sync")) return fail();
    mark("boundary lists");
    if (timing) fprintf(stderr, "[assembly] total %.3f ms\n", now_ms() - t0);

    v.nodes = nodes;
    v.reserved = rd;
    v.n = n;
    v.nnz = nnz;
    v.has_guess = has_guess ? 1 : 0;
    for (int i = 0; i < 8; i++) v.vef[i] = i < rd ? vef.v[i] : 0.0;
    v.d_row_ptr = row_ptr;
    v.d_col_ind = col_ind;
    v.d_values = values;
    v.d_diag = diag;
    v.d_b = b;
    v.d_x0 = x0;
    v.d_r0 = r0;
    v.n_points = dn->n_points;
    v.n_gnp_points = dn->n_gpts;
    v.d_lmm = lmm;
    v.d_lmm_gnp = lmm_g;
    v.gnp_nodes = nodes - first_gnp;
    v.n_blocks = (int64_t)as->blocks.size() / 2;
    v.blocks = as->blocks.data();
    *out = as;
    return 1;
}

int cicy_asm_get(const cicy_asm *a, cicy_asm_view *out) {
    if (!a || !out) return 0;
    *out = a->v;
    return 1;
}

void cicy_asm_free(cicy_asm *a) {
    if (!a) return;
    (void)hipStreamSynchronize(a->ctx->stream);
    for (void *p : a->owned) dev_release(a->ctx, p);
    delete a;
}

}  // extern "C"
