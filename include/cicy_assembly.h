/*
 * cicy_assembly.h -- C-ABI of the device assembly of a cluster's DEA system (libcicy_dea.so,
 * gfx950).
 *
 * Replaces what Direct_Electrifying::Compute_voltage_field does before its solve
 * (Direct_Electrifying.cpp:11-113), on the device, from a network held in HBM:
 *   LM_matrix_for_cnts            Direct_Electrifying.h:40, Direct_Electrifying.cpp:139-215
 *   Map_points_at_boundaries      Direct_Electrifying.h:41, Direct_Electrifying.cpp:217-247
 *   LM_matrix_for_gnps            Direct_Electrifying.h:43, Direct_Electrifying.cpp:307-349
 *   Fill_sparse_stiffness_matrix  Direct_Electrifying.h:44, Direct_Electrifying.cpp:351-487
 *     Fill_2d_matrices_cnts / _cnt_junctions / _mixed_junctions / _gnp / _gnp_junctions
 *                                 Direct_Electrifying.cpp:489-1327
 *     From_2d_to_1d_vectors       Direct_Electrifying.h:57, Direct_Electrifying.cpp:1329-1394
 *   Initial_guess_for_CG          Direct_Electrifying.h:58, Direct_Electrifying.cpp:1396-1470
 *   Update_residual_vector        Direct_Electrifying.h:62, Direct_Electrifying.cpp:1502-1535
 * and the boundary element lists of Electrical_analysis::Calculate_parallel_resistor
 * (Electrical_analysis.cpp:499-675).
 *
 * The reference fills one std::map per row and walks the maps in key order.  Here every
 * resistor becomes a (row, col, value, rule) record at its position in the reference's
 * insertion order; a stable radix sort on (row, col) and one fold per (row, col) run with the
 * reference's rule (an add_new overwrites, an add_to_existing subtracts) give the maps' walk;
 * each node's diagonal is summed in insertion order from a stable sort of its contributions.
 * Node numbers, row_ptr, col_ind and the boundary lists equal the host assembler's
 * (cicy_net_assemble, libcicy_net.so) exactly; so do values, diagonal, b, x0 and r0 except
 * where a tunnelling junction's exp() is involved, which the device evaluates with its own
 * library to within an ulp or so of the host's.
 *
 * Conventions as in cicy_dea.h: 1 on success, 0 on failure with the message in
 * cicy_last_error(ctx); the reference's error conditions give its messages.
 */
#ifndef CICY_ASSEMBLY_H
#define CICY_ASSEMBLY_H

#include <stdint.h>

#include "cicy_dea.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef CICY_ELEC_PARAMS_DEFINED
#define CICY_ELEC_PARAMS_DEFINED
/* Electric_para's fields that the element laws use (Direct_Electrifying.cpp:611-646,
 * :854-901, :1198-1232). */
typedef struct cicy_elec_params {
    double applied_voltage;    /* V_app */
    double resistivity_cnt;    /* Ohm.um */
    int junction_exponential;  /* 1: C1*d*exp(C2*d)/A; 0: constant junction_resistance */
    double junction_resistance;
    double c1, c2;             /* tunnelling constants */
    double resistivity_gnp_t;     /* Ohm.um, across the plate */
    double resistivity_gnp_surf;  /* Ohm.um, in the plate */
} cicy_elec_params;
#endif

/* A junction record: two points and the two particles they lie on, and the gap (um).  Tube
 * junctions (the reference's junctions_cnt): points_cnt ids, tubes e1 < e2.  Tube-plate
 * junctions (junctions_mixed): tube point, plate point, tube, plate.  Plate-plate junctions
 * (junctions_gnp): plate points, plates e1 < e2. */
typedef struct cicy_junction_rec {
    int64_t p1, p2;
    int64_t e1, e2;
    double dist;
} cicy_junction_rec;

/* A network in flat arrays (host memory), what Hoshen_Kopelman, Cutoff_Wins and the
 * generator hand to Compute_voltage_field.  A list of lists is an offset array [count + 1]
 * (starting at 0) into an entry array.  Face lists are numbered as Cutoff_Wins does (x: 4
 * and 2, y: 5 and 3, z: 1 and 0) and have 6 + 1 offsets. */
typedef struct cicy_net_arrays {
    int64_t n_points;           /* points_cnt */
    const double *xyz;          /* [3 n_points] */
    int64_t n_tubes;
    const double *radius;       /* radii [n_tubes] */
    const int64_t *elem_ptr;    /* elements_cnt: ascending point ids of each tube */
    const int64_t *elem_pts;
    int64_t n_junctions;
    const cicy_junction_rec *junctions;
    int64_t n_gnp_points;       /* points_gnp */
    const double *gxyz;         /* [3 n_gnp_points] */
    int64_t n_gnps;
    const double *gnp_thickness;      /* [n_gnps] */
    const int64_t *gstruct_ptr;       /* structure_gnp */
    const int64_t *gstruct;
    const int64_t *gtri_ptr;          /* triangulation edges of each plate, 2 ids per edge */
    const int64_t *gtri;
    int64_t n_mixed;
    const cicy_junction_rec *mixed;
    int64_t n_gjunctions;
    const cicy_junction_rec *gjunctions;
    const int64_t *bpts_ptr, *bpts;    /* boundary_cnt_pts */
    const int64_t *bcnt_ptr, *bcnt;    /* boundary_cnt */
    const int64_t *gbpts_ptr, *gbpts;  /* boundary_gnp_pts */
    const int64_t *gbnd_ptr, *gbnd;    /* boundary_gnp */
    int64_t n_clusters;
    const int64_t *cl_tubes_ptr, *cl_tubes;  /* clusters_cnt */
    const int64_t *cl_junc_ptr, *cl_junc;    /* each cluster's tube junctions */
    const int64_t *cl_gnps_ptr, *cl_gnps;    /* clusters_gnp, ascending */
    const int64_t *cl_mixed_ptr, *cl_mixed;  /* each cluster's tube-plate junctions */
    const int64_t *cl_gjunc_ptr, *cl_gjunc;  /* each cluster's plate-plate junctions */
    double window_lo[3], window_hi[3];       /* the observation window (ramp guess) */
} cicy_net_arrays;

typedef struct cicy_dnet cicy_dnet;
typedef struct cicy_asm cicy_asm;

/* Copy a network to the device once; every assembly of its clusters then reads it there.
 * Every index is checked against the array it points into (an out-of-range index is refused
 * with its array's name).  The host arrays are not kept. */
int cicy_dnet_create(cicy_ctx *ctx, const cicy_net_arrays *net, cicy_dnet **out);
void cicy_dnet_destroy(cicy_dnet *dn);

/* Assemble cluster k with resistor flag r_flag (0 unit resistors, 1 element laws) for
 * percolation family `family` (0..6; Get_vector_of_boundaries, Direct_Electrifying.cpp:
 * 249-305), and, for direction 0..2, the two boundaries' element lists (-1: none).  Every
 * array is built on the device; the call waits for the few counts it reads back. */
int cicy_dnet_assemble(cicy_dnet *dn, int64_t k, int r_flag, int family,
                       const cicy_elec_params *ep, int direction, cicy_asm **out);

/* What an assembly holds.  d_ arrays are device memory owned by the assembly (valid until
 * cicy_asm_free); the others are host memory owned by it. */
typedef struct cicy_asm_view {
    int64_t nodes;          /* global nodes (reserved + free) */
    int64_t reserved;       /* R_d */
    int64_t n;              /* free nodes */
    int64_t nnz;            /* stored strictly-lower entries */
    int has_guess;          /* R_flag == 1 && R_d == 2: x0 is the ramp, r0 = b - A x0 */
    int direction;          /* -1 when no boundary lists were built */
    double vef[8];          /* prescribed voltages of the reserved nodes (R_d <= 6) */
    int64_t *d_row_ptr;     /* [n + 1] */
    int64_t *d_col_ind;     /* [nnz] */
    double *d_values;       /* [nnz] */
    double *d_diag;         /* [n] */
    double *d_b;            /* [n] */
    double *d_x0;           /* [n] (zeros without a guess) */
    double *d_r0;           /* [n] */
    int64_t n_points, n_gnp_points;
    int64_t *d_lmm;         /* [n_points]: point -> global node or -1 (LMM_cnts) */
    int64_t *d_lmm_gnp;     /* [n_gnp_points]: plate point -> global node or -1 (LMM_gnps) */
    int64_t gnp_nodes;      /* nodes numbered by plates */
    int64_t n_blocks;       /* plates of the cluster with free nodes */
    const int64_t *blocks;  /* [2 n_blocks]: each plate's free-node rows [start, end), the
                               groups cicy_sys_set_blocks takes */
    int64_t ne[2];          /* boundary element counts, boundaries b1 and b2 */
    int64_t *d_e_bnd[2];    /* boundary-side global node of each element */
    int64_t *d_e_int[2];    /* the other end's global node */
    double *d_e_re[2];      /* element resistance (Ohm) */
} cicy_asm_view;

int cicy_asm_get(const cicy_asm *a, cicy_asm_view *out);
void cicy_asm_free(cicy_asm *a);

#ifdef __cplusplus
}
#endif

#endif
