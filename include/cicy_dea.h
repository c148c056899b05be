/*
 * cicy_dea.h -- C-ABI of the MI355X (gfx950) Direct Electrifying conductance solver.
 *
 * Replaces the CPU solve of rnaimehaom/CICY_geom's Direct_Electrifying class:
 *   Solve_DEA_equations_CG_SSS   Direct_Electrifying.h:63, Direct_Electrifying.cpp:1538-1659
 *   spM_V_SSS                    Direct_Electrifying.h:67, Direct_Electrifying.cpp:1678-1695
 *   Update_residual_vector       Direct_Electrifying.h:62, Direct_Electrifying.cpp:1502-1535
 * and the boundary current sum of Electrical_analysis::Calculate_parallel_resistor
 * (Electrical_analysis.cpp:355-473, :547-605).
 *
 * Conventions (mirroring the reference):
 *   - every entry point returns 1 on success and 0 on failure, as the reference's int
 *     functions do; cicy_last_error() holds the message the reference would write to
 *     hout, and the stats struct carries a status code so that the max-iteration case
 *     (a reference "success") and a breakdown are distinguishable;
 *   - matrices are the reference's SSS arrays: row_ptr[n+1] (row_ptr[0] == 0), col_ind
 *     and values of the strictly-lower triangle (col < row), diagonal[n]; indices are
 *     int64 (`long` on LP64), values double;
 *   - the caller owns every host array; the library borrows them for the duration of
 *     the call.  Device copies live in the context or a cicy_sys handle;
 *   - a context is not re-entrant: calls on one context are serialised by the caller.
 *     Work inside a call is ordered on the context's stream and complete at return.
 */
#ifndef CICY_DEA_H
#define CICY_DEA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cicy_ctx cicy_ctx;
typedef struct cicy_sys cicy_sys;

/* status codes (cicy_solve_stats.status) */
#define CICY_OK 0            /* converged */
#define CICY_MAXITER 1       /* iteration cap reached; the reference returns success here */
#define CICY_BREAKDOWN 2     /* p.Ap not finite or not positive (singular/indefinite) */
#define CICY_BAD_INPUT 3     /* malformed arrays (checked on the host before any launch) */
#define CICY_DEVICE_ERROR 4  /* a HIP call failed */

/* solve methods (cicy_sys_solve) */
#define CICY_METHOD_PCG_F64 0    /* plain Jacobi-PCG in FP64, the reference's algorithm */
#define CICY_METHOD_MIXED 1      /* FP32 Jacobi-PCG with FP64 reliable updates/refinement */
#define CICY_METHOD_RESIDENT 2   /* the reference's FP64 Jacobi-PCG run as register-resident
                                    cooperative launches (band + remainder systems that fit
                                    the device's registers; cicy_sys_size 6 > 0) */
#define CICY_METHOD_CONDENSED 3  /* exact static condensation: chain-interior rows (two band
                                    neighbours, no junction entries) are eliminated by
                                    tridiagonal sweeps, then rows left with at most two
                                    neighbours by rounds of series elimination; the
                                    reference's FP64 Jacobi-PCG runs on the remaining core
                                    (the Schur complement), the eliminated rows are recovered
                                    by back-substitution (cicy_sys_size 9 > 0) */
#define CICY_METHOD_CONDENSED_RESIDENT 4  /* the same with the resident CG on the reduced
                                             system */
#define CICY_METHOD_CONDENSED_BLOCK 5  /* the same eliminations; the core is solved by one
                                          persistent cooperative launch per chunk of
                                          iterations (Chronopoulos-Gear PCG, two grid
                                          barriers per iteration) preconditioned by exact
                                          dense inverses of the row groups given by
                                          cicy_sys_set_blocks (e.g. the nodes of each
                                          nanoplatelet) and Jacobi on the other rows */

/* stopping rule (cicy_sys_solve rel_to) */
#define CICY_REL_R0 0  /* ||r_k|| <= tol*||r_0||, r_0 = b - A x0: the reference (:1571,:1619) */
#define CICY_REL_B 1   /* ||b - A x|| <= tol*||b||, b the zero-guess right-hand side */

typedef struct cicy_solve_stats {
    int64_t iters;          /* CG iterations (the reference's k, Direct_Electrifying.cpp:1642) */
    int status;             /* CICY_OK, CICY_MAXITER, CICY_BREAKDOWN, ... */
    int refinements;        /* FP64 true-residual recomputations (mixed method) */
    double rel_res;         /* final ||r_k|| / ||r_0|| of the recursive residual */
    double true_rel_res;    /* ||b - A x|| / ||b|| when b was given, else -1 */
    double solve_ms;        /* device time of the solve (HIP events), inputs resident */
    double total_ms;        /* wall time of the call including host<->device copies */
    int64_t fp64_iters;     /* mixed method: iterations finished in FP64 after the FP32
                               recursion stopped making progress (0: none) */
    double kernel_ms;       /* device time (HIP events) of the launches of the solve's
                               dominant kernel, the persistent core CG of
                               CICY_METHOD_CONDENSED_BLOCK; 0 when not measured */
} cicy_solve_stats;

/* ---- context ---------------------------------------------------------------------- */
int cicy_create(cicy_ctx **out, int device);
void cicy_destroy(cicy_ctx *ctx);
const char *cicy_last_error(const cicy_ctx *ctx);
/* rows per work-group of the SSS kernels (0 = autotuned/default); see cicy_autotune */
int cicy_set_rows_per_block(cicy_ctx *ctx, int rows_per_block);
/* device layout of systems created afterwards: 0 auto (band + remainder when at least half
 * of the stored entries sit on the sub-diagonal, as for tube-by-tube numbering), 1 blocked
 * SSS with a remote transpose list, 2 band + remainder */
int cicy_set_format(cicy_ctx *ctx, int format);
/* method cicy_solve_dea_cg_sss solves with (default CICY_METHOD_PCG_F64, the reference's
 * algorithm, iteration for iteration).  CICY_METHOD_CONDENSED / _CONDENSED_RESIDENT /
 * _RESIDENT keep the call's contract (stop when ||r|| <= tol ||r0||, R out = b - A x) and are
 * far faster on CNT networks; the iteration count then counts CG iterations of the system
 * the method iterates on.  No counterpart in the reference. */
int cicy_set_dropin_method(cicy_ctx *ctx, int method);
/* Row groups (ranges[2 b] .. ranges[2 b + 1], free-node rows) for the next
 * cicy_solve_dea_cg_sss call with CICY_METHOD_CONDENSED_BLOCK, as cicy_sys_set_blocks takes
 * them: the caller that assembled the system knows the nodes of each nanoplatelet (the
 * reference's LMM_gnps over structure_gnp[g], Direct_Electrifying.cpp:307-349), the seam
 * itself only sees the matrix.  Consumed (cleared) by that call.  No counterpart in the
 * reference. */
int cicy_set_dropin_blocks(cicy_ctx *ctx, int64_t nblocks, const int64_t *ranges);
/* rows per thread of the register-resident layout built for systems created afterwards
 * (0: the smallest of 4 8 16 24 32 40 44 48 whose grid is co-resident).  No counterpart in
 * the reference: a layout knob like cicy_set_rows_per_block. */
int cicy_set_resident_rows(cicy_ctx *ctx, int rows_per_thread);

/* ---- drop-in seam: Solve_DEA_equations_CG_SSS (Direct_Electrifying.cpp:1538-1659) -----
 * nodes          global node count (reserved + free), as the reference passes it
 * reserved_nodes number of Dirichlet nodes R_d (2, 4 or 6)
 * tolerance      relative reduction of the initial recursive residual
 * col_ind, row_ptr, values, diagonal   SSS arrays of the free block (n = nodes-R_d)
 * R              [n] in: b - A x0; out: final recursive residual
 * VEF            [R_d] prescribed voltages
 * voltages_sol   [n] in: x0; out: solution
 * voltages       [nodes] out: [VEF | x] (the class member `voltages`, :1646-1656)
 * nnz            length of col_ind/values (the reference's vector sizes)
 * max_iter <= 0 selects the reference's cap 300*nodes (:1559). */
int cicy_solve_dea_cg_sss(cicy_ctx *ctx, int64_t nodes, int64_t reserved_nodes,
                          double tolerance, const int64_t *col_ind, const int64_t *row_ptr,
                          const double *values, const double *diagonal, int64_t nnz,
                          double *R, const double *VEF, double *voltages_sol,
                          double *voltages, int64_t max_iter, cicy_solve_stats *stats);

/* ---- Update_residual_vector (Direct_Electrifying.cpp:1502-1535): R <- R - A x0 ---------- */
int cicy_update_residual(cicy_ctx *ctx, int64_t n, const int64_t *row_ptr,
                         const int64_t *col_ind, const double *values,
                         const double *diagonal, int64_t nnz, const double *x0, double *R);

/* ---- spM_V_SSS (Direct_Electrifying.cpp:1678-1695): out = A v, host vectors ------------- */
int cicy_spmv_sss_f64(cicy_ctx *ctx, int64_t n, const int64_t *row_ptr,
                      const int64_t *col_ind, const double *values, const double *diagonal,
                      int64_t nnz, const double *v, double *out);

/* ---- boundary current (Electrical_analysis.cpp:547-605) ------------------------------
 * I = sum_e (voltages[n_int[e]] - voltages[n_bnd[e]]) / re[e] in FP64 on the device.
 * voltages has nv entries (the reference's DEA->voltages, reserved nodes first). */
int cicy_boundary_current_f64(cicy_ctx *ctx, int64_t ne, const int64_t *n_bnd,
                              const int64_t *n_int, const double *re,
                              const double *voltages, int64_t nv, double *current);
/* The same sum with every array in device memory (n_bnd, n_int, re: ne entries; voltages: nv
 * entries; current: one double, written) and no allocation at all: no hipMalloc, hipFree or
 * allocating library call.  It has no counterpart in the reference; it is for callers that
 * cannot allocate (code under graph capture, tight serving loops).  Its reduction words sit in
 * a file-scope device array of about 20 KB, one copy per GPU, held until the process ends;
 * every call on a GPU shares that GPU's copy, so calls that overlap on one GPU race.  Runs on
 * the context's stream and returns after it has finished.  Sizes above 2^40 and a node index
 * outside [0, nv) are refused: the call returns 0 and cicy_last_error says why.  The result
 * is bit-identical to cicy_boundary_current_f64's on the same data (same kernel and grid). */
int cicy_boundary_current_f64_noalloc(cicy_ctx *ctx, int64_t ne, const int64_t *n_bnd,
                                      const int64_t *n_int, const double *re,
                                      const double *voltages, int64_t nv, double *current);

/* ---- backbone extraction (Backbone_Network.cpp) ----------------------------------------
 * The data-parallel part of Backbone_Network::Determine_backbone_network
 * (Backbone_Network.cpp:10-61) from the voltages [nvolt] of a cluster's unit-resistor solve:
 * zero_out = 1e-9 * max(1e-10, largest |V[a] - V[b]|) over the elements of every tube
 * (tube t's element points' nodes are tube_nodes[tube_off[t] .. tube_off[t + 1]], in
 * order) and over the npairs other elements (tube and plate junctions, plate triangulation
 * edges: pair_a[q], pair_b[q]) -- Find_zero_current, :64-221 -- and, per tube, the run of
 * its elements from the first whose |dV| exceeds zero_out by more than 1e-10 to the next one
 * that does not (seg_out[2 t], seg_out[2 t + 1], -1 when absent; Find_backbone_cnts,
 * :317-409).  The ordered pruning that uses them is the host's (cicy_net_backbone in the
 * workload library).  Node ids are checked against nvolt on the host. */
int cicy_backbone_scan(cicy_ctx *ctx, int64_t nvolt, const double *volt, int64_t ntubes,
                       const int64_t *tube_off, const int64_t *tube_nodes, int64_t npairs,
                       const int64_t *pair_a, const int64_t *pair_b, double *zero_out,
                       int32_t *seg_out);

/* ---- tube contact search and cluster labels (Contact_grid.cpp, Hoshen_Kopelman.cpp) -----
 * Replaces Contact_grid::Generate_contact_grid's tube part (Contact_grid.cpp:18-295) and
 * Hoshen_Kopelman::Make_cnt_clusters (Hoshen_Kopelman.cpp:203-247: Label_cnts_in_window,
 * Cleanup_labels, Compress_cnt_cnt_contact_segments, Complete_cnt_elements).
 * Points: xyz[3 n_points]; tube i holds struct_pts[struct_off[i] .. struct_off[i + 1]] (the
 * reference's structure[i]; a point belongs to one tube), radius radii[i]; cnts_inside lists
 * the tubes of the window in the reference's order.  geom[11] = window corner[3], window
 * lengths[3], gs_min[3] (background grid), gs_overlap_cnt, gs_overlap_gnp (the larger of the
 * two sizes the subregions, the first places points in neighbours).  Cutoffs: vdw, tunnel,
 * min_points.
 * Out: labels[n_cnts] (the reference's labels_cnt: proper labels 0.. in its numbering, -1
 * without contact), *n_labels; junctions in the reference's order, junc[4 q..] = (Pi, tube
 * i, Pj, tube j), junc_d[q] = junction distance, *n_junc (when it exceeds junc_cap the call
 * fails after setting it); elements (elements_cnt) as elem_off[n_cnts + 1] into elem_pts,
 * ascending point ids (elem_pts may be null; it needs at most struct_off[n_cnts] entries).
 * stats[8] (may be null): subregions, slots, pair tests, contacts, pair-kernel ms, host
 * subregion ms, device ms (copies and kernels), host labelling ms.  Results are the
 * reference's bit for bit: the pairs are visited in its order and the distance uses the
 * same IEEE operations. */
int cicy_cnt_contacts(cicy_ctx *ctx, int64_t n_points, const double *xyz, int64_t n_cnts,
                      const double *radii, const int64_t *struct_off, const int64_t *struct_pts,
                      int64_t n_inside, const int64_t *cnts_inside, const double *geom,
                      double vdw, double tunnel, int64_t min_points, int32_t *labels,
                      int64_t *n_labels, int64_t junc_cap, int64_t *n_junc, int64_t *junc,
                      double *junc_d, int64_t *elem_off, int64_t *elem_pts, double *stats);

/* ---- device-resident systems (repeated solves, benchmarking) --------------------------
 * cicy_sys_create uploads one SSS system (validated on the host) and keeps it in HBM.
 * The reference instead rebuilds and re-solves a fresh object per direction/cluster
 * (Electrical_analysis.cpp:42, :252); here the device copy is reused across solves. */
int cicy_sys_create(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                    const int64_t *col_ind, const double *values, const double *diagonal,
                    cicy_sys **out);
/* Builds what `method` needs on top of the uploaded system (the FP32 copy, the resident
 * layout, the condensation and its core solver's layout); idempotent.  cicy_sys_solve
 * calls it, so it is optional: it moves that set-up out of the first solve.  0 on a device
 * error or when the method cannot run on this system (e.g. nothing to condense).  No
 * counterpart in the reference. */
int cicy_sys_prepare(cicy_sys *sys, int method);
void cicy_sys_destroy(cicy_sys *sys);
int64_t cicy_sys_size(const cicy_sys *sys, int what); /* 0 n, 1 nnz, 2 remote entries,
                                                           3 device bytes, 4 format,
                                                           5 off-layout entries,
                                                           6 resident rows per thread (0:
                                                           no resident layout), 7 resident
                                                           work-groups, 8 published rows,
                                                           9 rows after path elimination
                                                           (0: no condensed system), 10 their
                                                           stored entries, 11 core rows (the
                                                           system CG runs on), 12 core stored
                                                           entries, 13 series rounds,
                                                           14 core rows in dense blocks,
                                                           15 dense blocks, 16 work-groups
                                                           of the block-preconditioned core
                                                           CG (0: not built), 17 entries of
                                                           its dense block inverses;
                                                           6-17 describe layouts built by
                                                           cicy_sys_prepare (0 before) */
/* r0 = b - A x0 (the reference's R), x0 (may be NULL = zeros), b (may be NULL; needed for
 * the true residual and for CICY_REL_B). */
int cicy_sys_set_rhs(cicy_sys *sys, const double *r0, const double *x0, const double *b);

/* The same system from the reference's arrays already in device memory (the arrays
 * From_2d_to_1d_vectors builds, Direct_Electrifying.cpp:1329-1394, with int64 indices as
 * its std::vector<long>): they are checked and laid out (band + remainder) on the device,
 * and the condensation, series rounds and core layout of the condensed methods are built on
 * the device too, so a new system costs no host work proportional to its size.  The
 * caller's arrays are only read during the call.  Malformed arrays give 0 with the first
 * bad row in cicy_last_error, as cicy_sys_create.  No host copy is kept: the methods whose
 * layouts are built on the host (CICY_METHOD_MIXED, CICY_METHOD_RESIDENT, the blocked SSS
 * format) are refused for such a system.  No counterpart in the reference. */
int cicy_sys_create_device(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *d_row_ptr,
                           const int64_t *d_col_ind, const double *d_values,
                           const double *d_diagonal, cicy_sys **out);
/* cicy_sys_set_rhs with device arrays (d_x0, d_b may be NULL); the reference norms are
 * summed on the device.  No counterpart in the reference. */
int cicy_sys_set_rhs_device(cicy_sys *sys, const double *d_r0, const double *d_x0,
                            const double *d_b);
/* Where the set-up of systems created afterwards runs: 0 on the device (default; the
 * condensation of systems in the band layout), 1 on the host (the host forms, bit-identical;
 * for checking one against the other).  No counterpart in the reference. */
int cicy_set_setup(cicy_ctx *ctx, int where);
/* Plain device memory for callers that hold no device allocator of their own (tests,
 * benchmarks): allocate, copy in and out, free.  No counterpart in the reference. */
int cicy_device_alloc(cicy_ctx *ctx, int64_t bytes, void **out);
int cicy_device_free(cicy_ctx *ctx, void *p);
int cicy_device_upload(cicy_ctx *ctx, void *dst, const void *src, int64_t bytes);
int cicy_device_download(cicy_ctx *ctx, void *dst, const void *src, int64_t bytes);
/* Solve on device-resident data.  max_iter <= 0: 300*(n+2). */
int cicy_sys_solve(cicy_sys *sys, int method, double tol, int64_t max_iter, int rel_to,
                   cicy_solve_stats *stats);
int cicy_sys_get(cicy_sys *sys, double *x, double *r);
/* Row groups of the full system whose core rows CICY_METHOD_CONDENSED_BLOCK
 * preconditions with the exact inverse of their diagonal block: ranges[2 b] .. ranges[2 b
 * + 1] (end exclusive) of free-node rows, non-overlapping.  The reference numbers the
 * points of each nanoplatelet consecutively (LM_matrix_for_gnps, Direct_Electrifying.cpp:
 * 307-349), so each plate's node range is one group; its edges conduct ~1e7 times more than
 * its junctions, which Jacobi scaling cannot balance.  Groups longer than 128 core rows are
 * split.  nblocks = 0 gives Jacobi everywhere.  Needs a condensed system (cicy_sys_size 9
 * > 0).  No counterpart in the reference. */
int cicy_sys_set_blocks(cicy_sys *sys, int64_t nblocks, const int64_t *ranges);
/* The core system a CICY_METHOD_CONDENSED solve runs CG on, as a borrowed handle (owned by
 * sys, valid while sys lives; do not destroy it), for measuring its kernels; 0 if sys has
 * no condensed system.  No counterpart in the reference. */
int cicy_sys_core(cicy_sys *sys, cicy_sys **core);
/* Time `reps` back-to-back SSS SpMV launches on resident data; average ms per launch. */
int cicy_sys_spmv_bench(cicy_sys *sys, int precision /*64 or 32*/, int reps, double *avg_ms);
/* The same with the memory-side (Infinity) cache flushed before every launch: a device copy
 * of flush_bytes (use more than twice its 256 MB) runs before each timed launch, and each
 * launch is timed alone, so the matrix and vectors come from HBM. */
int cicy_sys_spmv_bench_cold(cicy_sys *sys, int precision, int reps, int64_t flush_bytes,
                             double *avg_ms);
/* Time a plain device copy of `bytes` (read `bytes`, write `bytes`); the reference kernel for
 * traffic counters and the achievable-bandwidth yardstick. */
int cicy_copy_bench(cicy_ctx *ctx, int64_t bytes, int reps, double *avg_ms);
/* Time a read-mostly stream over `bytes` (read `bytes`, write a word per work-group): the
 * achievable-bandwidth yardstick of a kernel that mostly reads, as an SpMV does. */
int cicy_read_bench(cicy_ctx *ctx, int64_t bytes, int reps, double *avg_ms);

/* ---- batches of independent systems (clusters x directions x realisations) -----------
 * The reference solves each (cluster, direction) system in turn with a fresh object
 * (Electrical_analysis.cpp:34-105, :249-283).  A batch holds many systems of any sizes in
 * HBM and advances all of them in the same launches; each keeps its own scalars, stopping
 * test and iteration count.  Arrays are per system, as in cicy_sys_create. */
typedef struct cicy_batch cicy_batch;
int cicy_batch_create(cicy_ctx *ctx, int nsys, const int64_t *n, const int64_t *nnz,
                      const int64_t *const *row_ptr, const int64_t *const *col_ind,
                      const double *const *values, const double *const *diag,
                      cicy_batch **out);
void cicy_batch_destroy(cicy_batch *batch);
/* 0 total rows, 1 total nnz, 2 remote entries, 3 device bytes, 4 systems, 5 row blocks */
int64_t cicy_batch_size(const cicy_batch *batch, int what);
/* r0 = b - A x0; x0 may be NULL (zeros, then b defaults to r0); b is required otherwise */
int cicy_batch_set_rhs(cicy_batch *batch, int sys, const double *r0, const double *x0,
                       const double *b);
/* per_sys[nsys] receives each system's stats; total_ms the device time of the batch */
int cicy_batch_solve(cicy_batch *batch, int method, double tol, int64_t max_iter, int rel_to,
                     cicy_solve_stats *per_sys, double *total_ms);
int cicy_batch_get(cicy_batch *batch, int sys, double *x, double *r);

/* ---- learned warm start (no counterpart in the reference; its warm start is the ramp of
 * Initial_guess_for_CG, Direct_Electrifying.cpp:1396-1470) ------------------------------
 * A 3-layer message-passing network (32 channels, BF16 weights, FP32 accumulation) over
 * the conductance graph predicts the voltage field.  Parameters are a flat FP32 vector of
 * cicy_mpnn_num_params() entries (layout in dea_mpnn.hip); training runs on the device:
 * forward (with loss) -> backward (accumulates gradients) -> Adam step. */
typedef struct cicy_mpnn cicy_mpnn;
typedef struct cicy_graph cicy_graph;
int64_t cicy_mpnn_num_params(void);
int cicy_mpnn_create(cicy_ctx *ctx, const float *theta /* may be NULL: zeros */,
                     cicy_mpnn **out);
void cicy_mpnn_destroy(cicy_mpnn *mpnn);
int cicy_mpnn_set_params(cicy_mpnn *mpnn, const float *theta);
int cicy_mpnn_get_params(cicy_mpnn *mpnn, float *theta);
int cicy_mpnn_get_grads(cicy_mpnn *mpnn, float *grad);
int cicy_mpnn_zero_grads(cicy_mpnn *mpnn);
/* Graph of one SSS system: node_xyz[n*3] are the free nodes' coordinates, box the RVE
 * side, direction the current direction (0 x, 1 y, 2 z), b the zero-guess right-hand
 * side.  training != 0 keeps the activations needed by backward. */
int cicy_graph_create(cicy_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr,
                      const int64_t *col_ind, const double *values, const double *diag,
                      const double *b, const double *node_xyz, double box, int direction,
                      double v_app, int training, cicy_graph **out);
void cicy_graph_destroy(cicy_graph *graph);
int cicy_graph_set_target(cicy_graph *graph, const double *v_true);
/* v_out[n] (may be NULL) receives the predicted voltages; loss (may be NULL) the MSE of
 * V^/V_app against the target (training graphs). */
int cicy_mpnn_forward(cicy_mpnn *mpnn, cicy_graph *graph, double *v_out, double *loss);
int cicy_mpnn_backward(cicy_mpnn *mpnn, cicy_graph *graph);
int cicy_mpnn_adam_step(cicy_mpnn *mpnn, double lr, double beta1, double beta2, double eps,
                        double grad_scale);
/* Predict straight into a device-resident system: x0 := V^, r0 := b - A x0 (the system
 * must have b).  ms (may be NULL) receives the device time. */
int cicy_mpnn_warm_start(cicy_mpnn *mpnn, cicy_graph *graph, cicy_sys *sys, double *ms);

/* ---- test hook: the blocked SSS layout applied on the host (no GPU) -------------------
 * Builds the device layout for rows_per_block and applies it with host loops, so the
 * layout can be checked where no GPU is present. */
int cicy_host_blocked_spmv(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                           const double *values, const double *diagonal, int64_t nnz,
                           int rows_per_block, const double *v, double *out);
/* same for either layout (format 1 blocked SSS, 2 band + remainder) */
int cicy_host_layout_spmv(int format, int64_t n, const int64_t *row_ptr,
                          const int64_t *col_ind, const double *values, const double *diagonal,
                          int64_t nnz, int rows_per_block, const double *v, double *out);
/* One product of the resident layout on the host: out = A^ (r + beta pold), A^ = D^-1/2 A
 * D^-1/2, walking lanes, waves and published rows as the resident kernel does (phase 0:
 * remote p recomputed from the published r and p_old; phase 1: published p). */
int cicy_host_resident_apply(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                             const double *values, const double *diagonal, int64_t nnz,
                             int rows_per_thread, int phase, const double *r, const double *pold,
                             double beta, double *out);
/* The condensation of CICY_METHOD_CONDENSED on the host: anchors (rows with junction
 * entries, [m]), the reduced SSS matrix S (red_row_ptr [m+1], red_col_ind/red_values
 * [*nnz_out], red_diagonal [m]) and, when rhs is given, the reduced right-hand side [m].
 * Buffers are sized by the caller for the worst case (m <= n, nnz_out <= nnz + n).
 * Returns 0 when there is nothing worth condensing.  For layout tests without a GPU. */
/* The block-preconditioned core CG of CICY_METHOD_CONDENSED_BLOCK on the host: the layout
 * (blocks, dense inverses) and the kernel's recurrences with sequential sums, on an SSS
 * system of m rows solved from x = 0 until ||r|| <= tol sqrt(ref2).  Returns 1 when it
 * converged.  For tests without a GPU. */
int cicy_host_core_block_cg(int64_t m, const int64_t *row_ptr, const int64_t *col_ind,
                            const double *values, const double *diagonal, int64_t nnz,
                            int64_t nblocks, const int64_t *ranges, const double *rhs,
                            double tol, double ref2, int64_t max_iter, double *x,
                            int64_t *iters, int64_t *blocked_rows);
int cicy_host_condense(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                       const double *values, const double *diagonal, int64_t nnz,
                       const double *rhs, int64_t *m_out, int64_t *nnz_out, int64_t *anchors,
                       int64_t *red_row_ptr, int64_t *red_col_ind, double *red_values,
                       double *red_diagonal, double *red_rhs);
/* Back-substitution of the condensation on the host: x [n] from the anchor values [m]. */
int cicy_host_expand(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                     const double *values, const double *diagonal, int64_t nnz,
                     const double *rhs, const double *x_anchor, double *x);
/* Both stages on the host: the core system the condensed method runs CG on (core rows as
 * indices of the full system [*m_out], SSS arrays, diagonal and right-hand side) and the
 * number of series rounds.  Buffer sizes as for cicy_host_condense. */
int cicy_host_core(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                   const double *values, const double *diagonal, int64_t nnz, const double *rhs,
                   int64_t *m_out, int64_t *nnz_out, int64_t *rounds_out, int64_t *core_rows,
                   int64_t *core_row_ptr, int64_t *core_col_ind, double *core_values,
                   double *core_diagonal, double *core_rhs);
/* The full solution x [n] from the core solution x_core [m] of cicy_host_core. */
int cicy_host_core_expand(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                          const double *values, const double *diagonal, int64_t nnz,
                          const double *rhs, const double *x_core, double *x);

#ifdef __cplusplus
}
#endif

#endif /* CICY_DEA_H */
