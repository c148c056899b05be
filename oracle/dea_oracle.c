/*
 * dea_oracle.c -- CPU oracle for the Direct Electrifying (DEA) conductance solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load liboracle.so.  The product path (cicy_geom_amd, libcicy_dea.so) never
 * links or calls anything here.
 *
 * It restates, in plain C with the same loop and summation order, the CPU
 * solver of rnaimehaom/CICY_geom (Direct_Electrifying.cpp), plus an FP64
 * "truth" solver used to state the error of every result:
 *
 *   oracle_voltage_vector   <- Get_voltage_vector        Direct_Electrifying.cpp:1472-1500
 *   oracle_update_residual  <- Update_residual_vector    Direct_Electrifying.cpp:1502-1535
 *   oracle_pcg_sss          <- Solve_DEA_equations_CG_SSS Direct_Electrifying.cpp:1538-1659
 *                              + Jacobi_preconditioner / Apply_preconditioner :1662-1676
 *                              + spM_V_SSS :1678-1695, V_dot_v :1697-1712,
 *                                V_plus_aW :1715-1720, W_plus_aV :1723-1728,
 *                                Componentwise_multiply :1730-1735
 *   oracle_spmv_sss         <- spM_V_SSS                  Direct_Electrifying.cpp:1678-1695
 *   oracle_boundary_current <- I = I + V/Re accumulation  Electrical_analysis.cpp:547-605
 *   oracle_parallel_resistance <- 1/R = sum 1/R_k         Electrical_analysis.cpp:873-932
 *   oracle_resistance_cnt   <- Calculate_resistance_cnt   Direct_Electrifying.cpp:611-646
 *   oracle_junction_resistance <- Calculate_junction_resistance Direct_Electrifying.cpp:854-901
 *   oracle_tunnel_constants <- C1/C2                       Input_Reader.cpp:1450-1456
 *   oracle_cg_truth         <- not in the reference: FP64 Jacobi-CG that recomputes the
 *                              true residual b-Ax every 50 iterations, accumulates dots
 *                              with Neumaier compensation, and stops on ||b-Ax||/||b||.
 *
 * Build with -O2 -ffp-contract=off (no FMA contraction) so that every operation is a
 * single IEEE double operation in source order, as in the reference's x86-64 build.
 *
 * Parity pins (tests/test_oracle.py): closed-form chains, parallel chains and cubic
 * lattices; dense numpy/scipy solves committed as fixtures under tests/golden/;
 * current conservation, Tellegen's identity and the maximum principle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Status codes shared with include/cicy_dea.h. */
#define ORC_OK 0
#define ORC_MAXITER 1
#define ORC_BREAKDOWN 2
#define ORC_BAD_INPUT 3

/* Get_voltage_vector (Direct_Electrifying.cpp:1472-1500): VEF[0]=0, VEF[i]=VEF[i-1]+V,
 * V = applied voltage (R_flag=1) or the number of nodes (R_flag=0). */
int oracle_voltage_vector(int64_t reserved_nodes, int64_t nodes, int r_flag,
                          double applied_voltage, double *vef)
{
    double v;
    if (r_flag == 1)
        v = applied_voltage;
    else if (r_flag == 0)
        v = (double)nodes;
    else
        return 0;
    for (int64_t i = 0; i < reserved_nodes; i++)
        vef[i] = 0.0;
    for (int64_t i = 1; i < reserved_nodes; i++)
        vef[i] = vef[i - 1] + v;
    return 1;
}

/* spM_V_SSS (Direct_Electrifying.cpp:1678-1695).  Row r: out[r] = diag[r]*v[r], then for
 * each stored (r,c<r): out[r] += a*v[c]; out[c] += a*v[r].  Later rows scatter into
 * earlier ones, so the summation order is fixed by the loop. */
void oracle_spmv_sss(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                     const double *values, const double *diag, const double *v,
                     double *out)
{
    for (int64_t i = 0; i < n; i++)
        out[i] = 1.0;
    for (int64_t r = 0; r < n; r++) {
        out[r] = diag[r] * v[r];
        for (int64_t j = row_ptr[r]; j < row_ptr[r + 1]; j++) {
            int64_t c = col_ind[j];
            out[r] = out[r] + values[j] * v[c];
            out[c] = out[c] + values[j] * v[r];
        }
    }
}

/* Update_residual_vector (Direct_Electrifying.cpp:1502-1535): R <- R - A*x0, with A*x0
 * formed by the same SSS loop. */
int oracle_update_residual(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                           const double *values, const double *diag,
                           const double *x0, double *r)
{
    double *ax = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!ax)
        return 0;
    oracle_spmv_sss(n, row_ptr, col_ind, values, diag, x0, ax);
    for (int64_t i = 0; i < n; i++)
        r[i] = r[i] - ax[i];
    free(ax);
    return 1;
}

static double dot_seq(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s = s + a[i] * b[i];
    return s;
}

/* Solve_DEA_equations_CG_SSS (Direct_Electrifying.cpp:1538-1659).
 *   nodes          = global node count (reserved + free), as the reference passes it
 *   reserved_nodes = Dirichlet node count R_d
 *   r    [N]  in: b - A x0, out: final recursive residual
 *   x    [N]  in: x0,       out: solution
 *   voltages [nodes] out: [VEF | x]                         (:1646-1656)
 * Loop: k = 1..300*nodes (:1559); stop when sqrt(r.r) <= tol*||r_init|| (:1571, :1619).
 * rr0 = Y.R is recomputed at the top of each iteration (:1582); it equals the previous
 * iteration's beta numerator bit for bit, so it is kept only for fidelity.
 * Returns 1 (reference success convention, including the max-iteration case) and
 * writes the iteration count k as the reference prints it (:1642).
 * status: ORC_OK, ORC_MAXITER, ORC_BREAKDOWN (p.Ap not finite, the case the reference
 * spins on until max_iter), ORC_BAD_INPUT. */
int oracle_pcg_sss(int64_t nodes, int64_t reserved_nodes, double tol,
                   const int64_t *row_ptr, const int64_t *col_ind,
                   const double *values, const double *diag, double *r,
                   const double *vef, double *x, double *voltages,
                   int64_t max_iter_override, int64_t *iters_out,
                   int *status_out, double *rel_res_out)
{
    int64_t n = nodes - reserved_nodes;
    if (n < 0 || reserved_nodes < 0) {
        if (status_out)
            *status_out = ORC_BAD_INPUT;
        return 0;
    }
    size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double *minv = (double *)malloc(bytes);
    double *y = (double *)malloc(bytes);
    double *p = (double *)malloc(bytes);
    double *ap = (double *)malloc(bytes);
    if (!minv || !y || !p || !ap) {
        free(minv); free(y); free(p); free(ap);
        return 0;
    }
    /* Jacobi_preconditioner / Apply_preconditioner (:1662-1676) */
    for (int64_t i = 0; i < n; i++)
        minv[i] = 1 / diag[i];
    for (int64_t i = 0; i < n; i++) {
        p[i] = minv[i] * r[i];
        y[i] = p[i];
    }
    int64_t max_iter = max_iter_override > 0 ? max_iter_override : 300 * nodes;
    double rdotr = dot_seq(n, r, r);
    double r0 = tol * sqrt(rdotr);
    double rr0 = 0.0, rr = rdotr, alpha, beta;
    int status = ORC_MAXITER;
    int64_t k;
    for (k = 1; k <= max_iter; k++) {
        oracle_spmv_sss(n, row_ptr, col_ind, values, diag, p, ap);
        rr0 = dot_seq(n, y, r);
        double pap = dot_seq(n, p, ap);
        alpha = rr0 / pap;
        if (!isfinite(alpha)) {
            status = ORC_BREAKDOWN;
            break;
        }
        for (int64_t i = 0; i < n; i++)
            x[i] = x[i] + alpha * p[i];
        for (int64_t i = 0; i < n; i++)
            r[i] = r[i] + (-alpha) * ap[i];
        rr = dot_seq(n, r, r);
        if (sqrt(rr) <= r0) {
            status = ORC_OK;
            break;
        }
        for (int64_t i = 0; i < n; i++)
            y[i] = minv[i] * r[i];
        beta = dot_seq(n, y, r);
        beta = beta / rr0;
        for (int64_t i = 0; i < n; i++)
            p[i] = y[i] + beta * p[i];
    }
    if (voltages) {
        for (int64_t i = 0; i < reserved_nodes; i++)
            voltages[i] = vef[i];
        for (int64_t i = reserved_nodes; i < nodes; i++)
            voltages[i] = x[i - reserved_nodes];
    }
    if (iters_out)
        *iters_out = k;
    if (status_out)
        *status_out = status;
    if (rel_res_out)
        *rel_res_out = rdotr > 0 ? sqrt(rr) / sqrt(rdotr) : 0.0;
    free(minv); free(y); free(p); free(ap);
    return 1;
}

/* Neumaier-compensated dot product. */
static double dot_comp(int64_t n, const double *a, const double *b)
{
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double t = a[i] * b[i];
        double u = s + t;
        if (fabs(s) >= fabs(t))
            c += (s - u) + t;
        else
            c += (t - u) + s;
        s = u;
    }
    return s + c;
}

/* FP64 truth solver: Jacobi-CG from x (in/out) on A x = b, recomputing the true
 * residual b - A x every `refresh` iterations, stopping when ||b-Ax||/||b|| <= tol or
 * when the true residual has not halved over the last `stall` iterations (the attainable
 * accuracy of FP64 CG on the system has been reached).  The best iterate seen at a
 * true-residual check is returned.  Returns the number of iterations; writes the final
 * true relative residual. */
int64_t oracle_cg_truth(int64_t n, const int64_t *row_ptr, const int64_t *col_ind,
                        const double *values, const double *diag, const double *b,
                        double *x, double tol, int64_t max_iter, double *rel_res_out)
{
    size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
    double *r = (double *)malloc(bytes), *z = (double *)malloc(bytes);
    double *p = (double *)malloc(bytes), *ap = (double *)malloc(bytes);
    int64_t k = 0;
    const int refresh = 50;
    double bnorm = sqrt(dot_comp(n, b, b));
    if (bnorm == 0.0) {
        for (int64_t i = 0; i < n; i++)
            x[i] = 0.0;
        if (rel_res_out)
            *rel_res_out = 0.0;
        free(r); free(z); free(p); free(ap);
        return 0;
    }
    oracle_spmv_sss(n, row_ptr, col_ind, values, diag, x, ap);
    for (int64_t i = 0; i < n; i++) {
        r[i] = b[i] - ap[i];
        z[i] = r[i] / diag[i];
        p[i] = z[i];
    }
    double rz = dot_comp(n, r, z);
    double rel = sqrt(dot_comp(n, r, r)) / bnorm;
    double *xbest = (double *)malloc(bytes);
    memcpy(xbest, x, bytes);
    double best = rel;
    int64_t k_best = 0;
    const int64_t stall = 20 * refresh + n / 4;
    while (rel > tol && k < max_iter) {
        oracle_spmv_sss(n, row_ptr, col_ind, values, diag, p, ap);
        double alpha = rz / dot_comp(n, p, ap);
        for (int64_t i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * ap[i];
        }
        k++;
        if (k % refresh == 0) {
            oracle_spmv_sss(n, row_ptr, col_ind, values, diag, x, ap);
            for (int64_t i = 0; i < n; i++)
                r[i] = b[i] - ap[i];
            double tr = sqrt(dot_comp(n, r, r)) / bnorm;
            if (tr < 0.5 * best) {
                best = tr;
                k_best = k;
                memcpy(xbest, x, bytes);
            } else if (tr < best) {
                best = tr;
                memcpy(xbest, x, bytes);
            }
            if (k - k_best > stall)
                break;
        }
        rel = sqrt(dot_comp(n, r, r)) / bnorm;
        if (rel <= tol) {
            /* confirm with the true residual before stopping */
            oracle_spmv_sss(n, row_ptr, col_ind, values, diag, x, ap);
            for (int64_t i = 0; i < n; i++)
                r[i] = b[i] - ap[i];
            rel = sqrt(dot_comp(n, r, r)) / bnorm;
            if (rel <= tol)
                break;
        }
        for (int64_t i = 0; i < n; i++)
            z[i] = r[i] / diag[i];
        double rz_new = dot_comp(n, r, z);
        double beta = rz_new / rz;
        rz = rz_new;
        for (int64_t i = 0; i < n; i++)
            p[i] = z[i] + beta * p[i];
    }
    if (rel > best) {
        memcpy(x, xbest, bytes);
        rel = best;
    }
    if (rel_res_out)
        *rel_res_out = rel;
    free(r); free(z); free(p); free(ap); free(xbest);
    return k;
}

/* Boundary current (Electrical_analysis.cpp:547-605): for each boundary element e in
 * caller order, I = I + (V[n_int[e]] - V[n_bnd[e]]) / R_e.  Takes resistances, as the
 * reference divides by Re. */
double oracle_boundary_current(int64_t ne, const int64_t *n_bnd, const int64_t *n_int,
                               const double *re, const double *voltages)
{
    double cur = 0.0;
    for (int64_t e = 0; e < ne; e++) {
        double v = voltages[n_int[e]] - voltages[n_bnd[e]];
        cur = cur + v / re[e];
    }
    return cur;
}

/* Calculate_resistances_and_resistivities (Electrical_analysis.cpp:873-932) for one
 * direction: 0 clusters -> matrix value, 1 -> that R, >1 -> 1/sum(1/R_k) accumulated
 * in order. */
double oracle_parallel_resistance(int64_t nk, const double *rk, double r_matrix)
{
    if (nk == 0)
        return r_matrix;
    if (nk == 1)
        return rk[0];
    double inv = 0.0;
    for (int64_t k = 0; k < nk; k++)
        inv = inv + 1 / rk[k];
    return 1 / inv;
}

/* Calculate_resistance_cnt (Direct_Electrifying.cpp:611-646): R = rho*L/(PI*r*r), L the
 * polyline length from point p1 to p2 (p1 < p2), points as xyz triples. */
double oracle_resistance_cnt(const double *xyz, int64_t p1, int64_t p2, double radius,
                             double rho)
{
    const double pi = 3.1415926535897932384626433832795;
    double length = 0;
    for (int64_t i = p1; i < p2; i++) {
        double dx = xyz[3 * i] - xyz[3 * (i + 1)];
        double dy = xyz[3 * i + 1] - xyz[3 * (i + 1) + 1];
        double dz = xyz[3 * i + 2] - xyz[3 * (i + 1) + 2];
        length = length + sqrt(dx * dx + dy * dy + dz * dz);
    }
    return rho * length / (pi * radius * radius);
}

/* C1, C2 of the tunnelling law (Input_Reader.cpp:1450-1456), inputs as the mantissas
 * the input file gives (h e-34, e e-19, m e-31, lambda eV). */
void oracle_tunnel_constants(double h, double e, double m, double lambda, double *c1,
                             double *c2)
{
    const double pi = 3.1415926535897932384626433832795;
    double s = sqrt(2.0 * m * lambda * e);
    *c1 = 10.0 * h * h / (m * m * s);
    *c2 = 4000.0 * pi * s / h;
}

/* Calculate_junction_resistance, exponential branch (Direct_Electrifying.cpp:869-890):
 * l1, l2 are radii (CNT) or thicknesses (GNP); a CNT side uses its diameter.  The area
 * is the square of the smaller of the two. */
double oracle_junction_resistance(double dist, double l1, int cnt1, double l2, int cnt2,
                                  double c1, double c2)
{
    const double zero = 1.0e-10;
    double jl1 = l1, jl2 = l2;
    if (cnt1)
        jl1 = jl1 + l1;
    if (cnt2)
        jl2 = jl2 + l2;
    double a = (jl1 - jl2 < zero) ? jl1 * jl1 : jl2 * jl2;
    return c1 * dist * exp(c2 * dist) / a;
}
