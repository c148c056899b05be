// Drives the C++ drop-in (include/cicy_dea.hpp) the way Direct_Electrifying::
// Compute_voltage_field drives its CPU solver (Direct_Electrifying.cpp:80-99): a uniform
// series chain of m interior nodes between node 1 (V = 1) and node 0 (V = 0), with the
// linear-ramp style warm start, residual update, solve and boundary current.
// Prints "OK <iters> <max |V - V_exact|> <|R - R_exact|/R_exact>" and exits 0 when the
// closed form holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "cicy_dea.hpp"

int main(int argc, char **argv) {
    const long int m = argc > 1 ? atol(argv[1]) : 2000;  // interior nodes
    const double g = 2.0;                                 // element conductance
    const long int reserved = 2, nodes = m + reserved, n = m;
    // interior node i (0-based, free index) sits between i-1 and i+1; free 0 touches V=1,
    // free m-1 touches V=0.  SSS lower triangle: row i has (i, i-1).
    std::vector<long int> row_ptr(n + 1, 0), col_ind;
    std::vector<double> values, diagonal(n, 2 * g), R(n, 0.0), VEF;
    for (long int i = 0; i < n; i++) {
        if (i > 0) {
            col_ind.push_back(i - 1);
            values.push_back(-g);
        }
        row_ptr[i + 1] = (long int)col_ind.size();
    }
    cicy::DirectElectrifyingGPU dea;
    if (!dea.Get_voltage_vector(reserved, nodes, 1, 1.0, VEF)) return 2;
    R[0] = g * VEF[1];  // b = -K_EF^T VEF
    std::vector<double> V_guess(n);
    for (long int i = 0; i < n; i++) V_guess[i] = 1.0 - (double)i / n;  // ramp
    if (!dea.Update_residual_vector(col_ind, row_ptr, values, diagonal, V_guess, R)) return 3;
    if (!dea.Solve_DEA_equations_CG_SSS(nodes, reserved, 1e-12, col_ind, row_ptr, values,
                                        diagonal, R, VEF, V_guess)) {
        std::printf("FAIL solve: %s\n", dea.last_error().c_str());
        return 4;
    }
    double err = 0.0;
    for (long int i = 0; i < n; i++) {
        double exact = 1.0 - (double)(i + 1) / (m + 1);
        err = std::fmax(err, std::fabs(dea.voltages[i + reserved] - exact));
    }
    // current into node 0 through the last element
    std::vector<long int> nb = {0}, ni = {nodes - 1};
    std::vector<double> re = {1.0 / g};
    double I = 0.0;
    if (!dea.Current_through_elements(nb, ni, re, I)) return 5;
    double R_exact = (m + 1) / g, R_got = 1.0 / I;
    double rel = std::fabs(R_got - R_exact) / R_exact;
    std::printf("%s %lld %.3e %.3e\n", (err < 1e-8 && rel < 1e-8) ? "OK" : "FAIL",
                (long long)dea.stats.iters, err, rel);
    return (err < 1e-8 && rel < 1e-8) ? 0 : 1;
}
