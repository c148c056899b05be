// cicy_dea.hpp -- C++ drop-in for the solve half of rnaimehaom/CICY_geom's
// Direct_Electrifying class, header-only over the C-ABI of cicy_dea.h.
//
// The member functions keep the reference's names, argument types and meaning
// (Direct_Electrifying.h:58-70) and its 1/0 return convention, so a maintainer can swap
// the CPU calls in Direct_Electrifying::Compute_voltage_field (Direct_Electrifying.cpp:
// 89-99) for this class without touching the assembly or Electrical_analysis:
//
//   Solve_DEA_equations_CG_SSS  <- Direct_Electrifying.cpp:1538-1659
//   Update_residual_vector      <- Direct_Electrifying.cpp:1502-1535
//   spM_V_SSS                   <- Direct_Electrifying.cpp:1678-1695
//   Get_voltage_vector          <- Direct_Electrifying.cpp:1472-1500
//   Current_through_elements    <- the I = I + V/Re sum of Electrical_analysis.cpp:547-605
//
// `voltages` is filled exactly as the reference fills its member (VEF first, then x).
// Messages the reference writes to hout are available from last_error().
#ifndef CICY_DEA_HPP
#define CICY_DEA_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "cicy_dea.h"

static_assert(sizeof(long int) == sizeof(int64_t), "LP64 expected: long int is 64-bit");

namespace cicy {

class DirectElectrifyingGPU {
public:
    std::vector<double> voltages;
    cicy_solve_stats stats{};

    explicit DirectElectrifyingGPU(int device = 0) {
        if (!cicy_create(&ctx_, device)) throw std::runtime_error("cicy_create failed");
    }
    ~DirectElectrifyingGPU() { cicy_destroy(ctx_); }
    DirectElectrifyingGPU(const DirectElectrifyingGPU &) = delete;
    DirectElectrifyingGPU &operator=(const DirectElectrifyingGPU &) = delete;

    std::string last_error() const { return cicy_last_error(ctx_); }
    // solver behind Solve_DEA_equations_CG_SSS: CICY_METHOD_PCG_F64 (default, the
    // reference's algorithm), CICY_METHOD_CONDENSED or CICY_METHOD_CONDENSED_BLOCK (same
    // stopping rule, far faster)
    bool use_method(int method) { return cicy_set_dropin_method(ctx_, method) != 0; }
    // node ranges [first, end) of free nodes (node - reserved_nodes) of each nanoplatelet,
    // flattened, for the next Solve_DEA_equations_CG_SSS with CICY_METHOD_CONDENSED_BLOCK
    bool set_plate_blocks(const std::vector<long int> &ranges) {
        if (ranges.size() % 2) return false;
        return cicy_set_dropin_blocks(ctx_, (int64_t)ranges.size() / 2,
                                      reinterpret_cast<const int64_t *>(ranges.data())) != 0;
    }
    cicy_ctx *context() { return ctx_; }

    int Get_voltage_vector(const long int &reserved_nodes, const long int &nodes,
                           const int &R_flag, const double &applied_voltage,
                           std::vector<double> &VEF) {
        double V;
        if (R_flag == 1) V = applied_voltage;
        else if (R_flag == 0) V = (double)nodes;
        else return 0;
        VEF.assign((size_t)reserved_nodes, 0.0);
        for (long int i = 1; i < reserved_nodes; i++) VEF[i] = VEF[i - 1] + V;
        return 1;
    }

    int Update_residual_vector(const std::vector<long int> &col_ind,
                               const std::vector<long int> &row_ptr,
                               const std::vector<double> &values,
                               const std::vector<double> &diagonal,
                               const std::vector<double> &V_guess, std::vector<double> &R) {
        if (R.size() != diagonal.size() || V_guess.size() != diagonal.size()) return 0;
        return cicy_update_residual(ctx_, (int64_t)diagonal.size(), ptr(row_ptr), ptr(col_ind),
                                    values.data(), diagonal.data(), (int64_t)values.size(),
                                    V_guess.data(), R.data());
    }

    void spM_V_SSS(const std::vector<double> &V, const std::vector<long int> &rowptr,
                   const std::vector<long int> &colind, const std::vector<double> &diagonal,
                   const std::vector<double> &values, std::vector<double> &R) {
        R.assign(V.size(), 1.0);
        if (!cicy_spmv_sss_f64(ctx_, (int64_t)V.size(), ptr(rowptr), ptr(colind), values.data(),
                               diagonal.data(), (int64_t)values.size(), V.data(), R.data()))
            throw std::runtime_error("spM_V_SSS: " + last_error());
    }

    int Solve_DEA_equations_CG_SSS(const long int &nodes, const long int &reserved_nodes,
                                   const double &tolerance, const std::vector<long int> &col_ind,
                                   const std::vector<long int> &row_ptr,
                                   const std::vector<double> &values,
                                   const std::vector<double> &diagonal, std::vector<double> &R,
                                   std::vector<double> &VEF, std::vector<double> &voltages_sol) {
        const long int n = nodes - reserved_nodes;
        if (n < 0 || (long int)R.size() != n || (long int)voltages_sol.size() != n ||
            (long int)diagonal.size() != n || (long int)VEF.size() != reserved_nodes ||
            (long int)row_ptr.size() != n + 1)
            return 0;
        voltages.assign((size_t)nodes, 0.0);
        return cicy_solve_dea_cg_sss(ctx_, nodes, reserved_nodes, tolerance, ptr(col_ind),
                                     ptr(row_ptr), values.data(), diagonal.data(),
                                     (int64_t)values.size(), R.data(), VEF.data(),
                                     voltages_sol.data(), voltages.data(), 0, &stats);
    }

    // I = sum_e (voltages[n_int[e]] - voltages[n_bnd[e]]) / Re[e], in FP64 on the device.
    int Current_through_elements(const std::vector<long int> &n_bnd,
                                 const std::vector<long int> &n_int,
                                 const std::vector<double> &Re, double &I) {
        if (n_bnd.size() != n_int.size() || n_bnd.size() != Re.size()) return 0;
        return cicy_boundary_current_f64(ctx_, (int64_t)n_bnd.size(), ptr(n_bnd), ptr(n_int),
                                         Re.data(), voltages.data(), (int64_t)voltages.size(),
                                         &I);
    }

private:
    static const int64_t *ptr(const std::vector<long int> &v) {
        return reinterpret_cast<const int64_t *>(v.data());
    }
    cicy_ctx *ctx_ = nullptr;
};

}  // namespace cicy

#endif  // CICY_DEA_HPP
