// Host representation of the scaled working problem handed to the device
// solver, and the diagonal scaling that produces it.
#pragma once

#include <vector>

#include "host_common.h"

namespace pdhcg {

struct HostCSR {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> v;
    int nnz() const { return (int)ci.size(); }
};

// Stable transpose: rows of the result list source rows in increasing order
// (the layout cusparseCsr2cscEx2 yields at reference src/solver_state.cu:735-777).
HostCSR transpose(const HostCSR &a);

enum class MiddleKind { None = 0, Diag = 1, Dense = 2 };

struct ScaledProblem {
    int n = 0, m = 0, k = 0;
    HostCSR A, Q, R;  // scaled [A;F], Q (n x n), R (k x n)
    std::vector<double> lb, ub, c, cl, cu, g;
    double obj_const = 0.0;
    quad_obj_type_t qtype = PDHCG_NON_Q;
    MiddleKind dkind = MiddleKind::None;
    std::vector<double> ddiag, ddense;  // D (diag or row-major k x k)
    std::vector<double> qdiag;          // DIAG_Q: diag(Q) + 1e-12
    std::vector<double> con_rescale, var_rescale;
    double cbr = 1.0, ovr = 1.0;
    double rescale_time = 0.0;
    // norms of the unscaled working problem used by the relative KKT tests
    double obj_norm = 0.0, bound_norm = 0.0;
    std::vector<double> primal_start, dual_start;  // unscaled, empty when absent
    // cone blocks of the working problem (variable and affine)
    const cone_blocks_t *cones = nullptr;
    const cone_blocks_t *affine_cones = nullptr;
};

// Builds the working problem (adds the single free row of reference
// src/utils.cu:82-134 when there are no constraints) and applies Curtis-Reid,
// Ruiz, Pock-Chambolle and bound/objective scaling as reference
// src/preconditioner.c:654-830 does, in FP64 on the host.
ScaledProblem scale_problem(const pdhg_parameters_t &params, const qp_problem_t &prob);

}  // namespace pdhcg
