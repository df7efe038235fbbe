// Host representation of the scaled working problem handed to the device
// solver, and the diagonal scaling that produces it.
#pragma once

#include <memory>
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
    HostCSR At;       // the stable transpose of A when the device built it (else empty)
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

// The passes of the scaling that touch every matrix entry, for a device
// implementation (scale.hip): the host keeps the factor vectors and the
// vector data and hands the entry-wise work to these.
class MatrixPasses {
  public:
    virtual ~MatrixPasses() = default;
    // takes the working A, Q and R (after any Curtis-Reid pass) and builds
    // the transpose of A
    virtual void load(const ScaledProblem &w) = 0;
    // row and column maxima of |A|
    virtual void abs_max(std::vector<double> &row, std::vector<double> &col) = 0;
    // row sums of |a|^alpha and column sums of |a|^(2 - alpha) of A
    virtual void pow_sums(double alpha, std::vector<double> &row, std::vector<double> &col) = 0;
    // A /= cr vr' (and its transpose alike), Q /= vr vr', R /= 1 vr'
    virtual void divide(const std::vector<double> &cr, const std::vector<double> &vr) = 0;
    // Q *= fq, R *= fr
    virtual void multiply_qr(double fq, double fr) = 0;
    // copies the scaled A, Q and R values and the transpose of A into w
    virtual void finish(ScaledProblem &w) = 0;
};

// The device implementation (scale.hip).
std::unique_ptr<MatrixPasses> make_device_passes();

// Builds the working problem (adds the single free row of reference
// src/utils.cu:82-134 when there are no constraints) and applies Curtis-Reid,
// Ruiz, Pock-Chambolle and bound/objective scaling as reference
// src/preconditioner.c:654-830 does, in FP64.  With `dev` the passes over
// the matrix entries of Ruiz, Pock-Chambolle and the bound/objective
// scaling run there (Curtis-Reid stays on the host); the result is the same
// bit for bit.
ScaledProblem scale_problem(const pdhg_parameters_t &params, const qp_problem_t &prob, MatrixPasses *dev = nullptr);

}  // namespace pdhcg
