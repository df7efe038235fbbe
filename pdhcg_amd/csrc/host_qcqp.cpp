// Lifting of diagonal quadratic constraints into second-order cones
// (reference src/qcqp_transform.c:96-552; entry point include/pdhcg.h:53).
//
// A quadratic constraint on row r,  a'x + x'Q x <= b  with Q diagonal and
// positive semidefinite (or  a'x + x'Q x >= b  with Q negative semidefinite,
// which is negated first), becomes one cone block over new variables:
//   v_j = sqrt(2 |q_jj|) x_j            (one equality row per nonzero q_jj)
//   rotated:  (v, s, t) in RSOC,  a'x + s = b,      t = 1   (|v|^2 <= 2 s t)
//   standard: (v, w, z) in SOC,   sqrt2 a'x + w + z = sqrt2 b,  z - w = sqrt2
// so that |v|^2 / 2 = sum |q_jj| x_j^2 is bounded by b - a'x.  A row of A
// with no entries (the constraint is x'Q x <= b alone) pins the two aux
// slots instead of adding coupling entries: (s, t) = (b, 1), or
// (w, z) = ((b - 1)/sqrt2, (b + 1)/sqrt2).  The lifted problem has
// n + sum(k_i) + 2K variables and m + sum(k_i) + K rows, keeps the
// objective, and its cone blocks are only the lifted ones (the layout and the
// rejections follow the reference, so results map back by truncation).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_common.h"

namespace pdhcg {

namespace {

struct QcBlock {
    std::vector<int> cols;      // variables with a nonzero diagonal entry, in row order
    std::vector<double> scale;  // sqrt(2 |q_jj|)
    bool flip = false;          // the >= form of a negative semidefinite Q
    double b = 0.0;             // right-hand side of the <= form
    bool pinned = false;        // the constraint's row of A is empty
    int start = 0;              // first lifted slot
};

// Nonzero diagonal of Q_i with one sign; false for an off-diagonal entry or
// mixed signs (reference src/qcqp_transform.c:59-94).
bool diagonal_of(const CsrComponent *Q, int n, int nnz, QcBlock &blk, int &sign) {
    sign = 0;
    for (int r = 0; r < n; ++r)
        for (int e = Q->row_ptr[r]; e < Q->row_ptr[r + 1]; ++e) {
            if (Q->col_ind[e] != r) return false;
            const double v = Q->val[e];
            if (v == 0.0) continue;
            const int s = v > 0.0 ? 1 : -1;
            if (sign == 0) sign = s;
            else if (sign != s) return false;
            if ((int)blk.cols.size() >= nnz) return false;
            blk.cols.push_back(r);
            blk.scale.push_back(std::fabs(v));
        }
    return true;
}

CsrComponent *copy_csr(const CsrComponent *src, int rows, int rows_out, int nnz) {
    CsrComponent *c = ccalloc<CsrComponent>(1);
    c->row_ptr = cmalloc<int>((size_t)rows_out + 1);
    std::memcpy(c->row_ptr, src->row_ptr, ((size_t)rows + 1) * sizeof(int));
    for (int r = rows + 1; r <= rows_out; ++r) c->row_ptr[r] = src->row_ptr[rows];
    if (nnz > 0) {
        c->col_ind = cmalloc<int>(nnz);
        c->val = cmalloc<double>(nnz);
        std::memcpy(c->col_ind, src->col_ind, (size_t)nnz * sizeof(int));
        std::memcpy(c->val, src->val, (size_t)nnz * sizeof(double));
    }
    return c;
}

}  // namespace

qp_problem_t *lift_quadratic_constraints(const qp_problem_t *in, cone_type_t form) {
    if (in->num_quadratic_constraints == 0) {
        std::fprintf(stderr, "[qcqp_to_socp_qp] no quadratic constraints; nothing to do.\n");
        return nullptr;
    }
    if (in->affine_cones.num_cones > 0) {
        std::fprintf(stderr, "[qcqp_to_socp_qp] native affine cones combined with quadratic constraints "
                             "are not supported by this transform.\n");
        return nullptr;
    }
    const int n = in->num_variables, m = in->num_constraints, K = in->num_quadratic_constraints;
    const bool standard = form == CONE_STANDARD_SOC;
    const double kRoot2 = 1.4142135623730951;
    std::vector<QcBlock> blocks((size_t)K);
    long long lifted = 0;
    for (int i = 0; i < K; ++i) {
        QcBlock &blk = blocks[i];
        int sign = 0;
        if (!diagonal_of(in->quadratic_constraint_matrices[i], n, in->quadratic_constraint_matrix_num_nonzeros[i], blk,
                         sign)) {
            std::fprintf(stderr, "[qcqp_to_socp_qp] Q_%d is non-diagonal or has mixed signs; diagonal PSD (or "
                                 "all-NSD with >= sense) required.\n", i);
            return nullptr;
        }
        const int row = in->quadratic_constraint_row_indices[i];
        const double lo = in->constraint_lower_bound[row], hi = in->constraint_upper_bound[row];
        if (sign >= 0) {
            if (std::isfinite(lo) || !std::isfinite(hi)) {
                std::fprintf(stderr, "[qcqp_to_socp_qp] QC row %d (Q PSD) requires one-sided <= (lhs=-inf, rhs "
                                     "finite); got lhs=%.3g rhs=%.3g.\n", row, lo, hi);
                return nullptr;
            }
            blk.b = hi;
        } else {
            if (!std::isfinite(lo) || std::isfinite(hi)) {
                std::fprintf(stderr, "[qcqp_to_socp_qp] QC row %d (Q NSD) requires one-sided >= (lhs finite, rhs="
                                     "+inf); got lhs=%.3g rhs=%.3g.\n", row, lo, hi);
                return nullptr;
            }
            blk.flip = true;
            blk.b = -lo;
        }
        for (double &s : blk.scale) s = std::sqrt(2.0 * s);
        blk.pinned = in->constraint_matrix->row_ptr[row + 1] == in->constraint_matrix->row_ptr[row];
        lifted += (long long)blk.cols.size();
    }
    const long long n_out = (long long)n + lifted + 2LL * K, m_out = (long long)m + lifted + K;
    const int per_row = standard ? 2 : 1;  // coupling entries of a QC row and of its closing row
    long long unpinned = 0;
    for (const QcBlock &b : blocks) unpinned += b.pinned ? 0 : 1;
    const long long nnz_out = (long long)in->constraint_matrix_num_nonzeros + 2LL * per_row * unpinned + 2LL * lifted;
    if (n_out > INT32_MAX || m_out > INT32_MAX || nnz_out > INT32_MAX) {
        std::fprintf(stderr, "[qcqp_to_socp_qp] extended problem size overflows int32.\n");
        return nullptr;
    }

    qp_problem_t *out = ccalloc<qp_problem_t>(1);
    try {
        out->num_variables = (int)n_out;
        out->num_constraints = (int)m_out;
        out->num_original_variables = n;
        out->objective_constant = in->objective_constant;
        out->objective_vector = ccalloc<double>((size_t)n_out);
        std::memcpy(out->objective_vector, in->objective_vector, (size_t)n * sizeof(double));
        out->variable_lower_bound = cmalloc<double>((size_t)n_out);
        out->variable_upper_bound = cmalloc<double>((size_t)n_out);
        std::memcpy(out->variable_lower_bound, in->variable_lower_bound, (size_t)n * sizeof(double));
        std::memcpy(out->variable_upper_bound, in->variable_upper_bound, (size_t)n * sizeof(double));
        for (long long j = n; j < n_out; ++j) {
            out->variable_lower_bound[j] = -INFINITY;
            out->variable_upper_bound[j] = INFINITY;
        }
        // cone blocks: (v_1 .. v_k, aux0, aux1) per constraint, in order
        out->cones.num_cones = K;
        out->cones.start_idx = cmalloc<int>(K);
        out->cones.v_dim = cmalloc<int>(K);
        out->cones.type = cmalloc<cone_type_t>(K);
        int pos = n;
        for (int i = 0; i < K; ++i) {
            blocks[i].start = pos;
            out->cones.start_idx[i] = pos;
            out->cones.v_dim[i] = (int)blocks[i].cols.size();
            out->cones.type[i] = form;
            pos += (int)blocks[i].cols.size() + 2;
        }
        // row bounds: original rows, then the v-definition rows (= 0), then
        // the closing rows (t = 1 or z - w = sqrt2; 0 when pinned)
        out->constraint_lower_bound = cmalloc<double>((size_t)m_out);
        out->constraint_upper_bound = cmalloc<double>((size_t)m_out);
        out->affine_cone_offset = ccalloc<double>((size_t)m_out);
        std::memcpy(out->constraint_lower_bound, in->constraint_lower_bound, (size_t)m * sizeof(double));
        std::memcpy(out->constraint_upper_bound, in->constraint_upper_bound, (size_t)m * sizeof(double));
        if (in->affine_cone_offset) std::memcpy(out->affine_cone_offset, in->affine_cone_offset, (size_t)m * sizeof(double));
        std::vector<int> block_of_row((size_t)m, -1);
        for (int i = 0; i < K; ++i) {
            const int row = in->quadratic_constraint_row_indices[i];
            block_of_row[row] = i;
            const double rhs = blocks[i].pinned ? 0.0 : (standard ? blocks[i].b * kRoot2 : blocks[i].b);
            out->constraint_lower_bound[row] = out->constraint_upper_bound[row] = rhs;
        }
        long long r = m;
        for (; r < m + lifted; ++r) out->constraint_lower_bound[r] = out->constraint_upper_bound[r] = 0.0;
        for (int i = 0; i < K; ++i, ++r) {
            const double v = blocks[i].pinned ? 0.0 : (standard ? kRoot2 : 1.0);
            out->constraint_lower_bound[r] = out->constraint_upper_bound[r] = v;
        }
        // [A; v rows; closing rows] as CSR
        const CsrComponent *A = in->constraint_matrix;
        CsrComponent *M = ccalloc<CsrComponent>(1);
        out->constraint_matrix = M;
        out->constraint_matrix_num_nonzeros = (int)nnz_out;
        M->row_ptr = ccalloc<int>((size_t)m_out + 1);
        M->col_ind = cmalloc<int>((size_t)std::max<long long>(nnz_out, 1));
        M->val = cmalloc<double>((size_t)std::max<long long>(nnz_out, 1));
        int e = 0;
        for (int row = 0; row < m; ++row) {
            const int i = block_of_row[row];
            double s = 1.0;
            if (i >= 0) {
                if (blocks[i].flip) s = -1.0;
                if (standard) s *= kRoot2;
            }
            for (int k = A->row_ptr[row]; k < A->row_ptr[row + 1]; ++k, ++e) {
                M->col_ind[e] = A->col_ind[k];
                M->val[e] = s * A->val[k];
            }
            if (i >= 0 && !blocks[i].pinned) {
                const int aux = blocks[i].start + (int)blocks[i].cols.size();
                M->col_ind[e] = aux;
                M->val[e++] = 1.0;
                if (standard) {
                    M->col_ind[e] = aux + 1;
                    M->val[e++] = 1.0;
                }
            }
            M->row_ptr[row + 1] = e;
        }
        int row = m;
        for (const QcBlock &b : blocks)
            for (size_t j = 0; j < b.cols.size(); ++j) {
                M->col_ind[e] = b.cols[j];
                M->val[e++] = -b.scale[j];
                M->col_ind[e] = b.start + (int)j;
                M->val[e++] = 1.0;
                M->row_ptr[++row] = e;
            }
        for (const QcBlock &b : blocks) {
            if (!b.pinned) {
                const int aux = b.start + (int)b.cols.size();
                if (standard) {
                    M->col_ind[e] = aux;
                    M->val[e++] = -1.0;
                }
                M->col_ind[e] = aux + 1;
                M->val[e++] = 1.0;
            }
            M->row_ptr[++row] = e;
        }
        if (e != nnz_out) fail("lifted matrix holds %d entries, expected %lld", e, nnz_out);
        // objective terms unchanged (the sparse Q padded with empty rows)
        out->num_rank_lowrank_obj = in->num_rank_lowrank_obj;
        out->objective_sparse_matrix_num_nonzeros = in->objective_sparse_matrix_num_nonzeros;
        out->objective_lowrank_matrix_num_nonzeros = in->objective_lowrank_matrix_num_nonzeros;
        out->objective_lowrank_middle_matrix_num_nonzeros = in->objective_lowrank_middle_matrix_num_nonzeros;
        if (in->objective_sparse_matrix)
            out->objective_sparse_matrix =
                copy_csr(in->objective_sparse_matrix, n, (int)n_out, in->objective_sparse_matrix_num_nonzeros);
        if (in->objective_lowrank_matrix)
            out->objective_lowrank_matrix = copy_csr(in->objective_lowrank_matrix, in->num_rank_lowrank_obj,
                                                     in->num_rank_lowrank_obj, in->objective_lowrank_matrix_num_nonzeros);
        if (in->objective_lowrank_middle_matrix)
            out->objective_lowrank_middle_matrix =
                copy_csr(in->objective_lowrank_middle_matrix, in->num_rank_lowrank_obj, in->num_rank_lowrank_obj,
                         in->objective_lowrank_middle_matrix_num_nonzeros);
        // pinned aux slots
        if (unpinned < K) {
            out->cones.fixed_mask_size = (int)n_out;
            out->cones.is_fixed = ccalloc<char>((size_t)n_out);
            out->primal_start = ccalloc<double>((size_t)n_out);
            for (const QcBlock &b : blocks) {
                if (!b.pinned) continue;
                const int s = b.start + (int)b.cols.size();
                out->cones.is_fixed[s] = out->cones.is_fixed[s + 1] = 1;
                out->primal_start[s] = standard ? (b.b - 1.0) * kRoot2 * 0.5 : b.b;
                out->primal_start[s + 1] = standard ? (b.b + 1.0) * kRoot2 * 0.5 : 1.0;
            }
        }
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "[qcqp_to_socp_qp] %s.\n", ex.what());
        qp_problem_free(out);
        return nullptr;
    }
    return out;
}

// Truncates a result of the lifted problem to the original's variables and
// rows (reference src/qcqp_transform.c:20-57).
void restore_result_dimensions(pdhcg_result_t *r, const qp_problem_t *orig) {
    if (!r || !orig) return;
    const int n = orig->num_variables, m = orig->num_constraints;
    auto shrink = [](double *&a, int keep) {
        double *b = keep > 0 ? cmalloc<double>(keep) : nullptr;
        if (keep > 0) std::memcpy(b, a, (size_t)keep * sizeof(double));
        std::free(a);
        a = b;
    };
    if (n >= 0 && n < r->num_variables) {
        if (r->primal_solution) shrink(r->primal_solution, n);
        if (r->reduced_cost) shrink(r->reduced_cost, n);
        r->num_variables = n;
    }
    if (m >= 0 && m < r->num_constraints) {
        if (r->dual_solution) shrink(r->dual_solution, m);
        r->num_constraints = m;
    }
    r->num_nonzeros = orig->constraint_matrix_num_nonzeros;
}

}  // namespace pdhcg

extern "C" qp_problem_t *qcqp_to_socp_qp(const qp_problem_t *orig, cone_type_t default_type) {
    if (!orig) return nullptr;
    try {
        return pdhcg::lift_quadratic_constraints(orig, default_type);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[qcqp_to_socp_qp] %s.\n", e.what());
        return nullptr;
    }
}
