// Problem assembly for libpdhcg_amd.so: descriptor validation, conversion of
// dense / CSC / COO / CSR input to one CSR per matrix, stacking of the affine
// rows F under A, cone block bookkeeping, warm starts and fixed cone slots.
//
// Interface and observable behaviour follow reference src/pdhcg.c:35-666 and
// src/utils.cu:580-804: same validation rules, same zero-tolerance rules,
// entries of each output row in input order, same error returns.  The code is
// a counting-sort design of this project: every input format is reduced to a
// stream of (row, col, value) triples visited in input order and scattered
// into row buckets.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_common.h"

namespace pdhcg {

int cone_slots(cone_type_t type, int v_dim) {
    switch (type) {
    case CONE_EXPONENTIAL:
    case CONE_POWER:
        return 3;
    case CONE_STANDARD_SOC:
    case CONE_ROTATED_SOC:
        return (v_dim >= 0 && v_dim <= INT_MAX - 2) ? v_dim + 2 : -1;
    case CONE_PSD: {
        if (v_dim <= 0) return -1;
        long long len = (long long)v_dim * (v_dim + 1LL) / 2LL;
        return len <= INT_MAX ? (int)len : -1;
    }
    default:
        return -1;
    }
}

int cone_block_slots(const cone_blocks_t *b, int k) {
    if (!b || k < 0 || k >= b->num_cones || !b->type || !b->v_dim) return -1;
    return cone_slots(b->type[k], b->v_dim[k]);
}

namespace {

void release_blocks(cone_blocks_t *b) {
    std::free(b->start_idx);
    std::free(b->v_dim);
    std::free(b->type);
    std::free(b->power_alpha);
    std::free(b->is_fixed);
    std::memset(b, 0, sizeof(*b));
}

void release_csr(CsrComponent *c) {
    if (!c) return;
    std::free(c->row_ptr);
    std::free(c->col_ind);
    std::free(c->val);
    std::memset(c, 0, sizeof(*c));
}

// Shape/content checks of one descriptor (reference src/pdhcg.c:35-168).
void check_descriptor(const matrix_desc_t *d, const char *name) {
    if (!d) return;
    if (d->m < 0 || d->n < 0) fail("%s matrix has negative shape (%d, %d)", name, d->m, d->n);
    switch (d->fmt) {
    case matrix_dense:
        if (d->m > 0 && d->n > INT_MAX / d->m) fail("%s dense matrix is too large to index with int", name);
        if (d->m > 0 && d->n > 0 && !d->data.dense.A) fail("%s dense matrix data is NULL", name);
        return;
    case matrix_csr: {
        const auto &c = d->data.csr;
        if (c.nnz < 0 || !c.row_ptr || (c.nnz > 0 && (!c.col_ind || !c.vals))) fail("%s CSR storage is incomplete", name);
        if (c.row_ptr[0] != 0 || c.row_ptr[d->m] != c.nnz) fail("%s CSR row pointers do not span [0, nnz]", name);
        for (int r = 0; r < d->m; ++r)
            if (c.row_ptr[r] > c.row_ptr[r + 1] || c.row_ptr[r] < 0 || c.row_ptr[r + 1] > c.nnz)
                fail("%s CSR row pointers are invalid at row %d", name, r);
        for (int e = 0; e < c.nnz; ++e)
            if (c.col_ind[e] < 0 || c.col_ind[e] >= d->n)
                fail("%s CSR column index %d is out of range [0, %d)", name, c.col_ind[e], d->n);
        return;
    }
    case matrix_csc: {
        const auto &c = d->data.csc;
        if (c.nnz < 0 || !c.col_ptr || (c.nnz > 0 && (!c.row_ind || !c.vals))) fail("%s CSC storage is incomplete", name);
        if (c.col_ptr[0] != 0 || c.col_ptr[d->n] != c.nnz) fail("%s CSC column pointers do not span [0, nnz]", name);
        for (int j = 0; j < d->n; ++j)
            if (c.col_ptr[j] > c.col_ptr[j + 1] || c.col_ptr[j] < 0 || c.col_ptr[j + 1] > c.nnz)
                fail("%s CSC column pointers are invalid at column %d", name, j);
        for (int e = 0; e < c.nnz; ++e)
            if (c.row_ind[e] < 0 || c.row_ind[e] >= d->m)
                fail("%s CSC row index %d is out of range [0, %d)", name, c.row_ind[e], d->m);
        return;
    }
    case matrix_coo: {
        const auto &c = d->data.coo;
        if (c.nnz < 0 || (c.nnz > 0 && (!c.row_ind || !c.col_ind || !c.vals))) fail("%s COO storage is incomplete", name);
        for (int e = 0; e < c.nnz; ++e)
            if (c.row_ind[e] < 0 || c.row_ind[e] >= d->m || c.col_ind[e] < 0 || c.col_ind[e] >= d->n)
                fail("%s COO index (%d, %d) is out of range for shape (%d, %d)", name, c.row_ind[e], c.col_ind[e],
                     d->m, d->n);
        return;
    }
    default:
        fail("%s matrix has unsupported format %d", name, (int)d->fmt);
    }
}

// Visits the stored entries of a descriptor in the order the CSR output must
// list them within each row, after the format's zero filter:
//   dense  : row-major, keeps |v| > tol, tol = zero_tolerance or 1e-12
//   CSC    : column-major, drops |v| <= tol only when zero_tolerance > 0
//   COO    : input order, same filter as CSC
template <class F> void for_each_entry(const matrix_desc_t *d, F &&visit) {
    if (d->fmt == matrix_dense) {
        const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 1e-12;
        const double *a = d->data.dense.A;
        for (int i = 0; i < d->m; ++i)
            for (int j = 0; j < d->n; ++j) {
                double v = a[(size_t)i * d->n + j];
                if (std::fabs(v) > tol) visit(i, j, v);
            }
        return;
    }
    const double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
    auto keep = [tol](double v) { return !(tol > 0 && std::fabs(v) <= tol); };
    if (d->fmt == matrix_csc) {
        const auto &c = d->data.csc;
        for (int j = 0; j < d->n; ++j)
            for (int e = c.col_ptr[j]; e < c.col_ptr[j + 1]; ++e)
                if (keep(c.vals[e])) visit(c.row_ind[e], j, c.vals[e]);
        return;
    }
    const auto &c = d->data.coo;
    for (int e = 0; e < c.nnz; ++e)
        if (keep(c.vals[e])) visit(c.row_ind[e], c.col_ind[e], c.vals[e]);
}

// Converts one descriptor to CSR; CSR input is copied verbatim
// (reference src/pdhcg.c:232-272, src/utils.cu:591-804).
void descriptor_to_csr(const matrix_desc_t *d, CsrComponent *out, int *nnz_out) {
    std::memset(out, 0, sizeof(*out));
    const int m = d->m;
    if (d->fmt == matrix_csr) {
        const int nnz = d->data.csr.nnz;
        out->row_ptr = cmalloc<int>((size_t)m + 1);
        std::memcpy(out->row_ptr, d->data.csr.row_ptr, ((size_t)m + 1) * sizeof(int));
        if (nnz > 0) {
            out->col_ind = cmalloc<int>(nnz);
            out->val = cmalloc<double>(nnz);
            std::memcpy(out->col_ind, d->data.csr.col_ind, (size_t)nnz * sizeof(int));
            std::memcpy(out->val, d->data.csr.vals, (size_t)nnz * sizeof(double));
        }
        *nnz_out = nnz;
        return;
    }
    // pass 1: bucket sizes; pass 2: scatter into buckets in visiting order
    std::vector<long long> start((size_t)m + 1, 0);
    for_each_entry(d, [&](int r, int, double) { start[(size_t)r + 1]++; });
    for (int r = 0; r < m; ++r) start[(size_t)r + 1] += start[r];
    const long long total = start[m];
    if (total > INT_MAX) fail("matrix has more than INT_MAX nonzeros");
    out->row_ptr = cmalloc<int>((size_t)m + 1);
    for (int r = 0; r <= m; ++r) out->row_ptr[r] = (int)start[r];
    out->col_ind = cmalloc<int>((size_t)total);
    out->val = cmalloc<double>((size_t)total);
    std::vector<long long> cursor(start.begin(), start.end() - 1);
    for_each_entry(d, [&](int r, int c, double v) {
        long long pos = cursor[r]++;
        out->col_ind[pos] = c;
        out->val[pos] = v;
    });
    *nnz_out = (int)total;
}

// cone_spec_t list -> cone_blocks_t (reference src/cone_utils.c:47-167).
void blocks_from_specs(cone_blocks_t *b, int count, const cone_spec_t *specs, int ambient, bool allow_fixed,
                       const char *label) {
    if (count < 0 || ambient < 0 || (count > 0 && !specs))
        fail("invalid %s cone metadata (num_cones=%d, ambient_dimension=%d)", label, count, ambient);
    release_blocks(b);
    if (count == 0) return;
    std::vector<char> taken((size_t)ambient, 0);
    bool any_fixed = false, any_power = false;
    for (int k = 0; k < count; ++k) {
        const cone_spec_t &s = specs[k];
        const int len = cone_slots(s.type, s.v_dim);
        if (len <= 0 || s.start_idx < 0 || (long long)s.start_idx + len > ambient)
            fail("%s cone %d has invalid type, size, or range (start=%d, length=%d, ambient_dimension=%d)", label, k,
                 s.start_idx, len, ambient);
        if (!allow_fixed && s.is_fixed) fail("%s cone %d does not support fixed slots", label, k);
        if (s.type == CONE_PSD && s.is_fixed)
            for (int j = 0; j < len; ++j)
                if (s.is_fixed[j]) fail("%s PSD cone %d does not support fixed slots", label, k);
        if (s.type == CONE_POWER && !(std::isfinite(s.power_alpha) && s.power_alpha > 0.0 && s.power_alpha < 1.0))
            fail("%s power cone %d requires alpha in (0,1); got %.6g", label, k, s.power_alpha);
        for (int i = s.start_idx; i < s.start_idx + len; ++i) {
            if (taken[i]) fail("%s cone %d overlaps at index %d", label, k, i);
            taken[i] = 1;
        }
        any_fixed |= s.is_fixed != nullptr;
        any_power |= s.type == CONE_POWER;
    }
    b->num_cones = count;
    b->start_idx = cmalloc<int>(count);
    b->v_dim = cmalloc<int>(count);
    b->type = cmalloc<cone_type_t>(count);
    b->power_alpha = any_power ? ccalloc<double>(count) : nullptr;
    if (any_fixed) {
        b->fixed_mask_size = ambient;
        b->is_fixed = ccalloc<char>(ambient);
    }
    for (int k = 0; k < count; ++k) {
        const cone_spec_t &s = specs[k];
        b->start_idx[k] = s.start_idx;
        b->v_dim[k] = (s.type == CONE_EXPONENTIAL || s.type == CONE_POWER) ? 1 : s.v_dim;
        b->type[k] = s.type;
        if (b->power_alpha && s.type == CONE_POWER) b->power_alpha[k] = s.power_alpha;
        if (b->is_fixed && s.is_fixed) {
            const int len = cone_slots(s.type, s.v_dim);
            for (int j = 0; j < len; ++j) b->is_fixed[s.start_idx + j] = s.is_fixed[j] ? 1 : 0;
        }
    }
}

double *copy_or_fill(int n, const double *src, double fill) {
    double *d = cmalloc<double>((size_t)(n > 0 ? n : 0));
    for (int i = 0; i < n; ++i) d[i] = src ? src[i] : fill;
    return d;
}

void free_problem(qp_problem_t *p) {
    if (!p) return;
    CsrComponent *mats[] = {p->objective_sparse_matrix, p->objective_lowrank_matrix, p->constraint_matrix,
                            p->objective_lowrank_middle_matrix};
    for (CsrComponent *c : mats) {
        release_csr(c);
        std::free(c);
    }
    std::free(p->variable_lower_bound);
    std::free(p->variable_upper_bound);
    std::free(p->objective_vector);
    std::free(p->constraint_lower_bound);
    std::free(p->constraint_upper_bound);
    std::free(p->affine_cone_offset);
    std::free(p->primal_start);
    std::free(p->dual_start);
    if (p->quadratic_constraint_matrices) {
        for (int i = 0; i < p->num_quadratic_constraints; ++i) {
            release_csr(p->quadratic_constraint_matrices[i]);
            std::free(p->quadratic_constraint_matrices[i]);
        }
        std::free(p->quadratic_constraint_matrices);
    }
    std::free(p->quadratic_constraint_row_indices);
    std::free(p->quadratic_constraint_matrix_num_nonzeros);
    release_blocks(&p->cones);
    release_blocks(&p->affine_cones);
    std::free(p);
}

qp_problem_t *build_problem(const double *objective_c, const matrix_desc_t *Q, const matrix_desc_t *R,
                            const matrix_desc_t *D, const matrix_desc_t *A, const double *con_lb,
                            const double *con_ub, const double *var_lb, const double *var_ub,
                            const double *objective_constant, int num_var_cones, const cone_spec_t *var_cones,
                            const matrix_desc_t *F, const double *g, int num_affine_cones,
                            const cone_spec_t *affine_cones, qp_problem_t *p) {
    if (!F && (g || num_affine_cones != 0 || affine_cones)) fail("affine cone data requires affine_cone_matrix_desc");
    if (!A && !F && !Q && !R) fail("at least one of A, F, Q, or R must be provided");
    check_descriptor(A, "A");
    check_descriptor(F, "F");
    check_descriptor(Q, "Q");
    check_descriptor(R, "R");
    const int n = A ? A->n : (F ? F->n : (Q ? Q->n : R->n));
    const int m = A ? A->m : 0;
    const int q = F ? F->m : 0;
    if (F && F->n != n) fail("F matrix shape (%d, %d) must have %d columns", F->m, F->n, n);
    if (Q && (Q->m != n || Q->n != n)) fail("Q matrix shape (%d, %d) must be (%d, %d)", Q->m, Q->n, n, n);
    if (R && R->n != n) fail("R matrix shape (%d, %d) must have %d columns", R->m, R->n, n);
    if (D && R && R->m > 0) {
        check_descriptor(D, "D");
        if (D->m != R->m || D->n != R->m) fail("D matrix shape (%d, %d) must be (%d, %d)", D->m, D->n, R->m, R->m);
    }
    if (q > INT_MAX - m) fail("combined A and F matrices have too many rows");

    p->num_variables = n;
    p->num_constraints = m + q;
    p->affine_cone_offset = p->num_constraints > 0 ? ccalloc<double>(p->num_constraints) : nullptr;

    // [A; F] as one CSR
    p->constraint_matrix = ccalloc<CsrComponent>(1);
    if (A) {
        descriptor_to_csr(A, p->constraint_matrix, &p->constraint_matrix_num_nonzeros);
    } else {
        p->constraint_matrix->row_ptr = ccalloc<int>((size_t)m + 1);
        p->constraint_matrix_num_nonzeros = 0;
    }
    if (F) {
        CsrComponent f{};
        int fnnz = 0;
        descriptor_to_csr(F, &f, &fnnz);
        const int base = p->constraint_matrix_num_nonzeros;
        if (fnnz > INT_MAX - base) {
            release_csr(&f);
            fail("combined A and F matrices have too many nonzeros");
        }
        CsrComponent *c = p->constraint_matrix;
        int *rp = cmalloc<int>((size_t)m + q + 1);
        std::memcpy(rp, c->row_ptr, ((size_t)m + 1) * sizeof(int));
        for (int r = 1; r <= q; ++r) rp[m + r] = base + f.row_ptr[r];
        std::free(c->row_ptr);
        c->row_ptr = rp;
        if (fnnz > 0) {
            int *ci = cmalloc<int>((size_t)base + fnnz);
            double *v = cmalloc<double>((size_t)base + fnnz);
            if (base > 0) {
                std::memcpy(ci, c->col_ind, (size_t)base * sizeof(int));
                std::memcpy(v, c->val, (size_t)base * sizeof(double));
            }
            std::memcpy(ci + base, f.col_ind, (size_t)fnnz * sizeof(int));
            std::memcpy(v + base, f.val, (size_t)fnnz * sizeof(double));
            std::free(c->col_ind);
            std::free(c->val);
            c->col_ind = ci;
            c->val = v;
        }
        p->constraint_matrix_num_nonzeros = base + fnnz;
        release_csr(&f);
    }

    p->objective_sparse_matrix = ccalloc<CsrComponent>(1);
    if (Q) descriptor_to_csr(Q, p->objective_sparse_matrix, &p->objective_sparse_matrix_num_nonzeros);
    else p->objective_sparse_matrix->row_ptr = ccalloc<int>((size_t)n + 1);

    p->objective_lowrank_matrix = ccalloc<CsrComponent>(1);
    if (R) {
        p->num_rank_lowrank_obj = R->m;
        descriptor_to_csr(R, p->objective_lowrank_matrix, &p->objective_lowrank_matrix_num_nonzeros);
    } else {
        p->objective_lowrank_matrix->row_ptr = ccalloc<int>(1);
    }
    if (D) {
        if (p->num_rank_lowrank_obj <= 0) {
            std::fprintf(stderr, "[interface] D matrix ignored: problem has no low-rank component.\n");
        } else {
            p->objective_lowrank_middle_matrix = ccalloc<CsrComponent>(1);
            descriptor_to_csr(D, p->objective_lowrank_middle_matrix, &p->objective_lowrank_middle_matrix_num_nonzeros);
        }
    }

    p->objective_constant = objective_constant ? *objective_constant : 0.0;
    p->objective_vector = copy_or_fill(n, objective_c, 0.0);
    p->variable_lower_bound = copy_or_fill(n, var_lb, -INFINITY);
    p->variable_upper_bound = copy_or_fill(n, var_ub, INFINITY);
    p->constraint_lower_bound = copy_or_fill(p->num_constraints, nullptr, -INFINITY);
    p->constraint_upper_bound = copy_or_fill(p->num_constraints, nullptr, INFINITY);
    if (m > 0 && con_lb) std::memcpy(p->constraint_lower_bound, con_lb, (size_t)m * sizeof(double));
    if (m > 0 && con_ub) std::memcpy(p->constraint_upper_bound, con_ub, (size_t)m * sizeof(double));

    // affine cones index rows of F; stored with global row indices of [A; F]
    blocks_from_specs(&p->affine_cones, num_affine_cones, affine_cones, q, false, "affine");
    long long covered = 0;
    for (int k = 0; k < num_affine_cones; ++k) {
        covered += cone_block_slots(&p->affine_cones, k);
        p->affine_cones.start_idx[k] += m;
    }
    if (covered != q) fail("affine cone blocks cover %lld of %d rows of F", covered, q);
    if (g && q > 0) std::memcpy(p->affine_cone_offset + m, g, (size_t)q * sizeof(double));

    blocks_from_specs(&p->cones, num_var_cones, var_cones, n, true, "variable");
    // cone slots must be free (reference src/pdhcg.c:468-499)
    int first_cone_slot = n;
    for (int k = 0; k < p->cones.num_cones; ++k) {
        const int s = p->cones.start_idx[k], len = cone_block_slots(&p->cones, k);
        for (int v = s; v < s + len; ++v) {
            const double lo = p->variable_lower_bound[v], hi = p->variable_upper_bound[v];
            if ((std::isfinite(lo) && lo > -1e30) || (std::isfinite(hi) && hi < 1e30))
                fail("cone %d slot %d has a finite box bound (lb=%.6g, ub=%.6g); cone variables must be free", k, v,
                     lo, hi);
        }
        if (s < first_cone_slot) first_cone_slot = s;
    }
    if (p->cones.num_cones > 0) p->num_original_variables = first_cone_slot;
    return p;
}

double start_value(const qp_problem_t *p, int i) { return p->primal_start ? p->primal_start[i] : 0.0; }

double fixed_slot_norm(const qp_problem_t *p, int start, int len) {
    double r = 0.0;
    for (int j = 0; j < len; ++j)
        if (p->cones.is_fixed[start + j]) r = std::hypot(r, start_value(p, start + j));
    return r;
}

bool exp_section_nonempty(const qp_problem_t *p, int s) {
    const bool fx = p->cones.is_fixed[s], fy = p->cones.is_fixed[s + 1], fz = p->cones.is_fixed[s + 2];
    const double x = start_value(p, s), y = start_value(p, s + 1), z = start_value(p, s + 2);
    auto near_le = [](double a, double b) { return a <= b + 64.0 * DBL_EPSILON * (1.0 + std::fabs(a) + std::fabs(b)); };
    if ((fx && !std::isfinite(x)) || (fy && !std::isfinite(y)) || (fz && !std::isfinite(z))) return false;
    if (fy) {
        if (y < 0.0) return false;
        if (y == 0.0) return (!fx || x <= 0.0) && (!fz || z >= 0.0);
        if (fz && !(z > 0.0)) return false;
        if (fx && fz) return near_le(std::log(y) + x / y, std::log(z));
        return true;
    }
    if (!fz) return true;
    if (z < 0.0) return false;
    if (z == 0.0) return !fx || x <= 0.0;
    if (!fx || x <= 0.0) return true;
    return near_le(1.0 + std::log(x), std::log(z));  // min_{y>0} y exp(x/y) = e x
}

}  // namespace

int validate_fixed_sections(const qp_problem_t *p) {
    if (!p || !p->cones.is_fixed) return 0;
    if (p->cones.fixed_mask_size != p->num_variables) {
        std::fprintf(stderr, "[solve_qp_problem] variable cone fixed mask has size %d; expected %d.\n",
                     p->cones.fixed_mask_size, p->num_variables);
        return -1;
    }
    for (int k = 0; k < p->cones.num_cones; ++k) {
        const int s = p->cones.start_idx[k], vd = p->cones.v_dim[k], len = cone_block_slots(&p->cones, k);
        bool any = false;
        for (int j = 0; j < len; ++j) any |= p->cones.is_fixed[s + j] != 0;
        if (!any) continue;
        const cone_type_t t = p->cones.type[k];
        if (t == CONE_EXPONENTIAL) {
            if (!exp_section_nonempty(p, s)) {
                std::fprintf(stderr, "[solve_qp_problem] exponential cone %d has an empty or non-finite fixed section.\n", k);
                return -1;
            }
            continue;
        }
        if (t == CONE_STANDARD_SOC || t == CONE_ROTATED_SOC) {
            for (int j = 0; j < vd + 2; ++j)
                if (p->cones.is_fixed[s + j] && !std::isfinite(start_value(p, s + j))) {
                    std::fprintf(stderr, "[solve_qp_problem] second-order cone %d has a non-finite fixed value.\n", k);
                    return -1;
                }
        }
        if (t == CONE_STANDARD_SOC) {
            const int zi = s + vd + 1;
            const double z = start_value(p, zi);
            if (p->cones.is_fixed[zi] && (!(z >= 0.0) || fixed_slot_norm(p, s, vd + 1) > z)) {
                std::fprintf(stderr, "[solve_qp_problem] standard SOC %d has an empty fixed section.\n", k);
                return -1;
            }
            continue;
        }
        if (t == CONE_ROTATED_SOC) {
            const int si = s + vd, ti = si + 1;
            const bool fs = p->cones.is_fixed[si], ft = p->cones.is_fixed[ti];
            const double sv = start_value(p, si), tv = start_value(p, ti), fn = fixed_slot_norm(p, s, vd);
            bool empty = (fs && sv < 0.0) || (ft && tv < 0.0);
            if (!empty && fs && ft) empty = fn > 1.41421356237309504880 * std::sqrt(sv) * std::sqrt(tv);
            else if (fs && sv == 0.0) empty |= fn > 0.0;
            else if (ft && tv == 0.0) empty |= fn > 0.0;
            if (empty) {
                std::fprintf(stderr, "[solve_qp_problem] rotated SOC %d has an empty fixed section.\n", k);
                return -1;
            }
            continue;
        }
        if (t != CONE_POWER) continue;
        const bool fx = p->cones.is_fixed[s], fy = p->cones.is_fixed[s + 1], fz = p->cones.is_fixed[s + 2];
        const double x = start_value(p, s), y = start_value(p, s + 1), z = start_value(p, s + 2);
        if ((fx && (!std::isfinite(x) || x < 0.0)) || (fy && (!std::isfinite(y) || y < 0.0)) ||
            (fz && !std::isfinite(z))) {
            std::fprintf(stderr, "[solve_qp_problem] power cone %d has an invalid fixed value.\n", k);
            return -1;
        }
        if (fz && z != 0.0 && ((fx && x == 0.0) || (fy && y == 0.0))) {
            std::fprintf(stderr, "[solve_qp_problem] power cone %d has an empty fixed section.\n", k);
            return -1;
        }
        if (fx && fy && fz && z != 0.0) {
            const double a = p->cones.power_alpha[k];
            const double lb = (x > 0.0 && y > 0.0) ? a * std::log(x) + (1.0 - a) * std::log(y) : -INFINITY;
            const double lz = std::log(std::fabs(z));
            if (lb + 64.0 * DBL_EPSILON * (1.0 + std::fabs(lb) + std::fabs(lz)) < lz) {
                std::fprintf(stderr, "[solve_qp_problem] power cone %d has an infeasible fully fixed point.\n", k);
                return -1;
            }
        }
    }
    return 0;
}

void fill_default_parameters(pdhg_parameters_t *p) {
    std::memset(p, 0, sizeof(*p));
    p->curtis_reid_iterations = 0;
    p->l_inf_ruiz_iterations = 10;
    p->has_pock_chambolle_alpha = true;
    p->pock_chambolle_alpha = 1.0;
    p->bound_objective_rescaling = true;
    p->use_cone_preserving_scaling = true;
    p->verbose = 1;
    p->termination_evaluation_frequency = 200;
    p->feasibility_polishing = false;
    p->reflection_coefficient = 1.0;
    p->presolve = false;
    p->sv_max_iter = 5000;
    p->sv_tol = 1e-4;
    p->termination_criteria.eps_optimal_relative = 1e-4;
    p->termination_criteria.eps_feasible_relative = 1e-4;
    p->termination_criteria.eps_infeasible = 1e-12;
    p->termination_criteria.time_sec_limit = 3600.0;
    p->termination_criteria.iteration_limit = INT_MAX;
    p->termination_criteria.eps_feas_polish_relative = 1e-6;
    p->restart_params.artificial_restart_threshold = 0.36;
    p->restart_params.sufficient_reduction_for_restart = 0.2;
    p->restart_params.necessary_reduction_for_restart = 0.8;
    p->restart_params.k_p = 0.99;
    p->restart_params.k_i = 0.01;
    p->restart_params.k_d = 0.0;
    p->restart_params.i_smooth = 0.3;
    p->optimality_norm = NORM_TYPE_L_INF;
    p->inner_solver_parameters.iteration_limit = 1000;
    p->inner_solver_parameters.initial_tolerance = 1e-3;
    p->inner_solver_parameters.min_tolerance = 1e-9;
    p->grid_size.decided = false;
    p->partition_method = NNZ_BALANCE_PARTITION;
    p->permute_method = BLOCK_RANDOM_PERMUTATION;
    p->diag_jacobi_precond = true;
    p->default_cone_type = CONE_ROTATED_SOC;
}

}  // namespace pdhcg

using namespace pdhcg;

extern "C" {

qp_problem_t *create_qp_problem(const double *objective_c, const matrix_desc_t *Q_desc, const matrix_desc_t *R_desc,
                                const matrix_desc_t *D_desc, const matrix_desc_t *A_desc, const double *con_lb,
                                const double *con_ub, const double *var_lb, const double *var_ub,
                                const double *objective_constant, int num_var_cones, const cone_spec_t *var_cones,
                                const matrix_desc_t *affine_cone_matrix_desc, const double *affine_cone_offset,
                                int num_affine_cones, const cone_spec_t *affine_cones) {
    qp_problem_t *p = nullptr;
    try {
        p = ccalloc<qp_problem_t>(1);
        return build_problem(objective_c, Q_desc, R_desc, D_desc, A_desc, con_lb, con_ub, var_lb, var_ub,
                             objective_constant, num_var_cones, var_cones, affine_cone_matrix_desc,
                             affine_cone_offset, num_affine_cones, affine_cones, p);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[create_qp_problem] %s.\n", e.what());
        free_problem(p);
        return nullptr;
    }
}

void qp_problem_free(qp_problem_t *prob) { free_problem(prob); }

void pdhcg_result_free(pdhcg_result_t *r) {
    if (!r) return;
    std::free(r->primal_solution);
    std::free(r->dual_solution);
    std::free(r->reduced_cost);
    std::free(r);
}

void set_default_parameters(pdhg_parameters_t *params) {
    if (params) fill_default_parameters(params);
}

void set_start_values(qp_problem_t *p, const double *primal, const double *dual) {
    if (!p) return;
    const int n = p->num_variables, m = p->num_constraints;
    try {
        if (primal && p->cones.is_fixed && p->primal_start)
            for (int i = 0; i < n; ++i)
                if (p->cones.is_fixed[i] && primal[i] != p->primal_start[i]) {
                    std::fprintf(stderr,
                                 "[set_start_values] slot %d is fixed at %.17g but caller provided %.17g; "
                                 "rejecting (use set_cone_fixed to change fixed value).\n",
                                 i, p->primal_start[i], primal[i]);
                    return;
                }
        double *ps = nullptr, *ds = nullptr;
        if (primal) {
            ps = cmalloc<double>(n);
            std::memcpy(ps, primal, (size_t)n * sizeof(double));
        } else if (p->cones.is_fixed && p->primal_start) {
            ps = ccalloc<double>(n);
            for (int i = 0; i < n; ++i)
                if (p->cones.is_fixed[i]) ps[i] = p->primal_start[i];
        }
        if (dual) {
            ds = cmalloc<double>(m);
            std::memcpy(ds, dual, (size_t)m * sizeof(double));
        }
        std::free(p->primal_start);
        std::free(p->dual_start);
        p->primal_start = ps;
        p->dual_start = ds;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[set_start_values] %s.\n", e.what());
    }
}

int set_cone_fixed(qp_problem_t *p, int cone_idx, int slot, double value) {
    if (!p) {
        std::fprintf(stderr, "[set_cone_fixed] prob is NULL\n");
        return -1;
    }
    if (cone_idx < 0 || cone_idx >= p->cones.num_cones) {
        std::fprintf(stderr, "[set_cone_fixed] cone_idx %d out of range [0, %d)\n", cone_idx, p->cones.num_cones);
        return -1;
    }
    if (p->cones.type[cone_idx] == CONE_PSD) {
        std::fprintf(stderr, "[set_cone_fixed] fixed PSD cone sections are not supported.\n");
        return -1;
    }
    const int len = cone_block_slots(&p->cones, cone_idx);
    if (slot < 0 || slot >= len) {
        std::fprintf(stderr, "[set_cone_fixed] slot %d out of range [0, %d) for cone %d\n", slot, len, cone_idx);
        return -1;
    }
    const int idx = p->cones.start_idx[cone_idx] + slot;
    if (idx < 0 || idx >= p->num_variables) {
        std::fprintf(stderr, "[set_cone_fixed] computed index %d out of range [0, %d)\n", idx, p->num_variables);
        return -1;
    }
    if (!std::isfinite(value)) {
        std::fprintf(stderr, "[set_cone_fixed] fixed value must be finite; got %.17g\n", value);
        return -1;
    }
    try {
        if (!p->cones.is_fixed) {
            p->cones.is_fixed = ccalloc<char>(p->num_variables);
            p->cones.fixed_mask_size = p->num_variables;
        }
        p->cones.is_fixed[idx] = 1;
        if (!p->primal_start) p->primal_start = ccalloc<double>(p->num_variables);
        p->primal_start[idx] = value;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "[set_cone_fixed] %s.\n", e.what());
        return -1;
    }
    return 0;
}

pdhcg_result_t *solve_qp_problem_distributed(const pdhg_parameters_t *, const qp_problem_t *) {
    std::fprintf(stderr,
                 "Error: distributed solve is not available in this build (single-GPU library); "
                 "use solve_qp_problem.\n");
    return nullptr;
}

}  // extern "C"
