/*
 * ORACLE (test infrastructure, see oracle.h): problem assembly, validation,
 * warm start, fixed cone slots and default parameters, restated from the
 * reference in plain C.
 */
#include "oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static void *xcalloc(size_t k, size_t n) {
    void *p = calloc(k ? k : 1, n ? n : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* reference src/cone_utils.c:26-38 */
static int cone_len(cone_type_t type, int v_dim) {
    if (type == CONE_EXPONENTIAL || type == CONE_POWER) return 3;
    if (type == CONE_STANDARD_SOC || type == CONE_ROTATED_SOC)
        return (v_dim >= 0 && v_dim <= INT_MAX - 2) ? v_dim + 2 : -1;
    if (type == CONE_PSD && v_dim > 0) {
        long long L = (long long)v_dim * (v_dim + 1LL) / 2LL;
        return L <= INT_MAX ? (int)L : -1;
    }
    return -1;
}
int orc_cone_block_length(const cone_blocks_t *b, int k) {
    if (!b || k < 0 || k >= b->num_cones || !b->type || !b->v_dim) return -1;
    return cone_len(b->type[k], b->v_dim[k]);
}

static void cone_blocks_release(cone_blocks_t *b) {
    free(b->start_idx); free(b->v_dim); free(b->type); free(b->power_alpha); free(b->is_fixed);
    memset(b, 0, sizeof(*b));
}

/* reference src/cone_utils.c:47-167 */
static int cone_blocks_from_specs(cone_blocks_t *b, int nc, const cone_spec_t *specs, int ambient, int allow_fixed,
                                  const char *label) {
    if (nc < 0 || ambient < 0 || (nc > 0 && !specs)) {
        fprintf(stderr, "[create_qp_problem] Invalid %s cone metadata.\n", label);
        return -1;
    }
    cone_blocks_release(b);
    if (nc == 0) return 0;
    char *owner = ambient > 0 ? (char *)xcalloc((size_t)ambient, 1) : NULL;
    int any_fixed = 0, any_power = 0;
    for (int k = 0; k < nc; ++k) {
        int L = cone_len(specs[k].type, specs[k].v_dim);
        int s = specs[k].start_idx;
        if (L <= 0 || s < 0 || (long long)s + L > ambient) {
            fprintf(stderr, "[create_qp_problem] %s cone %d has invalid type, size, or range.\n", label, k);
            free(owner);
            return -1;
        }
        if (!allow_fixed && specs[k].is_fixed) {
            fprintf(stderr, "[create_qp_problem] %s cone %d does not support fixed slots.\n", label, k);
            free(owner);
            return -1;
        }
        if (specs[k].type == CONE_PSD && specs[k].is_fixed) {
            int any = 0;
            for (int j = 0; j < L; ++j) any |= specs[k].is_fixed[j] != 0;
            if (any) {
                fprintf(stderr, "[create_qp_problem] %s PSD cone %d does not support fixed slots.\n", label, k);
                free(owner);
                return -1;
            }
        }
        if (specs[k].type == CONE_POWER &&
            !(isfinite(specs[k].power_alpha) && specs[k].power_alpha > 0.0 && specs[k].power_alpha < 1.0)) {
            fprintf(stderr, "[create_qp_problem] %s power cone %d requires alpha in (0,1).\n", label, k);
            free(owner);
            return -1;
        }
        for (int i = s; i < s + L; ++i) {
            if (owner[i]) {
                fprintf(stderr, "[create_qp_problem] %s cone %d overlaps at index %d.\n", label, k, i);
                free(owner);
                return -1;
            }
            owner[i] = 1;
        }
        any_fixed |= specs[k].is_fixed != NULL;
        any_power |= specs[k].type == CONE_POWER;
    }
    free(owner);
    b->num_cones = nc;
    b->start_idx = (int *)xmalloc((size_t)nc * sizeof(int));
    b->v_dim = (int *)xmalloc((size_t)nc * sizeof(int));
    b->type = (cone_type_t *)xmalloc((size_t)nc * sizeof(cone_type_t));
    b->power_alpha = any_power ? (double *)xcalloc((size_t)nc, sizeof(double)) : NULL;
    if (any_fixed) {
        b->fixed_mask_size = ambient;
        b->is_fixed = (char *)xcalloc((size_t)ambient, 1);
    }
    for (int k = 0; k < nc; ++k) {
        int L = cone_len(specs[k].type, specs[k].v_dim);
        b->start_idx[k] = specs[k].start_idx;
        b->v_dim[k] = (specs[k].type == CONE_EXPONENTIAL || specs[k].type == CONE_POWER) ? 1 : specs[k].v_dim;
        b->type[k] = specs[k].type;
        if (b->power_alpha && specs[k].type == CONE_POWER) b->power_alpha[k] = specs[k].power_alpha;
        if (b->is_fixed && specs[k].is_fixed)
            for (int j = 0; j < L; ++j) b->is_fixed[specs[k].start_idx + j] = specs[k].is_fixed[j] ? 1 : 0;
    }
    return 0;
}

/* reference src/pdhcg.c:35-168 */
static int check_desc(const matrix_desc_t *d, const char *name) {
    if (!d) return 0;
    if (d->m < 0 || d->n < 0) {
        fprintf(stderr, "[create_qp_problem] %s matrix has negative shape.\n", name);
        return -1;
    }
    switch (d->fmt) {
    case matrix_dense:
        if (d->m > 0 && d->n > INT_MAX / d->m) return -1;
        if (d->m > 0 && d->n > 0 && !d->data.dense.A) return -1;
        return 0;
    case matrix_csr: {
        int nnz = d->data.csr.nnz;
        const int *rp = d->data.csr.row_ptr;
        if (nnz < 0 || !rp || (nnz > 0 && (!d->data.csr.col_ind || !d->data.csr.vals))) return -1;
        if (rp[0] != 0 || rp[d->m] != nnz) return -1;
        for (int r = 0; r < d->m; ++r)
            if (rp[r] > rp[r + 1] || rp[r] < 0 || rp[r + 1] > nnz) return -1;
        for (int e = 0; e < nnz; ++e)
            if (d->data.csr.col_ind[e] < 0 || d->data.csr.col_ind[e] >= d->n) return -1;
        return 0;
    }
    case matrix_csc: {
        int nnz = d->data.csc.nnz;
        const int *cp = d->data.csc.col_ptr;
        if (nnz < 0 || !cp || (nnz > 0 && (!d->data.csc.row_ind || !d->data.csc.vals))) return -1;
        if (cp[0] != 0 || cp[d->n] != nnz) return -1;
        for (int c = 0; c < d->n; ++c)
            if (cp[c] > cp[c + 1] || cp[c] < 0 || cp[c + 1] > nnz) return -1;
        for (int e = 0; e < nnz; ++e)
            if (d->data.csc.row_ind[e] < 0 || d->data.csc.row_ind[e] >= d->m) return -1;
        return 0;
    }
    case matrix_coo: {
        int nnz = d->data.coo.nnz;
        if (nnz < 0 || (nnz > 0 && (!d->data.coo.row_ind || !d->data.coo.col_ind || !d->data.coo.vals))) return -1;
        for (int e = 0; e < nnz; ++e) {
            int r = d->data.coo.row_ind[e], c = d->data.coo.col_ind[e];
            if (r < 0 || r >= d->m || c < 0 || c >= d->n) return -1;
        }
        return 0;
    }
    default:
        return -1;
    }
}

/* Dense/CSC/COO/CSR -> CSR.  reference src/utils.cu:591-804 and src/pdhcg.c:232-272.
 * Dense keeps |v| > tol with tol defaulting to 1e-12; CSC/COO drop |v| <= tol
 * only when tol > 0; CSR is copied verbatim.  Row order within a row follows
 * the input order (counting sort). */
static int to_csr(const matrix_desc_t *d, CsrComponent *out, int *nnz_out) {
    int m = d->m, n = d->n;
    memset(out, 0, sizeof(*out));
    if (d->fmt == matrix_dense) {
        double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 1e-12;
        const double *A = d->data.dense.A;
        long long total = (long long)m * n;
        int nnz = 0;
        for (long long i = 0; i < total; ++i) nnz += fabs(A[i]) > tol;
        out->row_ptr = (int *)xmalloc((size_t)(m + 1) * sizeof(int));
        out->col_ind = (int *)xmalloc((size_t)nnz * sizeof(int));
        out->val = (double *)xmalloc((size_t)nnz * sizeof(double));
        int p = 0;
        for (int i = 0; i < m; ++i) {
            out->row_ptr[i] = p;
            for (int j = 0; j < n; ++j) {
                double v = A[(size_t)i * n + j];
                if (fabs(v) > tol) { out->col_ind[p] = j; out->val[p] = v; ++p; }
            }
        }
        out->row_ptr[m] = p;
        *nnz_out = p;
        return 0;
    }
    if (d->fmt == matrix_csr) {
        int nnz = d->data.csr.nnz;
        out->row_ptr = (int *)xmalloc((size_t)(m + 1) * sizeof(int));
        memcpy(out->row_ptr, d->data.csr.row_ptr, (size_t)(m + 1) * sizeof(int));
        if (nnz > 0) {
            out->col_ind = (int *)xmalloc((size_t)nnz * sizeof(int));
            out->val = (double *)xmalloc((size_t)nnz * sizeof(double));
            memcpy(out->col_ind, d->data.csr.col_ind, (size_t)nnz * sizeof(int));
            memcpy(out->val, d->data.csr.vals, (size_t)nnz * sizeof(double));
        }
        *nnz_out = nnz;
        return 0;
    }
    /* CSC and COO share one counting sort over (row, col, val) triples visited in input order. */
    int nin;
    const int *ri, *ci = NULL;
    const double *v;
    const int *cp = NULL;
    if (d->fmt == matrix_csc) {
        nin = d->data.csc.nnz; ri = d->data.csc.row_ind; v = d->data.csc.vals; cp = d->data.csc.col_ptr;
    } else if (d->fmt == matrix_coo) {
        nin = d->data.coo.nnz; ri = d->data.coo.row_ind; ci = d->data.coo.col_ind; v = d->data.coo.vals;
    } else {
        return -1;
    }
    double tol = d->zero_tolerance > 0 ? d->zero_tolerance : 0.0;
    int *count = (int *)xcalloc((size_t)m + 1, sizeof(int));
    int kept = 0;
    for (int e = 0; e < nin; ++e) {
        if (tol > 0 && fabs(v[e]) <= tol) continue;
        if (ri[e] < 0 || ri[e] >= m) { free(count); return -1; }
        count[ri[e] + 1]++;
        kept++;
    }
    for (int i = 0; i < m; ++i) count[i + 1] += count[i];
    out->row_ptr = count;
    out->col_ind = (int *)xmalloc((size_t)kept * sizeof(int));
    out->val = (double *)xmalloc((size_t)kept * sizeof(double));
    int *next = (int *)xmalloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    for (int i = 0; i < m; ++i) next[i] = count[i];
    if (d->fmt == matrix_csc) {
        for (int j = 0; j < n; ++j)
            for (int e = cp[j]; e < cp[j + 1]; ++e) {
                if (tol > 0 && fabs(v[e]) <= tol) continue;
                int pos = next[ri[e]]++;
                out->col_ind[pos] = j;
                out->val[pos] = v[e];
            }
    } else {
        for (int e = 0; e < nin; ++e) {
            if (tol > 0 && fabs(v[e]) <= tol) continue;
            if (ci[e] < 0 || ci[e] >= n) { free(next); return -1; }
            int pos = next[ri[e]]++;
            out->col_ind[pos] = ci[e];
            out->val[pos] = v[e];
        }
    }
    free(next);
    *nnz_out = kept;
    return 0;
}

static void csr_release(CsrComponent *c) {
    if (!c) return;
    free(c->row_ptr); free(c->col_ind); free(c->val);
    memset(c, 0, sizeof(*c));
}

void orc_qp_problem_free(qp_problem_t *p) {
    if (!p) return;
    csr_release(p->objective_sparse_matrix); free(p->objective_sparse_matrix);
    csr_release(p->objective_lowrank_matrix); free(p->objective_lowrank_matrix);
    csr_release(p->constraint_matrix); free(p->constraint_matrix);
    csr_release(p->objective_lowrank_middle_matrix); free(p->objective_lowrank_middle_matrix);
    free(p->variable_lower_bound); free(p->variable_upper_bound); free(p->objective_vector);
    free(p->constraint_lower_bound); free(p->constraint_upper_bound); free(p->affine_cone_offset);
    free(p->primal_start); free(p->dual_start);
    cone_blocks_release(&p->cones);
    cone_blocks_release(&p->affine_cones);
    free(p);
}

static double *fill_copy(int n, const double *src, double fill) {
    double *d = (double *)xmalloc((size_t)n * sizeof(double));
    if (src) memcpy(d, src, (size_t)n * sizeof(double));
    else for (int i = 0; i < n; ++i) d[i] = fill;
    return d;
}

/* reference src/pdhcg.c:365-508 (with the shape checks of :170-230 and the
 * F-row append of :280-330). */
qp_problem_t *orc_create_qp_problem(const double *c, const matrix_desc_t *Qd, const matrix_desc_t *Rd,
                                    const matrix_desc_t *Dd, const matrix_desc_t *Ad, const double *con_lb,
                                    const double *con_ub, const double *var_lb, const double *var_ub,
                                    const double *objective_constant, int num_var_cones, const cone_spec_t *var_cones,
                                    const matrix_desc_t *Fd, const double *g, int num_affine_cones,
                                    const cone_spec_t *affine_cones) {
    qp_problem_t *p = (qp_problem_t *)xcalloc(1, sizeof(qp_problem_t));
    int n, m, q;
    if (!Fd && (g || num_affine_cones != 0 || affine_cones)) goto fail;
    if (!Ad && !Fd && !Qd && !Rd) goto fail;
    if (check_desc(Ad, "A") || check_desc(Fd, "F") || check_desc(Qd, "Q") || check_desc(Rd, "R")) goto fail;
    n = Ad ? Ad->n : (Fd ? Fd->n : (Qd ? Qd->n : Rd->n));
    m = Ad ? Ad->m : 0;
    q = Fd ? Fd->m : 0;
    if (Fd && Fd->n != n) goto fail;
    if (Qd && (Qd->m != n || Qd->n != n)) goto fail;
    if (Rd && Rd->n != n) goto fail;
    if (Dd && Rd && Rd->m > 0) {
        if (check_desc(Dd, "D") || Dd->m != Rd->m || Dd->n != Rd->m) goto fail;
    }
    if (q > INT_MAX - m) goto fail;

    p->num_variables = n;
    p->num_constraints = m + q;
    p->affine_cone_offset = p->num_constraints > 0 ? (double *)xcalloc((size_t)p->num_constraints, sizeof(double)) : NULL;

    p->constraint_matrix = (CsrComponent *)xcalloc(1, sizeof(CsrComponent));
    if (Ad) {
        if (to_csr(Ad, p->constraint_matrix, &p->constraint_matrix_num_nonzeros)) goto fail;
    } else {
        p->constraint_matrix->row_ptr = (int *)xcalloc((size_t)m + 1, sizeof(int));
        p->constraint_matrix_num_nonzeros = 0;
    }
    if (Fd) {
        CsrComponent f;
        int fnnz = 0;
        if (to_csr(Fd, &f, &fnnz)) goto fail;
        if (fnnz > INT_MAX - p->constraint_matrix_num_nonzeros) { csr_release(&f); goto fail; }
        int base = p->constraint_matrix_num_nonzeros, tot = base + fnnz;
        CsrComponent *C = p->constraint_matrix;
        C->row_ptr = (int *)realloc(C->row_ptr, ((size_t)m + q + 1) * sizeof(int));
        for (int r = 1; r <= q; ++r) C->row_ptr[m + r] = base + f.row_ptr[r];
        if (fnnz > 0) {
            C->col_ind = (int *)realloc(C->col_ind, (size_t)tot * sizeof(int));
            C->val = (double *)realloc(C->val, (size_t)tot * sizeof(double));
            memcpy(C->col_ind + base, f.col_ind, (size_t)fnnz * sizeof(int));
            memcpy(C->val + base, f.val, (size_t)fnnz * sizeof(double));
        }
        p->constraint_matrix_num_nonzeros = tot;
        csr_release(&f);
    }

    p->objective_sparse_matrix = (CsrComponent *)xcalloc(1, sizeof(CsrComponent));
    if (Qd) {
        if (to_csr(Qd, p->objective_sparse_matrix, &p->objective_sparse_matrix_num_nonzeros)) goto fail;
    } else {
        p->objective_sparse_matrix->row_ptr = (int *)xcalloc((size_t)n + 1, sizeof(int));
    }
    p->objective_lowrank_matrix = (CsrComponent *)xcalloc(1, sizeof(CsrComponent));
    if (Rd) {
        p->num_rank_lowrank_obj = Rd->m;
        if (to_csr(Rd, p->objective_lowrank_matrix, &p->objective_lowrank_matrix_num_nonzeros)) goto fail;
    } else {
        p->objective_lowrank_matrix->row_ptr = (int *)xcalloc(1, sizeof(int));
    }
    if (Dd) {
        if (p->num_rank_lowrank_obj <= 0) {
            fprintf(stderr, "[interface] D matrix ignored: problem has no low-rank component.\n");
        } else {
            p->objective_lowrank_middle_matrix = (CsrComponent *)xcalloc(1, sizeof(CsrComponent));
            if (to_csr(Dd, p->objective_lowrank_middle_matrix, &p->objective_lowrank_middle_matrix_num_nonzeros))
                goto fail;
        }
    }
    p->objective_constant = objective_constant ? *objective_constant : 0.0;
    p->objective_vector = fill_copy(n, c, 0.0);
    p->variable_lower_bound = fill_copy(n, var_lb, -INFINITY);
    p->variable_upper_bound = fill_copy(n, var_ub, INFINITY);
    p->constraint_lower_bound = fill_copy(p->num_constraints, NULL, -INFINITY);
    p->constraint_upper_bound = fill_copy(p->num_constraints, NULL, INFINITY);
    if (m > 0 && con_lb) memcpy(p->constraint_lower_bound, con_lb, (size_t)m * sizeof(double));
    if (m > 0 && con_ub) memcpy(p->constraint_upper_bound, con_ub, (size_t)m * sizeof(double));

    /* affine cones: reference src/pdhcg.c:332-363 */
    if (cone_blocks_from_specs(&p->affine_cones, num_affine_cones, affine_cones, q, 0, "affine")) goto fail;
    {
        int covered = 0;
        for (int k = 0; k < num_affine_cones; ++k) {
            covered += orc_cone_block_length(&p->affine_cones, k);
            p->affine_cones.start_idx[k] += m;
        }
        if (covered != q) goto fail;
        if (g && q > 0) memcpy(p->affine_cone_offset + m, g, (size_t)q * sizeof(double));
    }
    if (cone_blocks_from_specs(&p->cones, num_var_cones, var_cones, n, 1, "variable")) goto fail;
    {
        int n_orig = n;
        for (int k = 0; k < p->cones.num_cones; ++k) {
            int s = p->cones.start_idx[k], L = orc_cone_block_length(&p->cones, k);
            for (int v = s; v < s + L; ++v) {
                double lo = p->variable_lower_bound[v], hi = p->variable_upper_bound[v];
                if ((isfinite(lo) && lo > -1e30) || (isfinite(hi) && hi < 1e30)) {
                    fprintf(stderr, "[create_qp_problem] cone %d slot %d has a finite box bound.\n", k, v);
                    goto fail;
                }
            }
            if (s < n_orig) n_orig = s;
        }
        if (p->cones.num_cones > 0) p->num_original_variables = n_orig;
    }
    return p;
fail:
    fprintf(stderr, "[create_qp_problem] invalid problem data.\n");
    orc_qp_problem_free(p);
    return NULL;
}

/* reference src/pdhcg.c:569-618 */
void orc_set_start_values(qp_problem_t *p, const double *primal, const double *dual) {
    if (!p) return;
    int n = p->num_variables, m = p->num_constraints;
    if (primal && p->cones.is_fixed && p->primal_start)
        for (int i = 0; i < n; ++i)
            if (p->cones.is_fixed[i] && primal[i] != p->primal_start[i]) {
                fprintf(stderr, "[set_start_values] slot %d is fixed; rejecting.\n", i);
                return;
            }
    double *np_ = NULL, *nd = NULL;
    if (primal) {
        np_ = (double *)xmalloc((size_t)n * sizeof(double));
        memcpy(np_, primal, (size_t)n * sizeof(double));
    } else if (p->cones.is_fixed && p->primal_start) {
        np_ = (double *)xcalloc((size_t)n, sizeof(double));
        for (int i = 0; i < n; ++i)
            if (p->cones.is_fixed[i]) np_[i] = p->primal_start[i];
    }
    if (dual) {
        nd = (double *)xmalloc((size_t)m * sizeof(double));
        memcpy(nd, dual, (size_t)m * sizeof(double));
    }
    free(p->primal_start);
    free(p->dual_start);
    p->primal_start = np_;
    p->dual_start = nd;
}

/* reference src/pdhcg.c:620-666 */
int orc_set_cone_fixed(qp_problem_t *p, int k, int slot, double value) {
    if (!p || k < 0 || k >= p->cones.num_cones) return -1;
    if (p->cones.type[k] == CONE_PSD) return -1;
    int L = orc_cone_block_length(&p->cones, k);
    if (slot < 0 || slot >= L) return -1;
    int idx = p->cones.start_idx[k] + slot;
    if (idx < 0 || idx >= p->num_variables) return -1;
    if (!isfinite(value)) return -1;
    if (!p->cones.is_fixed) {
        p->cones.is_fixed = (char *)xcalloc((size_t)p->num_variables, 1);
        p->cones.fixed_mask_size = p->num_variables;
    }
    p->cones.is_fixed[idx] = 1;
    if (!p->primal_start) p->primal_start = (double *)xcalloc((size_t)p->num_variables, sizeof(double));
    p->primal_start[idx] = value;
    return 0;
}

static double fixed_norm(const qp_problem_t *p, int start, int len) {
    double r = 0.0;
    for (int j = 0; j < len; ++j)
        if (p->cones.is_fixed[start + j]) r = hypot(r, p->primal_start ? p->primal_start[start + j] : 0.0);
    return r;
}

/* reference src/pdhcg.c:683-727 */
static int exp_fixed_nonempty(const qp_problem_t *p, int s) {
    int fx = p->cones.is_fixed[s] != 0, fy = p->cones.is_fixed[s + 1] != 0, fz = p->cones.is_fixed[s + 2] != 0;
    double x = p->primal_start ? p->primal_start[s] : 0.0;
    double y = p->primal_start ? p->primal_start[s + 1] : 0.0;
    double z = p->primal_start ? p->primal_start[s + 2] : 0.0;
    if ((fx && !isfinite(x)) || (fy && !isfinite(y)) || (fz && !isfinite(z))) return 0;
    if (fy) {
        if (y < 0.0) return 0;
        if (y == 0.0) return (!fx || x <= 0.0) && (!fz || z >= 0.0);
        if (fz && !(z > 0.0)) return 0;
        if (fx && fz) {
            double lb = log(y) + x / y, lz = log(z);
            return lb <= lz + 64.0 * DBL_EPSILON * (1.0 + fabs(lb) + fabs(lz));
        }
        return 1;
    }
    if (!fz) return 1;
    if (z < 0.0) return 0;
    if (z == 0.0) return !fx || x <= 0.0;
    if (!fx || x <= 0.0) return 1;
    double lm = 1.0 + log(x), lz = log(z);
    return lm <= lz + 64.0 * DBL_EPSILON * (1.0 + fabs(lm) + fabs(lz));
}

/* reference src/pdhcg.c:729-892 */
int orc_validate_fixed_cone_sections(const qp_problem_t *p) {
    if (!p || !p->cones.is_fixed) return 0;
    if (p->cones.fixed_mask_size != p->num_variables) return -1;
    for (int k = 0; k < p->cones.num_cones; ++k) {
        int s = p->cones.start_idx[k], vd = p->cones.v_dim[k], L = orc_cone_block_length(&p->cones, k);
        int any = 0;
        for (int j = 0; j < L; ++j) any |= p->cones.is_fixed[s + j] != 0;
        if (!any) continue;
        cone_type_t t = p->cones.type[k];
        if (t == CONE_EXPONENTIAL) {
            if (!exp_fixed_nonempty(p, s)) return -1;
            continue;
        }
        if (t == CONE_STANDARD_SOC || t == CONE_ROTATED_SOC) {
            for (int j = 0; j < vd + 2; ++j)
                if (p->cones.is_fixed[s + j] && !isfinite(p->primal_start ? p->primal_start[s + j] : 0.0)) return -1;
        }
        if (t == CONE_STANDARD_SOC) {
            int zi = s + vd + 1;
            int fz = p->cones.is_fixed[zi] != 0;
            double z = p->primal_start ? p->primal_start[zi] : 0.0;
            double fn = fixed_norm(p, s, vd + 1);
            if (fz && (!(z >= 0.0) || fn > z)) return -1;
            continue;
        }
        if (t == CONE_ROTATED_SOC) {
            int si = s + vd, ti = si + 1;
            int fs = p->cones.is_fixed[si] != 0, ft = p->cones.is_fixed[ti] != 0;
            double sv = p->primal_start ? p->primal_start[si] : 0.0;
            double tv = p->primal_start ? p->primal_start[ti] : 0.0;
            double fn = fixed_norm(p, s, vd);
            int empty = (fs && sv < 0.0) || (ft && tv < 0.0);
            if (!empty && fs && ft) empty = fn > 1.41421356237309504880 * sqrt(sv) * sqrt(tv);
            else if (fs && sv == 0.0) empty |= fn > 0.0;
            else if (ft && tv == 0.0) empty |= fn > 0.0;
            if (empty) return -1;
            continue;
        }
        if (t != CONE_POWER) continue;
        int fx = p->cones.is_fixed[s] != 0, fy = p->cones.is_fixed[s + 1] != 0, fz = p->cones.is_fixed[s + 2] != 0;
        double x = p->primal_start ? p->primal_start[s] : 0.0;
        double y = p->primal_start ? p->primal_start[s + 1] : 0.0;
        double z = p->primal_start ? p->primal_start[s + 2] : 0.0;
        if ((fx && (!isfinite(x) || x < 0.0)) || (fy && (!isfinite(y) || y < 0.0)) || (fz && !isfinite(z))) return -1;
        if (fz && z != 0.0 && ((fx && x == 0.0) || (fy && y == 0.0))) return -1;
        if (fx && fy && fz && z != 0.0) {
            double a = p->cones.power_alpha[k];
            double lb = (x > 0.0 && y > 0.0) ? a * log(x) + (1.0 - a) * log(y) : -INFINITY;
            double lz = log(fabs(z));
            if (lb + 64.0 * DBL_EPSILON * (1.0 + fabs(lb) + fabs(lz)) < lz) return -1;
        }
    }
    return 0;
}

/* reference src/utils.cu:324-368 */
void orc_set_default_parameters(pdhg_parameters_t *p) {
    memset(p, 0, sizeof(*p));
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
    p->termination_criteria.iteration_limit = INT32_MAX;
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

/* Stable counting-sort transpose (what cusparseCsr2cscEx2 returns). */
void orc_csr_transpose(int rows, int cols, const int *rp, const int *ci, const double *v, int *trp, int *tci,
                       double *tv) {
    for (int j = 0; j <= cols; ++j) trp[j] = 0;
    int nnz = rp[rows];
    for (int e = 0; e < nnz; ++e) trp[ci[e] + 1]++;
    for (int j = 0; j < cols; ++j) trp[j + 1] += trp[j];
    int *next = (int *)xmalloc((size_t)(cols > 0 ? cols : 1) * sizeof(int));
    for (int j = 0; j < cols; ++j) next[j] = trp[j];
    for (int i = 0; i < rows; ++i)
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            int pos = next[ci[e]]++;
            tci[pos] = i;
            tv[pos] = v[e];
        }
    free(next);
}
