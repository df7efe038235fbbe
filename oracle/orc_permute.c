/*
 * ORACLE (test infrastructure, see oracle.h): cone-aware permutations of a
 * problem and partition cuts, restated in plain C.
 *   permutations : reference src/permute.cu:36-494
 *   cut optimiser: reference src/partition_utils.c:23-248
 *   balanced cuts: reference distributed/distributed_utils.cu:179-216,298-307
 * All integer work; the random permutations draw from the C library's rand()
 * exactly as the reference does, so a caller that seeds with srand() gets the
 * reference's permutation.
 */
#include "orc_internal.h"

#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void *pmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) abort();
    return p;
}

static void *dup_bytes(const void *src, size_t n) {
    if (!src) return NULL;
    void *d = pmalloc(n);
    memcpy(d, src, n);
    return d;
}

static CsrComponent *csr_copy(const CsrComponent *c, int rows, int nnz) {
    if (!c) return NULL;
    CsrComponent *d = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    d->row_ptr = (int *)dup_bytes(c->row_ptr, ((size_t)rows + 1) * sizeof(int));
    d->col_ind = nnz > 0 ? (int *)dup_bytes(c->col_ind, (size_t)nnz * sizeof(int)) : NULL;
    d->val = nnz > 0 ? (double *)dup_bytes(c->val, (size_t)nnz * sizeof(double)) : NULL;
    return d;
}

static void cones_copy(cone_blocks_t *d, const cone_blocks_t *s) {
    memset(d, 0, sizeof(*d));
    d->num_cones = s->num_cones;
    if (s->num_cones > 0) {
        size_t k = (size_t)s->num_cones;
        d->start_idx = (int *)dup_bytes(s->start_idx, k * sizeof(int));
        d->v_dim = (int *)dup_bytes(s->v_dim, k * sizeof(int));
        d->type = (cone_type_t *)dup_bytes(s->type, k * sizeof(cone_type_t));
        d->power_alpha = (double *)dup_bytes(s->power_alpha, k * sizeof(double));
    }
    d->fixed_mask_size = s->fixed_mask_size;
    d->is_fixed = s->fixed_mask_size > 0 ? (char *)dup_bytes(s->is_fixed, (size_t)s->fixed_mask_size) : NULL;
}

/* Deep copy of an assembled problem (what reference deepcopy_problem gives
 * permute_problem_return_new, src/permute.cu:420-433). */
qp_problem_t *orc_qp_problem_copy(const qp_problem_t *p) {
    if (!p) return NULL;
    int n = p->num_variables, m = p->num_constraints, k = p->num_rank_lowrank_obj;
    qp_problem_t *d = (qp_problem_t *)calloc(1, sizeof(qp_problem_t));
    *d = *p;
    d->variable_lower_bound = (double *)dup_bytes(p->variable_lower_bound, (size_t)n * sizeof(double));
    d->variable_upper_bound = (double *)dup_bytes(p->variable_upper_bound, (size_t)n * sizeof(double));
    d->objective_vector = (double *)dup_bytes(p->objective_vector, (size_t)n * sizeof(double));
    d->constraint_lower_bound = (double *)dup_bytes(p->constraint_lower_bound, (size_t)m * sizeof(double));
    d->constraint_upper_bound = (double *)dup_bytes(p->constraint_upper_bound, (size_t)m * sizeof(double));
    d->affine_cone_offset = (double *)dup_bytes(p->affine_cone_offset, (size_t)m * sizeof(double));
    d->primal_start = (double *)dup_bytes(p->primal_start, (size_t)n * sizeof(double));
    d->dual_start = (double *)dup_bytes(p->dual_start, (size_t)m * sizeof(double));
    d->constraint_matrix = csr_copy(p->constraint_matrix, m, p->constraint_matrix_num_nonzeros);
    d->objective_sparse_matrix = csr_copy(p->objective_sparse_matrix, n, p->objective_sparse_matrix_num_nonzeros);
    d->objective_lowrank_matrix = csr_copy(p->objective_lowrank_matrix, k, p->objective_lowrank_matrix_num_nonzeros);
    d->objective_lowrank_middle_matrix =
        csr_copy(p->objective_lowrank_middle_matrix, k, p->objective_lowrank_middle_matrix_num_nonzeros);
    cones_copy(&d->cones, &p->cones);
    cones_copy(&d->affine_cones, &p->affine_cones);
    d->num_quadratic_constraints = 0;
    d->quadratic_constraint_row_indices = NULL;
    d->quadratic_constraint_matrices = NULL;
    d->quadratic_constraint_matrix_num_nonzeros = NULL;
    return d;
}

/* ---- permutations: reference src/permute.cu ---- */

typedef struct {
    int new_col;
    double val;
} col_val_t;

static int cmp_col_val(const void *a, const void *b) {
    int x = ((const col_val_t *)a)->new_col, y = ((const col_val_t *)b)->new_col;
    return (x > y) - (x < y);
}

/* Renames columns by old_to_new and re-sorts each row by new column
 * (reference :41-84; columns within a row are distinct, so the order is
 * unique). */
static void permute_columns(int rows, const int *rp, int *ci, double *v, const int *old_to_new) {
    int longest = 0;
    for (int r = 0; r < rows; ++r)
        if (rp[r + 1] - rp[r] > longest) longest = rp[r + 1] - rp[r];
    col_val_t *buf = (col_val_t *)pmalloc((size_t)longest * sizeof(col_val_t));
    for (int r = 0; r < rows; ++r) {
        int s = rp[r], len = rp[r + 1] - s;
        for (int e = 0; e < len; ++e) {
            buf[e].new_col = old_to_new[ci[s + e]];
            buf[e].val = v[s + e];
        }
        if (len > 1) qsort(buf, (size_t)len, sizeof(col_val_t), cmp_col_val);
        for (int e = 0; e < len; ++e) {
            ci[s + e] = buf[e].new_col;
            v[s + e] = buf[e].val;
        }
    }
    free(buf);
}

/* New row i is old row perm[i] (reference :86-121). */
static void permute_rows(CsrComponent *c, int rows, int nnz, const int *perm) {
    if (!c || nnz == 0) return;
    int *rp = (int *)pmalloc(((size_t)rows + 1) * sizeof(int));
    int *ci = (int *)pmalloc((size_t)nnz * sizeof(int));
    double *v = (double *)pmalloc((size_t)nnz * sizeof(double));
    int at = 0;
    rp[0] = 0;
    for (int i = 0; i < rows; ++i) {
        int s = c->row_ptr[perm[i]], len = c->row_ptr[perm[i] + 1] - s;
        memcpy(ci + at, c->col_ind + s, (size_t)len * sizeof(int));
        memcpy(v + at, c->val + s, (size_t)len * sizeof(double));
        at += len;
        rp[i + 1] = at;
    }
    free(c->row_ptr); free(c->col_ind); free(c->val);
    c->row_ptr = rp; c->col_ind = ci; c->val = v;
}

static void gather(double *a, int n, const int *perm) {
    if (!a) return;
    double *t = (double *)pmalloc((size_t)n * sizeof(double));
    for (int i = 0; i < n; ++i) t[i] = a[perm[i]];
    memcpy(a, t, (size_t)n * sizeof(double));
    free(t);
}

/* inverse[perm[i]] = i, or NULL when perm is not a permutation of 0..n-1
 * (reference :235-262). */
static int *checked_inverse(int n, const int *perm, int *ok) {
    *ok = 1;
    if (n <= 0) return NULL;
    if (!perm) { *ok = 0; return NULL; }
    int *inv = (int *)pmalloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) inv[i] = -1;
    for (int i = 0; i < n; ++i) {
        int v = perm[i];
        if (v < 0 || v >= n || inv[v] >= 0) { free(inv); *ok = 0; return NULL; }
        inv[v] = i;
    }
    return inv;
}

/* Every block keeps its slots contiguous and in order (reference :264-316). */
static int blocks_kept_whole(const cone_blocks_t *b, int n, const int *perm) {
    int ok;
    int *inv = checked_inverse(n, perm, &ok);
    if (!ok) return 0;
    for (int k = 0; k < b->num_cones && ok; ++k) {
        int s = b->start_idx[k], L = orc_cone_block_length(b, k), ns = inv[s];
        for (int j = 1; j < L; ++j)
            if (inv[s + j] != ns + j) { ok = 0; break; }
    }
    free(inv);
    return ok;
}

int orc_validate_cone_permutation(const qp_problem_t *p, const int *col_perm) {
    return p ? blocks_kept_whole(&p->cones, p->num_variables, col_perm) : 0;
}

int orc_validate_affine_row_permutation(const qp_problem_t *p, const int *row_perm) {
    return p ? blocks_kept_whole(&p->affine_cones, p->num_constraints, row_perm) : 0;
}

/* reference :143-224 */
int orc_permute_problem(qp_problem_t *p, const int *row_perm, const int *col_perm) {
    if (!p || (p->num_constraints > 0 && !row_perm) || (p->num_variables > 0 && !col_perm)) return 0;
    int m = p->num_constraints, n = p->num_variables;
    if (!orc_validate_cone_permutation(p, col_perm)) {
        fprintf(stderr, "[oracle] column permutation splits or reorders a cone block.\n");
        return 0;
    }
    if (!orc_validate_affine_row_permutation(p, row_perm)) {
        fprintf(stderr, "[oracle] row permutation splits or reorders an affine cone block.\n");
        return 0;
    }
    gather(p->objective_vector, n, col_perm);
    gather(p->variable_lower_bound, n, col_perm);
    gather(p->variable_upper_bound, n, col_perm);
    gather(p->primal_start, n, col_perm);
    gather(p->constraint_lower_bound, m, row_perm);
    gather(p->constraint_upper_bound, m, row_perm);
    gather(p->affine_cone_offset, m, row_perm);
    gather(p->dual_start, m, row_perm);
    int *inv_row = (int *)pmalloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    for (int i = 0; i < m; ++i) inv_row[row_perm[i]] = i;
    for (int k = 0; k < p->affine_cones.num_cones; ++k) p->affine_cones.start_idx[k] = inv_row[p->affine_cones.start_idx[k]];
    int *inv_col = (int *)pmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < n; ++i) inv_col[col_perm[i]] = i;
    if (p->constraint_matrix && p->constraint_matrix_num_nonzeros > 0) {
        permute_rows(p->constraint_matrix, m, p->constraint_matrix_num_nonzeros, row_perm);
        permute_columns(m, p->constraint_matrix->row_ptr, p->constraint_matrix->col_ind, p->constraint_matrix->val,
                        inv_col);
    }
    if (p->objective_sparse_matrix && p->objective_sparse_matrix_num_nonzeros > 0) {
        permute_rows(p->objective_sparse_matrix, n, p->objective_sparse_matrix_num_nonzeros, col_perm);
        permute_columns(n, p->objective_sparse_matrix->row_ptr, p->objective_sparse_matrix->col_ind,
                        p->objective_sparse_matrix->val, inv_col);
    }
    if (p->objective_lowrank_matrix && p->objective_lowrank_matrix_num_nonzeros > 0)
        permute_columns(p->num_rank_lowrank_obj, p->objective_lowrank_matrix->row_ptr,
                        p->objective_lowrank_matrix->col_ind, p->objective_lowrank_matrix->val, inv_col);
    for (int k = 0; k < p->cones.num_cones; ++k) p->cones.start_idx[k] = inv_col[p->cones.start_idx[k]];
    if (p->cones.num_cones > 0 && p->cones.is_fixed) {
        char *t = (char *)pmalloc((size_t)n);
        for (int i = 0; i < n; ++i) t[i] = p->cones.is_fixed[col_perm[i]];
        memcpy(p->cones.is_fixed, t, (size_t)n);
        free(t);
    }
    free(inv_row);
    free(inv_col);
    return 1;
}

qp_problem_t *orc_permute_problem_copy(const qp_problem_t *p, const int *row_perm, const int *col_perm) {
    qp_problem_t *d = orc_qp_problem_copy(p);
    if (d && !orc_permute_problem(d, row_perm, col_perm)) {
        orc_qp_problem_free(d);
        return NULL;
    }
    return d;
}

typedef struct {
    int start, length;
} unit_t;

static int cmp_unit_start(const void *a, const void *b) {
    int x = ((const unit_t *)a)->start, y = ((const unit_t *)b)->start;
    return (x > y) - (x < y);
}

/* Fisher-Yates over whole units with the C library's rand() (reference
 * :335-406, :436-481). */
static void vector_permutation(int n, const cone_blocks_t *b, int method, int block_size, int *perm) {
    if (method == NO_PERMUTATION || n <= 1) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        return;
    }
    int K = b ? b->num_cones : 0;
    int free_len = method == FULL_RANDOM_PERMUTATION ? 1 : (block_size > 0 ? block_size : 1);
    if (K <= 0 && method == FULL_RANDOM_PERMUTATION) {
        for (int i = 0; i < n; ++i) perm[i] = i;
        for (int i = n - 1; i > 0; --i) {
            int j = rand() % (i + 1);
            int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
        }
        return;
    }
    if (K <= 0) { /* block permutation without cones: shuffle block ids */
        int nb = (n + free_len - 1) / free_len;
        int *ids = (int *)pmalloc((size_t)nb * sizeof(int));
        for (int i = 0; i < nb; ++i) ids[i] = i;
        for (int i = nb - 1; i > 0; --i) {
            int j = rand() % (i + 1);
            int t = ids[i]; ids[i] = ids[j]; ids[j] = t;
        }
        int at = 0;
        for (int i = 0; i < nb; ++i) {
            int s = ids[i] * free_len, e = s + free_len < n ? s + free_len : n;
            for (int v = s; v < e; ++v) perm[at++] = v;
        }
        free(ids);
        return;
    }
    unit_t *cones = (unit_t *)pmalloc((size_t)K * sizeof(unit_t));
    for (int k = 0; k < K; ++k) {
        cones[k].start = b->start_idx[k];
        cones[k].length = orc_cone_block_length(b, k);
    }
    qsort(cones, (size_t)K, sizeof(unit_t), cmp_unit_start);
    unit_t *units = (unit_t *)pmalloc(((size_t)n + (size_t)K) * sizeof(unit_t));
    int nu = 0, cursor = 0;
    for (int k = 0; k < K; ++k) {
        while (cursor < cones[k].start) {
            int len = cones[k].start - cursor < free_len ? cones[k].start - cursor : free_len;
            units[nu].start = cursor; units[nu].length = len; ++nu;
            cursor += len;
        }
        units[nu++] = cones[k];
        cursor = cones[k].start + cones[k].length;
    }
    while (cursor < n) {
        int len = n - cursor < free_len ? n - cursor : free_len;
        units[nu].start = cursor; units[nu].length = len; ++nu;
        cursor += len;
    }
    free(cones);
    for (int i = nu - 1; i > 0; --i) {
        int j = rand() % (i + 1);
        unit_t t = units[i]; units[i] = units[j]; units[j] = t;
    }
    int at = 0;
    for (int u = 0; u < nu; ++u)
        for (int j = 0; j < units[u].length; ++j) perm[at++] = units[u].start + j;
    free(units);
}

void orc_generate_cone_aware_permutation(const qp_problem_t *p, int method, int block_size, int *perm) {
    vector_permutation(p->num_variables, &p->cones, method, block_size, perm);
}

void orc_generate_affine_row_permutation(const qp_problem_t *p, int method, int block_size, int *perm) {
    vector_permutation(p->num_constraints, &p->affine_cones, method, block_size, perm);
}

/* reference :483-494 */
void orc_repermute_solution(pdhcg_result_t *r, const int *row_perm, const int *col_perm) {
    int n = r->num_variables, m = r->num_constraints;
    int *ic = (int *)pmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int *ir = (int *)pmalloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    for (int i = 0; i < n; ++i) ic[col_perm[i]] = i;
    for (int i = 0; i < m; ++i) ir[row_perm[i]] = i;
    gather(r->primal_solution, n, ic);
    gather(r->dual_solution, m, ir);
    gather(r->reduced_cost, n, ic);
    free(ic);
    free(ir);
}

/* ---- partition cuts ---- */

/* cuts[0..P]: cut p is the first index where the running weight reaches
 * p * total / P (reference distributed/distributed_utils.cu:179-216). */
void orc_balanced_cuts(const int *weights, int total, int parts, int *cuts) {
    cuts[0] = 0;
    cuts[parts] = total;
    if (parts == 1) return;
    long long sum = 0;
    for (int i = 0; i < total; ++i) sum += weights[i];
    double target = (double)sum / parts;
    long long run = 0;
    int p = 1;
    for (int i = 0; i < total; ++i) {
        run += weights[i];
        if (run >= p * target) {
            cuts[p++] = i + 1;
            if (p >= parts) break;
        }
    }
    while (p < parts) cuts[p++] = total;
}

typedef struct {
    int first, last;
} span_t;

static int cmp_span(const void *a, const void *b) {
    int x = ((const span_t *)a)->first, y = ((const span_t *)b)->first;
    return (x > y) - (x < y);
}

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Nearest legal boundary at or beyond `at` in `dir` (reference
 * src/partition_utils.c:43-80): boundaries lie in 1..total-1 outside every
 * forbidden span; -1 when there is none. */
static int legal_boundary(const span_t *s, int count, int total, int at, int dir) {
    if (total <= 1) return -1;
    if (at < 1) {
        if (dir < 0) return -1;
        at = 1;
    }
    if (at >= total) {
        if (dir > 0) return -1;
        at = total - 1;
    }
    int lo = 0, hi = count - 1;
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (at < s[mid].first) hi = mid - 1;
        else if (at > s[mid].last) lo = mid + 1;
        else {
            at = dir < 0 ? s[mid].first - 1 : s[mid].last + 1;
            break;
        }
    }
    return at > 0 && at < total ? at : -1;
}

/* Moves the interior cuts of cuts[0..P] to legal boundaries minimising the
 * sum of squared displacements (dynamic programme over candidate
 * boundaries near each target; reference src/partition_utils.c:88-248). */
int orc_optimize_partition_cuts(int total, int parts, const int *fs, const int *fe, int nf, int *cuts) {
    if (total < 0 || parts <= 0 || nf < 0 || !cuts || (nf > 0 && (!fs || !fe))) return 0;
    int need = parts - 1;
    if (need <= 0) return 1;
    if (total < parts) return 0;
    span_t *sp = (span_t *)pmalloc((size_t)(nf > 0 ? nf : 1) * sizeof(span_t));
    for (int i = 0; i < nf; ++i) {
        if (fs[i] < 1 || fe[i] < fs[i] || fe[i] >= total) {
            free(sp);
            return 0;
        }
        sp[i].first = fs[i];
        sp[i].last = fe[i];
    }
    if (nf > 1) qsort(sp, (size_t)nf, sizeof(span_t), cmp_span);
    int ns = 0;
    for (int i = 0; i < nf; ++i) {
        if (ns == 0 || sp[i].first > sp[ns - 1].last + 1) sp[ns++] = sp[i];
        else if (sp[i].last > sp[ns - 1].last) sp[ns - 1].last = sp[i].last;
    }
    size_t cap = 4u * (size_t)parts * (size_t)parts + 4u * (size_t)parts;
    int *cand = (int *)pmalloc(cap * sizeof(int));
    int nc = 0;
    for (int c = 1; c < parts; ++c) {
        int down = legal_boundary(sp, ns, total, cuts[c], -1), up = legal_boundary(sp, ns, total, cuts[c], 1);
        for (int s = 0; s < parts && down > 0; ++s) {
            cand[nc++] = down;
            down = legal_boundary(sp, ns, total, down - 1, -1);
        }
        for (int s = 0; s < parts && up > 0; ++s) {
            cand[nc++] = up;
            up = legal_boundary(sp, ns, total, up + 1, 1);
        }
    }
    int first = legal_boundary(sp, ns, total, 1, 1), last = legal_boundary(sp, ns, total, total - 1, -1);
    for (int s = 0; s < parts && first > 0; ++s) {
        cand[nc++] = first;
        first = legal_boundary(sp, ns, total, first + 1, 1);
    }
    for (int s = 0; s < parts && last > 0; ++s) {
        cand[nc++] = last;
        last = legal_boundary(sp, ns, total, last - 1, -1);
    }
    free(sp);
    qsort(cand, (size_t)nc, sizeof(int), cmp_int);
    int nu = 0;
    for (int i = 0; i < nc; ++i)
        if (nu == 0 || cand[i] != cand[nu - 1]) cand[nu++] = cand[i];
    if (nu < need) {
        free(cand);
        return 0;
    }
    long double *cost = (long double *)pmalloc((size_t)need * nu * sizeof(long double));
    int *pred = (int *)pmalloc((size_t)need * nu * sizeof(int));
    for (int j = 0; j < nu; ++j) {
        long double d = (long double)cand[j] - cuts[1];
        cost[j] = d * d;
        pred[j] = -1;
    }
    for (int c = 1; c < need; ++c) {
        long double best = LDBL_MAX;
        int arg = -1;
        for (int j = 0; j < nu; ++j) {
            if (j > 0) {
                long double prev = cost[(size_t)(c - 1) * nu + j - 1];
                if (prev < best) {
                    best = prev;
                    arg = j - 1;
                }
            }
            size_t e = (size_t)c * nu + j;
            if (arg < 0) {
                cost[e] = LDBL_MAX;
                pred[e] = -1;
                continue;
            }
            long double d = (long double)cand[j] - cuts[c + 1];
            cost[e] = best + d * d;
            pred[e] = arg;
        }
    }
    int arg = -1;
    long double best = LDBL_MAX;
    size_t row = (size_t)(need - 1) * nu;
    for (int j = 0; j < nu; ++j)
        if (cost[row + j] < best) {
            best = cost[row + j];
            arg = j;
        }
    if (arg >= 0)
        for (int c = need - 1; c >= 0; --c) {
            cuts[c + 1] = cand[arg];
            arg = pred[(size_t)c * nu + arg];
        }
    free(cost);
    free(pred);
    free(cand);
    return best < LDBL_MAX;
}
