/*
 * ORACLE (test infrastructure, see oracle.h): the restarted Halpern PDHG
 * iteration with the projected Barzilai-Borwein inner solve, adaptive restart,
 * PID primal weight, KKT residuals and infeasibility rays, in FP64 on the CPU.
 * Restated step for step from reference src/solver.cu:35-210,
 * src/pdhg_core_op.cu, src/kernels/pdhcg_kernels.cu and src/utils.cu:219-322.
 * Reductions are plain sequential sums.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc_cones.h"
#include "orc_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef enum { P_LP, P_QP } ptype_t;

struct orc_solver {
    pdhg_parameters_t params;
    orc_scaled *w;
    int n, m, k;
    orc_csr At, Rt;
    const qp_problem_t *orig;      /* caller problem (result dims) */
    qp_problem_t *dummy;           /* owned when a dummy row was added */
    double obj_norm, bound_norm;
    double q_norm, nonconvexity;
    ptype_t ptype;
    /* iterates (scaled space) */
    double *x0, *x, *xp, *xr, *y0, *y, *yp, *yr;
    double *aty, *ax, *dslack, *pslack, *pres, *dres, *dx, *dy;
    double *qx, *rx, *rxbuf;
    double *cl_fin, *cu_fin, *lb_fin, *ub_fin;
    /* BB inner solver (reference internal/internal_types.h:71-94) */
    int bb;
    double *grad, *dirn, *pbuf, *diag_h, *mdiag, *minv;
    int precond;
    double cached_inv_tau, tol_scale, inner_tol, inner_min_tol;
    int inner_limit, inner_total;
    /* scalars */
    double step_size, primal_weight, pw_err_sum, pw_last_err, best_pw, best_gap;
    double fpe, init_fpe, last_trial_fpe;
    int total_count, inner_count, is_major, do_restart;
    double abs_pr, rel_pr, abs_du, rel_du, pobj, dobj, gap, rel_gap;
    double max_pray, max_dray, pray_lin_obj, dray_obj;
    termination_reason_t reason;
    double cum_time;
    struct timespec t_start;
    /* variable cones (reference internal/internal_types.h:106-127) */
    const cone_blocks_t *cb;
    int has_cones, has_power, has_psd;
    double *proj_warm, *res_warm; /* per-block bisection warm starts */
    double *eff;                  /* c + Qx, the effective objective gradient (QP with cones) */
    double *snap;                 /* BB primal snapshot */
    double *pgp;                  /* projected-gradient point */
    /* affine cones on the rows of [A; F] (reference internal/internal_types.h:106-127,
       src/solver_state.cu:621-638): projections onto diag(1 / con_rescale) K */
    const cone_blocks_t *acb;
    int has_affine;
    double *aff_d, *aff_pw, *aff_rw; /* row scaling of the cones, warm starts */
    double *aproj, *amem, *acomp;    /* dual prox point, dual membership, complementarity */
};

/* Projection of one block (slots v[0..len)) onto diag(d) K. */
static void project_one(cone_type_t type, double *v, int vd, const double *d, double *warm, double alpha) {
    switch (type) {
    case CONE_STANDARD_SOC: orc_soc_project(v, vd, d, warm); break;
    case CONE_ROTATED_SOC: orc_rsoc_project(v, vd, d, warm); break;
    case CONE_EXPONENTIAL: orc_exp_project(v, d); break;
    case CONE_POWER: orc_power_project(v, d, alpha); break;
    case CONE_PSD: orc_psd_project(v, vd); break;
    default: break;
    }
}

/* Applies every block's projection onto diag(d) K to v (reference
 * src/cone_dispatch.cu:38-57). */
static void project_blocks(const cone_blocks_t *b, const double *d, double *v, double *warm) {
    for (int k = 0; k < b->num_cones; ++k) {
        const int st = b->start_idx[k];
        const double al = b->power_alpha ? b->power_alpha[k] : 0.0;
        if (b->is_fixed && orc_section_has_fixed(b->is_fixed + st, orc_cone_block_length(b, k)))
            orc_section_project(b->type[k], v + st, b->v_dim[k], d + st, NULL, b->is_fixed + st, al);
        else project_one(b->type[k], v + st, b->v_dim[k], d + st, warm + k, al);
    }
}

/* Variable cone blocks: projection onto diag(var_rescale) K. */
static void project_cones(orc_solver_t *s, double *v, double *warm) {
    project_blocks(s->cb, s->w->var_rescale, v, warm);
}

/* Cone step of the diagonal-Q proximal update: the projection onto D K in
 * the metric diag(w), w = 1 + tau q, computed as W^-1/2 P_{W^1/2 D K}(W^1/2 v)
 * (reference src/cone_dispatch.cu:59-85 with the *_diag_q kernels, e.g.
 * src/kernels/pdhcg_exp_cone_kernels.cu:624-700).  PSD blocks never get here
 * (they take the BB path). */
static void project_cones_diag_q(orc_solver_t *s, double *v, double tau) {
    const cone_blocks_t *b = s->cb;
    const double *d = s->w->var_rescale, *q = s->w->qdiag;
    for (int k = 0; k < b->num_cones; ++k) {
        int st = b->start_idx[k], L = orc_cone_block_length(b, k);
        if (b->is_fixed && orc_section_has_fixed(b->is_fixed + st, L)) {
            /* fixed section in the metric diag(w) directly (reference
               src/kernels/pdhcg_soc_cone_kernels.cu:1463-1467 and siblings) */
            double *mw = (double *)malloc((size_t)L * sizeof(double));
            for (int i = 0; i < L; ++i) {
                double wi = 1.0 + tau * q[st + i];
                mw[i] = wi > 0.0 ? wi : 1.0;
            }
            orc_section_project(b->type[k], v + st, b->v_dim[k], d + st, mw, b->is_fixed + st,
                                b->power_alpha ? b->power_alpha[k] : 0.0);
            free(mw);
            continue;
        }
        double *sw = (double *)malloc((size_t)L * sizeof(double)), *e = (double *)malloc((size_t)L * sizeof(double));
        for (int i = 0; i < L; ++i) {
            double wi = 1.0 + tau * q[st + i];
            if (!(wi > 0.0)) wi = 1.0;
            sw[i] = sqrt(wi);
            v[st + i] *= sw[i];
            e[i] = sw[i] * d[st + i];
        }
        project_one(b->type[k], v + st, b->v_dim[k], e, s->proj_warm + k, b->power_alpha ? b->power_alpha[k] : 0.0);
        for (int i = 0; i < L; ++i) v[st + i] /= sw[i];
        free(sw);
        free(e);
    }
}

/* x_bar = 2 x_p - x on every cone slot (reference src/cone_dispatch.cu:118-135). */
static void cone_reflection(orc_solver_t *s) {
    const cone_blocks_t *b = s->cb;
    for (int k = 0; k < b->num_cones; ++k) {
        int st = b->start_idx[k], L = orc_cone_block_length(b, k);
        for (int i = st; i < st + L; ++i) s->xr[i] = 2.0 * s->xp[i] - s->x[i];
    }
}

/* Threads used by the row-parallel products (OpenMP); 0 restores the
   default. */
void orc_set_threads(int threads) {
#ifdef _OPENMP
    static int deflt = -1;
    if (deflt < 0) deflt = omp_get_max_threads();
    omp_set_num_threads(threads > 0 ? threads : deflt);
#else
    (void)threads;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Rows per dynamically scheduled chunk: about eight chunks per thread, so
   that a product with few long rows (R x, k rows) is still shared out. */
static int par_chunk(int rows) {
#ifdef _OPENMP
    const int c = rows / (8 * omp_get_max_threads());
    return c > 1 ? c : 1;
#else
    return rows > 1 ? rows : 1;
#endif
}

static void *xm(size_t n) { void *p = malloc(n ? n : 1); if (!p) abort(); return p; }
static double *zeros(int n) { double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); if (!p) abort(); return p; }

/* y = M x over CSR.  Rows are independent: with OpenMP they are shared
   out among threads, each row summed in the same order as serially, so the
   result does not depend on the thread count. */
static void spmv(const orc_csr *M, const double *x, double *y) {
    const int chunk = par_chunk(M->rows);
#pragma omp parallel for schedule(dynamic, chunk) if (M->nnz > 200000)
    for (int r = 0; r < M->rows; ++r) {
        double s = 0.0;
        for (int e = M->rp[r]; e < M->rp[r + 1]; ++e) s += M->v[e] * x[M->ci[e]];
        y[r] = s;
    }
}
static double dot(int n, const double *a, const double *b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double nrm2(int n, const double *a) { return sqrt(dot(n, a, a)); }
/* reference src/utils.cu:548-567 (first index of max |x|; non-finite -> inf) */
static double inf_norm(int n, const double *a) {
    if (n <= 0) return 0.0;
    double mx = -1.0;
    int idx = 0;
    for (int i = 0; i < n; ++i) {
        double v = fabs(a[i]);
        if (isnan(a[i])) return INFINITY;
        if (v > mx) { mx = v; idx = i; }
    }
    if (!isfinite(a[idx])) return INFINITY;
    return fabs(a[idx]);
}
static double clampd(double v, double lo, double hi) { return fmax(lo, fmin(v, hi)); }

/* reference src/utils.cu:535-546 */
static int print_freq(int iter) {
    int step = 10;
    long long thr = 1000;
    while (iter >= thr) { step *= 10; thr *= 10; }
    return step;
}

static void csr_transpose(const orc_csr *A, orc_csr *T) {
    T->rows = A->cols; T->cols = A->rows; T->nnz = A->nnz;
    T->rp = (int *)xm((size_t)(T->rows + 1) * sizeof(int));
    T->ci = (int *)xm((size_t)A->nnz * sizeof(int));
    T->v = (double *)xm((size_t)A->nnz * sizeof(double));
    orc_csr_transpose(A->rows, A->cols, A->rp, A->ci, A->v, T->rp, T->ci, T->v);
}

/* ---- Q products: reference src/pdhg_core_op.cu:272-413 ---- */
static void apply_middle(orc_solver_t *s) {
    orc_scaled *w = s->w;
    int k = s->k;
    if (w->dkind == 0 || k <= 0) return;
    if (w->dkind == 1) {
        for (int r = 0; r < k; ++r) s->rx[r] *= w->ddiag[r];
        return;
    }
    /* Dsymv with the lower triangle of the column-major view of the row-major
       array, i.e. the upper triangle of the row-major D. */
    for (int i = 0; i < k; ++i) {
        double acc = 0.0;
        for (int j = 0; j < k; ++j) {
            int a = i < j ? i : j, b = i < j ? j : i;
            acc += w->ddense[(size_t)a * k + b] * s->rx[j];
        }
        s->rxbuf[i] = acc;
    }
    memcpy(s->rx, s->rxbuf, (size_t)k * sizeof(double));
}

static void obj_product(orc_solver_t *s, const double *x) {
    orc_scaled *w = s->w;
    int n = s->n;
    switch (w->qtype) {
    case PDHCG_NON_Q:
        return;
    case PDHCG_DIAG_Q:
        for (int i = 0; i < n; ++i) s->qx[i] = w->qdiag[i] * x[i];
        break;
    case PDHCG_SPARSE_Q:
        spmv(&w->Q, x, s->qx);
        break;
    case PDHCG_LOW_RANK_Q:
        spmv(&w->R, x, s->rx);
        apply_middle(s);
        spmv(&s->Rt, s->rx, s->qx);
        break;
    case PDHCG_LOW_RANK_PLUS_SPARSE_Q:
        spmv(&w->Q, x, s->qx);
        spmv(&w->R, x, s->rx);
        apply_middle(s);
        const int chunk = par_chunk(n);
#pragma omp parallel for schedule(dynamic, chunk) if (s->Rt.nnz > 200000)
        for (int r = 0; r < n; ++r) {
            double acc = 0.0;
            for (int e = s->Rt.rp[r]; e < s->Rt.rp[r + 1]; ++e) acc += s->Rt.v[e] * s->rx[s->Rt.ci[e]];
            s->qx[r] += acc;
        }
        break;
    }
    if (s->eff)  /* reference src/pdhg_core_op.cu:405-412 */
        for (int i = 0; i < n; ++i) s->eff[i] = w->c[i] + s->qx[i];
}

/* ---- spectral estimates: reference src/pdhg_core_op.cu:1498-1788 ---- */
static double est_max_eig(const orc_csr *Q, int maxit, double tol) {
    int n = Q->rows;
    double *v = zeros(n), *av = zeros(n);
    unsigned int seed = 1234;
    for (int i = 0; i < n; ++i) v[i] = (double)rand_r(&seed) / RAND_MAX;
    double lam = 0.0;
    for (int it = 0; it < maxit; ++it) {
        double nv = nrm2(n, v), inv = 1.0 / nv;
        for (int i = 0; i < n; ++i) v[i] *= inv;
        spmv(Q, v, av);
        double old = lam;
        lam = dot(n, v, av);
        if (it > 0 && fabs(lam - old) < tol) break;
        memcpy(v, av, (size_t)n * sizeof(double));
    }
    free(v); free(av);
    return lam;
}

static double est_min_eig(const orc_csr *Q, double lmax, int maxit, double tol) {
    int n = Q->rows;
    double *v = zeros(n), *av = zeros(n), *sh = zeros(n);
    unsigned int seed = 1234;
    for (int i = 0; i < n; ++i) v[i] = (double)rand_r(&seed) / RAND_MAX;
    double mu = 0.0;
    for (int it = 0; it < maxit; ++it) {
        double nv = nrm2(n, v), inv = 1.0 / nv;
        for (int i = 0; i < n; ++i) v[i] *= inv;
        spmv(Q, v, av);
        double old = mu;
        for (int i = 0; i < n; ++i) sh[i] = -av[i] + lmax * v[i];
        mu = dot(n, v, sh);
        if (it > 0 && fabs(mu - old) < tol) break;
        memcpy(v, sh, (size_t)n * sizeof(double));
    }
    free(v); free(av); free(sh);
    return lmax - mu;
}

static double est_max_sv(const orc_csr *A, const orc_csr *At, int maxit, double tol) {
    int m = A->rows, n = A->cols;
    double *e = zeros(m), *ne = zeros(m), *dp = zeros(n);
    unsigned int seed = 1234;
    for (int i = 0; i < m; ++i) e[i] = ((double)rand_r(&seed) / (double)RAND_MAX) * 2.0 - 1.0;
    double s2 = 1.0;
    double inv = 1.0 / nrm2(m, e);
    for (int i = 0; i < m; ++i) e[i] *= inv;
    for (int it = 0; it < maxit; ++it) {
        spmv(At, e, dp);
        spmv(A, dp, ne);
        s2 = dot(m, ne, e);
        for (int i = 0; i < m; ++i) ne[i] += -s2 * e[i];
        double res = nrm2(m, ne);
        if (res < tol) break;
        for (int i = 0; i < m; ++i) ne[i] += s2 * e[i];
        inv = 1.0 / nrm2(m, ne);
        for (int i = 0; i < m; ++i) ne[i] *= inv;
        double *t = e; e = ne; ne = t;
    }
    free(e); free(ne); free(dp);
    return sqrt(s2);
}

/* reference src/utils.cu:82-134 */
static qp_problem_t *with_dummy_row(const qp_problem_t *p) {
    qp_problem_t *d = (qp_problem_t *)calloc(1, sizeof(qp_problem_t));
    *d = *p; /* shallow: only the constraint part is replaced */
    d->num_constraints = 1;
    d->constraint_matrix_num_nonzeros = 1;
    d->constraint_matrix = (CsrComponent *)calloc(1, sizeof(CsrComponent));
    d->constraint_matrix->row_ptr = (int *)xm(2 * sizeof(int));
    d->constraint_matrix->row_ptr[0] = 0;
    d->constraint_matrix->row_ptr[1] = 1;
    d->constraint_matrix->col_ind = (int *)xm(sizeof(int));
    d->constraint_matrix->col_ind[0] = 0;
    d->constraint_matrix->val = (double *)xm(sizeof(double));
    d->constraint_matrix->val[0] = 1.0;
    d->constraint_lower_bound = (double *)xm(sizeof(double));
    d->constraint_lower_bound[0] = -INFINITY;
    d->constraint_upper_bound = (double *)xm(sizeof(double));
    d->constraint_upper_bound[0] = INFINITY;
    d->affine_cone_offset = zeros(1);
    d->dual_start = zeros(1);
    return d;
}
static void free_dummy(qp_problem_t *d) {
    if (!d) return;
    free(d->constraint_matrix->row_ptr); free(d->constraint_matrix->col_ind); free(d->constraint_matrix->val);
    free(d->constraint_matrix);
    free(d->constraint_lower_bound); free(d->constraint_upper_bound); free(d->affine_cone_offset); free(d->dual_start);
    free(d);
}

/* ---- state setup: reference src/solver_state.cu:280-420,688-1058 ---- */
static void init_inner(orc_solver_t *s) {
    const pdhg_parameters_t *pp = &s->params;
    int lim = pp->inner_solver_parameters.iteration_limit;
    double t0 = pp->inner_solver_parameters.initial_tolerance, tmin = pp->inner_solver_parameters.min_tolerance;
    if (lim < 1) lim = 1;
    if (t0 <= 0.0) t0 = 1e-3;
    if (tmin <= 0.0) tmin = 1e-9;
    quad_obj_type_t qt = s->w->qtype;
    s->inner_limit = lim;
    if (qt == PDHCG_NON_Q || (qt == PDHCG_DIAG_Q && !s->has_psd)) return;
    int n = s->n;
    s->bb = 1;
    s->pbuf = zeros(n);
    s->grad = zeros(n);
    s->dirn = zeros(n);
    if (s->has_cones) s->snap = zeros(n);
    s->inner_tol = t0;
    s->inner_min_tol = tmin;
    s->precond = pp->diag_jacobi_precond && !s->has_psd;
    if (s->precond) {
        s->diag_h = zeros(n);
        s->mdiag = zeros(n);
        s->minv = zeros(n);
        s->cached_inv_tau = -1.0;
        s->tol_scale = 1.0;
        if (qt == PDHCG_SPARSE_Q || qt == PDHCG_LOW_RANK_PLUS_SPARSE_Q)
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                for (int e = s->w->Q.rp[i]; e < s->w->Q.rp[i + 1]; ++e)
                    if (s->w->Q.ci[e] == i) acc += s->w->Q.v[e];
                s->diag_h[i] = acc;
            }
        if (qt == PDHCG_LOW_RANK_Q || qt == PDHCG_LOW_RANK_PLUS_SPARSE_Q) {
            /* diag(R' D R): rows of R' (reference src/kernels/pdhcg_kernels.cu:425-495) */
            const orc_csr *T = &s->Rt;
            for (int i = 0; i < n; ++i) {
                double acc = 0.0;
                if (s->w->dkind == 1) {
                    for (int e = T->rp[i]; e < T->rp[i + 1]; ++e) acc += s->w->ddiag[T->ci[e]] * T->v[e] * T->v[e];
                } else if (s->w->dkind == 2) {
                    for (int a = T->rp[i]; a < T->rp[i + 1]; ++a) {
                        const double *Dr = s->w->ddense + (size_t)T->ci[a] * s->k;
                        for (int b = T->rp[i]; b < T->rp[i + 1]; ++b) acc += T->v[a] * T->v[b] * Dr[T->ci[b]];
                    }
                } else {
                    for (int e = T->rp[i]; e < T->rp[i + 1]; ++e) acc += T->v[e] * T->v[e];
                }
                s->diag_h[i] += acc;
            }
        }
    }
}

orc_solver_t *orc_solver_create(const qp_problem_t *prob, const pdhg_parameters_t *params) {
    if (!prob) return NULL;
    if (prob->num_quadratic_constraints > 0) {
        fprintf(stderr, "[oracle] quadratic constraints are not restated.\n");
        return NULL;
    }
    orc_solver_t *s = (orc_solver_t *)calloc(1, sizeof(orc_solver_t));
    if (params) s->params = *params;
    else orc_set_default_parameters(&s->params);
    s->orig = prob;
    const qp_problem_t *work = prob;
    if (prob->num_constraints == 0 || prob->constraint_matrix == NULL) {
        s->dummy = with_dummy_row(prob);
        work = s->dummy;
    }
    s->w = orc_rescale(&s->params, work);
    orc_scaled *w = s->w;
    int n = w->n, m = w->m;
    s->n = n; s->m = m; s->k = w->k;
    csr_transpose(&w->A, &s->At);
    if (w->R.nnz > 0 || w->qtype == PDHCG_LOW_RANK_Q || w->qtype == PDHCG_LOW_RANK_PLUS_SPARSE_Q)
        csr_transpose(&w->R, &s->Rt);
    s->x0 = zeros(n); s->x = zeros(n); s->xp = zeros(n); s->xr = zeros(n);
    s->aty = zeros(n); s->dslack = zeros(n); s->dres = zeros(n); s->dx = zeros(n);
    s->y0 = zeros(m); s->y = zeros(m); s->yp = zeros(m); s->yr = zeros(m);
    s->ax = zeros(m); s->pslack = zeros(m); s->pres = zeros(m); s->dy = zeros(m);
    s->qx = zeros(n); s->rx = zeros(s->k); s->rxbuf = zeros(s->k);
    if (work->primal_start)
        for (int i = 0; i < n; ++i) s->x0[i] = s->x[i] = s->xp[i] = work->primal_start[i] * w->var_rescale[i] * w->cbr;
    if (work->dual_start)
        for (int i = 0; i < m; ++i) s->y0[i] = s->y[i] = s->yp[i] = work->dual_start[i] * w->con_rescale[i] * w->ovr;
    s->cl_fin = zeros(m); s->cu_fin = zeros(m); s->lb_fin = zeros(n); s->ub_fin = zeros(n);
    for (int i = 0; i < m; ++i) {
        s->cl_fin[i] = isfinite(w->cl[i]) ? w->cl[i] : 0.0;
        s->cu_fin[i] = isfinite(w->cu[i]) ? w->cu[i] : 0.0;
    }
    for (int j = 0; j < n; ++j) {
        s->lb_fin[j] = isfinite(w->lb[j]) ? w->lb[j] : 0.0;
        s->ub_fin[j] = isfinite(w->ub[j]) ? w->ub[j] : 0.0;
    }
    /* norms of the unscaled working problem: reference src/solver_state.cu:854-934 */
    int linf = s->params.optimality_norm == NORM_TYPE_L_INF;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) {
        double c = work->objective_vector[j];
        if (linf) { if (fabs(c) > acc) acc = fabs(c); }
        else acc += c * c;
    }
    s->obj_norm = linf ? acc : sqrt(acc);
    acc = 0.0;
    for (int i = 0; i < m; ++i) {
        double lo = work->constraint_lower_bound[i], hi = work->constraint_upper_bound[i];
        if (linf) {
            if (isfinite(lo) && lo != hi && fabs(lo) > acc) acc = fabs(lo);
            if (isfinite(hi) && fabs(hi) > acc) acc = fabs(hi);
        } else {
            if (isfinite(lo) && lo != hi) acc += lo * lo;
            if (isfinite(hi)) acc += hi * hi;
        }
    }
    for (int i = 0; i < m; ++i) {
        double g = work->affine_cone_offset[i];
        if (linf) acc = fmax(acc, fabs(g));
        else acc += g * g;
    }
    s->bound_norm = linf ? acc : sqrt(acc);
    s->best_gap = INFINITY;
    s->last_trial_fpe = INFINITY;
    /* variable cones: layout, initial auxiliary values (reference
       src/solver_state.cu:578-686) and the initial cone projection (:976-980) */
    s->cb = &work->cones;
    s->has_cones = work->cones.num_cones > 0;
    for (int k = 0; k < work->cones.num_cones; ++k) {
        s->has_power |= work->cones.type[k] == CONE_POWER;
        s->has_psd |= work->cones.type[k] == CONE_PSD;
    }
    s->acb = &work->affine_cones;
    s->has_affine = work->affine_cones.num_cones > 0 && m > 0;
    if (s->has_affine) {
        s->aff_d = zeros(m);
        for (int i = 0; i < m; ++i) s->aff_d[i] = 1.0 / w->con_rescale[i];
        s->aff_pw = zeros(work->affine_cones.num_cones);
        s->aff_rw = zeros(work->affine_cones.num_cones);
        s->aproj = zeros(m);
        s->amem = zeros(m);
        s->acomp = zeros(work->affine_cones.num_cones);
    }
    s->proj_warm = zeros(work->cones.num_cones);
    s->res_warm = zeros(work->cones.num_cones);
    s->pgp = zeros(n);
    if (s->has_cones && w->qtype != PDHCG_NON_Q) s->eff = zeros(n);
    const char *pin = work->cones.is_fixed;
    for (int k = 0; k < work->cones.num_cones; ++k) {
        int aux = work->cones.start_idx[k] + work->cones.v_dim[k];
        double *arrs[4] = {s->x0, s->x, s->xp, s->xr};
        /* fixed slots keep their start value (reference :650-651,673) */
        if (work->cones.type[k] == CONE_STANDARD_SOC) {
            double wv = -0.7071067811865475 * w->cbr * w->var_rescale[aux];
            double zv = 0.7071067811865475 * w->cbr * w->var_rescale[aux + 1];
            for (int a = 0; a < 4; ++a) {
                if (!(pin && pin[aux])) arrs[a][aux] = wv;
                if (!(pin && pin[aux + 1])) arrs[a][aux + 1] = zv;
            }
        } else if (work->cones.type[k] == CONE_ROTATED_SOC) {
            double tv = w->cbr * w->var_rescale[aux + 1];
            for (int a = 0; a < 4; ++a)
                if (!(pin && pin[aux + 1])) arrs[a][aux + 1] = tv;
        }
    }
    if (s->has_cones) project_cones(s, s->x0, s->proj_warm);
    /* initial projection onto the box: reference src/solver_state.cu:981-995 */
    for (int j = 0; j < n; ++j) s->x0[j] = fmax(w->lb[j], fmin(s->x0[j], w->ub[j]));
    memcpy(s->x, s->x0, (size_t)n * sizeof(double));
    memcpy(s->xp, s->x0, (size_t)n * sizeof(double));
    memcpy(s->xr, s->x0, (size_t)n * sizeof(double));
    /* quadratic term information: reference src/solver_state.cu:380-420 */
    if (w->qtype == PDHCG_SPARSE_Q) {
        s->q_norm = est_max_eig(&w->Q, s->params.sv_max_iter, s->params.sv_tol);
        s->nonconvexity = est_min_eig(&w->Q, s->q_norm, s->params.sv_max_iter, s->params.sv_tol);
    } else if (w->qtype == PDHCG_DIAG_Q) {
        double mx = 0.0, mn = 0.0;
        for (int j = 0; j < n; ++j) { mx = fmax(mx, fabs(w->qdiag[j])); mn = fmin(mn, w->qdiag[j]); }
        s->q_norm = mx;
        s->nonconvexity = mn;
    }
    init_inner(s);
    s->ptype = w->qtype == PDHCG_NON_Q ? P_LP : P_QP;
    if (s->nonconvexity < 0) s->inner_limit = 1;
    /* step size and primal weight: reference src/pdhg_core_op.cu:933-971 */
    if (w->A.nnz == 0) s->step_size = 1.0;
    else {
        double sv = est_max_sv(&w->A, &s->At, s->params.sv_max_iter, s->params.sv_tol);
        s->step_size = sv < 1e-12 ? 1.0 : 0.998 / sv;
    }
    s->primal_weight = s->params.bound_objective_rescaling ? 1.0 : (s->obj_norm + 1.0) / (s->bound_norm + 1.0);
    s->best_pw = s->primal_weight;
    clock_gettime(CLOCK_MONOTONIC, &s->t_start);
    return s;
}

/* ---- BB inner solve: reference src/pdhg_core_op.cu:495-669 ---- */
static void bb_prox(orc_solver_t *s, double tau) {
    int n = s->n;
    orc_scaled *w = s->w;
    double inv = 1.0 / tau, a0 = 1.0 / inv;
    int it = 1;
    if (s->precond && s->cached_inv_tau != inv) {
        double sum = 0.0;
        for (int i = 0; i < n; ++i) {
            double mm = s->diag_h[i] + inv;
            if (mm <= 0.0) mm = 1.0;
            s->mdiag[i] = mm;
            s->minv[i] = 1.0 / mm;
            sum += fabs(mm);
        }
        s->cached_inv_tau = inv;
        s->tol_scale = n > 0 ? sqrt(sum / (double)n) : 1.0;
    }
    if (s->precond) a0 *= s->tol_scale * s->tol_scale;
    obj_product(s, s->x);
#pragma omp parallel for schedule(static) if (n > 100000)
    for (int i = 0; i < n; ++i) {
        double g = s->qx[i] + w->c[i] - s->aty[i];
        double nx = s->precond ? s->x[i] - a0 * s->minv[i] * g : s->x[i] - a0 * g;
        nx = clampd(nx, w->lb[i], w->ub[i]);
        s->xp[i] = nx;
        s->grad[i] = g;
        s->dirn[i] = nx - s->x[i];
    }
    if (s->has_cones) { /* reference :561-566 */
        project_cones(s, s->xp, s->proj_warm);
        for (int i = 0; i < n; ++i) s->dirn[i] = s->xp[i] - s->x[i];
    }
    double stms = 0.0, nrm = 0.0;
    while (it < s->inner_limit) {
        if (s->precond) {
            stms = 0.0;
            for (int i = 0; i < n; ++i) stms += s->dirn[i] * (s->mdiag[i] * s->dirn[i]);
            nrm = sqrt(stms);
        } else {
            nrm = sqrt(dot(n, s->dirn, s->dirn));
        }
        if (nrm <= s->inner_tol) break;
        obj_product(s, s->xp);
#pragma omp parallel for schedule(static) if (n > 100000)
        for (int i = 0; i < n; ++i) {
            double last = s->grad[i];
            double cur = s->qx[i] + w->c[i] - s->aty[i] + inv * (s->xp[i] - s->x[i]);
            s->pbuf[i] = cur - last;
            s->grad[i] = cur;
        }
        double tmp = dot(n, s->dirn, s->pbuf);
        if (s->snap) memcpy(s->snap, s->xp, (size_t)n * sizeof(double));
        double alpha = s->precond ? stms / tmp : (nrm * nrm) / tmp;
#pragma omp parallel for schedule(static) if (n > 100000)
        for (int i = 0; i < n; ++i) {
            double cur = s->xp[i];
            double nx = s->precond ? cur - alpha * s->minv[i] * s->grad[i] : cur - alpha * s->grad[i];
            nx = clampd(nx, w->lb[i], w->ub[i]);
            s->dirn[i] = nx - cur;
            s->xp[i] = nx;
        }
        if (s->snap) { /* reference :652-657 */
            project_cones(s, s->xp, s->proj_warm);
            for (int i = 0; i < n; ++i) s->dirn[i] = s->xp[i] - s->snap[i];
        }
        it++;
    }
    for (int i = 0; i < n; ++i) s->xr[i] = 2.0 * s->xp[i] - s->x[i];
    s->inner_total += it - 1;
}

/* ---- one PDHG step: reference src/pdhg_core_op.cu:705-830 ---- */
static void pdhg_step(orc_solver_t *s) {
    orc_scaled *w = s->w;
    int n = s->n, m = s->m;
    double tau = s->step_size / s->primal_weight;
    if (s->nonconvexity < 0) {
        tau = fmax(tau, -1.01 * fmin(0.0, s->nonconvexity));
        tau /= 100;
    }
    double sigma = s->step_size * s->primal_weight;
    spmv(&s->At, s->y, s->aty);
    int tc = s->total_count;
    int store = s->is_major || ((tc + 2) % print_freq(tc + 2)) == 0;
    switch (w->qtype) {
    case PDHCG_NON_Q: {
        int store_x = store || s->has_cones; /* cones force the stored form (:428) */
        for (int i = 0; i < n; ++i) {
            double xi = s->x[i];
            double t = xi - tau * (w->c[i] - s->aty[i]);
            double pj = clampd(t, w->lb[i], w->ub[i]);
            s->xr[i] = 2.0 * pj - xi;
            if (store_x) {
                s->xp[i] = pj;
                s->dslack[i] = (pj - t) / tau;
            }
        }
        if (s->has_cones) { /* reference :730-735 */
            project_cones(s, s->xp, s->proj_warm);
            cone_reflection(s);
        }
        break;
    }
    case PDHCG_DIAG_Q:
        if (s->has_psd) { /* reference :740-741 */
            bb_prox(s, tau);
            break;
        }
        for (int i = 0; i < n; ++i) {
            double xi = s->x[i];
            double t = (xi - tau * (w->c[i] - s->aty[i])) / (1.0 + tau * w->qdiag[i]);
            double pj = clampd(t, w->lb[i], w->ub[i]);
            s->xr[i] = 2.0 * pj - xi;
            if (store || s->has_cones) s->xp[i] = pj; /* cones force the stored form (reference :460-463) */
        }
        if (s->has_cones) { /* reference :742-746 */
            project_cones_diag_q(s, s->xp, tau);
            cone_reflection(s);
        }
        break;
    default:
        bb_prox(s, tau);
        break;
    }
    s->inner_total++;
    spmv(&w->A, s->xr, s->ax);
    if (m == 0) return;
    if (s->has_affine) { /* reference src/pdhg_core_op.cu:809-829, kernels :252-281 */
        for (int i = 0; i < m; ++i) s->aproj[i] = clampd(s->ax[i] + w->g[i] - s->y[i] / sigma, w->cl[i], w->cu[i]);
        project_blocks(s->acb, s->aff_d, s->aproj, s->aff_pw);
    }
    for (int i = 0; i < m; ++i) {
        double cur = s->y[i];
        double v = s->ax[i] + w->g[i];
        double pj = s->has_affine ? s->aproj[i] : clampd(v - cur / sigma, w->cl[i], w->cu[i]);
        double nx = cur - sigma * v + sigma * pj;
        if (store) s->yp[i] = nx;
        s->yr[i] = 2.0 * nx - cur;
    }
}

/* reference src/pdhg_core_op.cu:832-845 */
static void halpern(orc_solver_t *s, double rho) {
    double wgt = (double)(s->inner_count + 1) / (s->inner_count + 2);
    for (int i = 0; i < s->n; ++i) {
        double r = rho * s->xr[i] + (1.0 - rho) * s->x[i];
        s->x[i] = wgt * r + (1.0 - wgt) * s->x0[i];
    }
    for (int i = 0; i < s->m; ++i) {
        double r = rho * s->yr[i] + (1.0 - rho) * s->y[i];
        s->y[i] = wgt * r + (1.0 - wgt) * s->y0[i];
    }
}

/* reference src/pdhg_core_op.cu:973-1035 */
static void fixed_point_error(orc_solver_t *s) {
    int n = s->n, m = s->m;
    for (int i = 0; i < n; ++i) s->dx[i] = s->xr[i] - s->x[i];
    for (int i = 0; i < m; ++i) s->dy[i] = s->yr[i] - s->y[i];
    spmv(&s->At, s->dy, s->aty);
    double dn = nrm2(m, s->dy), pn = nrm2(n, s->dx);
    dn = sqrt(dn * dn);
    pn = sqrt(pn * pn);
    double mov = pn * pn * s->primal_weight + dn * dn / s->primal_weight;
    double cross = dot(n, s->aty, s->dx);
    s->fpe = sqrt(mov + 2 * s->step_size * cross);
    if (s->ptype == P_QP && s->bb)
        s->inner_tol = fmin(s->inner_tol, fmax(0.0005 * pn / s->step_size * s->primal_weight, s->inner_min_tol));
}

/* reference src/pdhg_core_op.cu:859-931 */
static void restart(orc_solver_t *s) {
    int n = s->n, m = s->m;
    const restart_parameters_t *rp = &s->params.restart_params;
    for (int i = 0; i < n; ++i) s->dx[i] = s->xp[i] - s->x0[i];
    for (int i = 0; i < m; ++i) s->dy[i] = s->yp[i] - s->y0[i];
    double pd = nrm2(n, s->dx), dd = nrm2(m, s->dy);
    pd = sqrt(pd * pd);
    dd = sqrt(dd * dd);
    double ratio = s->rel_du / s->rel_pr;
    if (pd > 1e-16 && dd > 1e-16 && pd < 1e12 && dd < 1e12 && ratio > 1e-8 && ratio < 1e8) {
        double err = log(dd) - log(pd) - log(s->primal_weight);
        s->pw_err_sum *= rp->i_smooth;
        s->pw_err_sum += err;
        double derr = err - s->pw_last_err;
        s->primal_weight *= exp(rp->k_p * err + rp->k_i * s->pw_err_sum + rp->k_d * derr);
        s->pw_last_err = err;
    } else {
        s->primal_weight = s->best_pw;
        s->pw_err_sum = 0.0;
        s->pw_last_err = 0.0;
    }
    double g = fabs(log10(s->rel_du / s->rel_pr));
    if (g < s->best_gap) {
        s->best_gap = g;
        s->best_pw = s->primal_weight;
    }
    memcpy(s->x0, s->xp, (size_t)n * sizeof(double));
    memcpy(s->x, s->xp, (size_t)n * sizeof(double));
    memcpy(s->y0, s->yp, (size_t)m * sizeof(double));
    memcpy(s->y, s->yp, (size_t)m * sizeof(double));
    s->inner_count = 0;
    s->last_trial_fpe = INFINITY;
}

/* reference src/utils.cu:291-322 */
int orc_restart_decision(const restart_parameters_t *rp, int ef, int total_count, int inner_count, double fpe,
                         double init_fpe, double last_trial_fpe) {
    int r = 0;
    if (total_count == ef) r = 1;
    else if (total_count > ef) {
        if (fpe <= rp->sufficient_reduction_for_restart * init_fpe) r = 1;
        if (fpe <= rp->necessary_reduction_for_restart * init_fpe && fpe > last_trial_fpe) r = 1;
        if (inner_count >= rp->artificial_restart_threshold * total_count) r = 1;
    }
    return r;
}

static int should_restart(orc_solver_t *s) {
    int r = orc_restart_decision(&s->params.restart_params, s->params.termination_evaluation_frequency,
                                 s->total_count, s->inner_count, s->fpe, s->init_fpe, s->last_trial_fpe);
    s->last_trial_fpe = s->fpe;
    return r;
}

/* ---- KKT residuals: reference src/pdhg_core_op.cu:1037-1294, kernels :638-752 ---- */
static void residual(orc_solver_t *s) {
    orc_scaled *w = s->w;
    int n = s->n, m = s->m;
    int linf = s->params.optimality_norm == NORM_TYPE_L_INF;
    double tau = s->step_size / s->primal_weight;
    spmv(&w->A, s->xp, s->ax);
    spmv(&s->At, s->yp, s->aty);
    obj_product(s, s->xp);
    for (int i = 0; i < m; ++i) {
        double v = s->ax[i] + w->g[i];
        double pj = clampd(v, w->cl[i], w->cu[i]);
        s->pres[i] = (v - pj) * w->con_rescale[i];
        s->pslack[i] = fmax(s->yp[i], 0.0) * s->cl_fin[i] + fmin(s->yp[i], 0.0) * s->cu_fin[i] - w->g[i] * s->yp[i];
    }
    if (s->ptype == P_LP) {
        for (int j = 0; j < n; ++j) s->dres[j] = (w->c[j] - s->aty[j] - s->dslack[j]) * w->var_rescale[j];
    } else {
        for (int j = 0; j < n; ++j) {
            double gr = s->qx[j] + w->c[j] - s->aty[j];
            double t = s->xp[j] - tau * gr;
            double pj = fmax(w->lb[j], fmin(w->ub[j], t));
            double ds = (pj - t) / tau;
            s->dres[j] = (gr - ds) * w->var_rescale[j];
            s->dslack[j] = ds;
        }
    }
    const double *eff = s->eff ? s->eff : w->c;
    if (s->has_cones) {
        /* dual-cone residual of c_eff - A'y on the cone slots (reference
           src/cone_dispatch.cu:88-115) */
        const cone_blocks_t *b = s->cb;
        double *rc = s->pgp;
        for (int k = 0; k < b->num_cones; ++k) {
            int st = b->start_idx[k], vd = b->v_dim[k], L = orc_cone_block_length(b, k);
            for (int i = st; i < st + L; ++i) rc[i] = eff[i] - s->aty[i];
            const double *d = w->var_rescale + st;
            if (b->is_fixed && orc_section_has_fixed(b->is_fixed + st, L)) {
                /* prox residual of the section (reference
                   src/kernels/pdhcg_soc_cone_kernels.cu:489-507) */
                orc_section_dual_residual(b->type[k], s->xp + st, rc + st, vd, d, b->is_fixed + st,
                                          b->power_alpha ? b->power_alpha[k] : 0.0, s->dres + st);
                continue;
            }
            switch (b->type[k]) {
            case CONE_STANDARD_SOC: orc_soc_dual_residual(rc + st, vd, d, s->res_warm + k, s->dres + st); break;
            case CONE_ROTATED_SOC: orc_rsoc_dual_residual(rc + st, vd, d, s->res_warm + k, s->dres + st); break;
            case CONE_EXPONENTIAL: orc_exp_dual_residual(rc + st, d, s->dres + st); break;
            case CONE_POWER: orc_power_dual_residual(rc + st, d, b->power_alpha[k], s->dres + st); break;
            case CONE_PSD: orc_psd_dual_residual(rc + st, vd, d, s->dres + st); break;
            default: break;
            }
        }
        /* projected-gradient mapping (reference src/pdhg_core_op.cu:67-90,
           kernels :75-109) */
        double st = s->step_size / s->primal_weight;
        if (!(st > 0.0) || !isfinite(st)) st = 1.0;
        for (int j = 0; j < n; ++j) s->pgp[j] = fmax(w->lb[j], fmin(s->xp[j] - st * (eff[j] - s->aty[j]), w->ub[j]));
        project_cones(s, s->pgp, s->res_warm);
        for (int j = 0; j < n; ++j) {
            double r = (s->xp[j] - s->pgp[j]) / st * w->var_rescale[j];
            if (!isfinite(r)) r = copysign(INFINITY, r);
            if (fabs(r) > fabs(s->dres[j])) s->dres[j] = r;
        }
    }
    double amem_norm = 0.0, acomp_norm = 0.0;
    if (s->has_affine) {
        /* primal residual against the affine cones and dual-cone membership of y
           (reference src/pdhg_core_op.cu:1125-1129,114-201, kernels
           pdhcg_affine_cone_kernels.cu:23-75, pdhcg_kernels.cu:660-671) */
        const cone_blocks_t *b = s->acb;
        for (int i = 0; i < m; ++i) {
            s->pres[i] = clampd(s->ax[i] + w->g[i], w->cl[i], w->cu[i]);
            s->amem[i] = 0.0;
        }
        project_blocks(b, s->aff_d, s->pres, s->aff_rw);
        for (int k = 0; k < b->num_cones; ++k) {
            int st = b->start_idx[k], L = orc_cone_block_length(b, k);
            double dotv = 0.0;
            for (int i = st; i < st + L; ++i) {
                s->amem[i] = -s->yp[i];
                dotv += s->yp[i] * (s->ax[i] + w->g[i]);
            }
            s->acomp[k] = fabs(dotv) / w->cbr;
        }
        project_blocks(b, s->aff_d, s->amem, s->aff_rw);
        for (int i = 0; i < m; ++i) {
            s->pres[i] = (s->ax[i] + w->g[i] - s->pres[i]) * w->con_rescale[i];
            s->amem[i] *= s->aff_d[i];
        }
        amem_norm = linf ? inf_norm(m, s->amem) : nrm2(m, s->amem);
        acomp_norm = linf ? inf_norm(b->num_cones, s->acomp) : nrm2(b->num_cones, s->acomp);
    }
    s->abs_pr = (linf ? inf_norm(m, s->pres) : nrm2(m, s->pres)) / w->cbr;
    double lin_pr = s->abs_pr;
    double pv_abs = 0.0, pv_rel = 0.0;
    if (s->has_power) { /* reference src/pdhg_core_op.cu:214-270 */
        const cone_blocks_t *b = s->cb;
        for (int k = 0; k < b->num_cones; ++k) {
            if (b->type[k] != CONE_POWER) continue;
            int st = b->start_idx[k];
            double a, r;
            orc_power_violation(s->xp + st, w->var_rescale + st, b->power_alpha[k], w->cbr, &a, &r);
            if (linf) { pv_abs = fmax(pv_abs, a); pv_rel = fmax(pv_rel, r); }
            else { pv_abs += a * a; pv_rel += r * r; }
        }
        if (!linf) { pv_abs = sqrt(pv_abs); pv_rel = sqrt(pv_rel); }
        pv_abs /= w->cbr;
        s->abs_pr = linf ? fmax(s->abs_pr, pv_abs) : hypot(s->abs_pr, pv_abs);
    }
    s->abs_du = linf ? inf_norm(n, s->dres) : nrm2(n, s->dres);
    if (s->has_affine) /* reference src/pdhg_core_op.cu:1166-1189 */
        s->abs_du = linf ? fmax(fmax(s->abs_du, amem_norm), acomp_norm)
                         : sqrt(s->abs_du * s->abs_du + amem_norm * amem_norm + acomp_norm * acomp_norm);
    s->abs_du /= w->ovr;
    double half_xqx = w->qtype == PDHCG_NON_Q ? 0.0 : 0.5 * dot(n, s->xp, s->qx);
    s->pobj = (dot(n, w->c, s->xp) + half_xqx) / (w->cbr * w->ovr) + w->obj_const;
    if (s->has_cones) { /* cone dual slack (reference src/cone_dispatch.cu:137-154) */
        const cone_blocks_t *b = s->cb;
        for (int k = 0; k < b->num_cones; ++k) {
            int st = b->start_idx[k], L = orc_cone_block_length(b, k);
            for (int i = st; i < st + L; ++i) s->dslack[i] = eff[i] - s->aty[i];
        }
    }
    double base = dot(n, s->dslack, s->xp);
    double sps = 0.0;
    for (int i = 0; i < m; ++i) sps += s->pslack[i];
    s->dobj = (base + sps - half_xqx) / (w->cbr * w->ovr) + w->obj_const;
    s->rel_pr = lin_pr / (1.0 + s->bound_norm);
    if (s->has_power) s->rel_pr = linf ? fmax(s->rel_pr, pv_rel) : hypot(s->rel_pr, pv_rel);
    double den;
    if (s->ptype == P_LP) den = 1.0 + s->obj_norm;
    else {
        for (int j = 0; j < n; ++j) {
            s->aty[j] *= w->var_rescale[j];
            s->qx[j] *= w->var_rescale[j];
        }
        double qn = linf ? inf_norm(n, s->qx) : nrm2(n, s->qx);
        double an = linf ? inf_norm(n, s->aty) : nrm2(n, s->aty);
        den = 1.0 + fmax(s->obj_norm, fmax(qn / w->ovr, an / w->ovr));
    }
    s->rel_du = s->abs_du / den;
    s->gap = fabs(s->pobj - s->dobj);
    s->rel_gap = s->gap / (1.0 + fabs(s->pobj) + fabs(s->dobj));
}

/* reference src/pdhg_core_op.cu:1296-1429 */
static void infeasibility(orc_solver_t *s) {
    orc_scaled *w = s->w;
    int n = s->n, m = s->m;
    for (int j = 0; j < n; ++j) {
        if (isfinite(w->lb[j])) s->dx[j] = fmax(s->dx[j], 0.0);
        if (isfinite(w->ub[j])) s->dx[j] = fmin(s->dx[j], 0.0);
    }
    for (int i = 0; i < m; ++i) {
        if (!isfinite(w->cl[i])) s->dy[i] = fmin(s->dy[i], 0.0);
        if (!isfinite(w->cu[i])) s->dy[i] = fmax(s->dy[i], 0.0);
    }
    double prn = inf_norm(n, s->dx);
    if (prn > 0.0) {
        double sc = 1.0 / prn;
        for (int j = 0; j < n; ++j) s->dx[j] *= sc;
    }
    double drn = inf_norm(m, s->dy);
    spmv(&w->A, s->dx, s->ax);
    spmv(&s->At, s->dy, s->aty);
    s->pray_lin_obj = dot(n, w->c, s->dx) / (w->cbr * w->ovr);
    for (int i = 0; i < m; ++i)
        s->pslack[i] = fmax(s->dy[i], 0.0) * s->cl_fin[i] + fmin(s->dy[i], 0.0) * s->cu_fin[i] - w->g[i] * s->dy[i];
    for (int j = 0; j < n; ++j)
        s->dslack[j] = fmax(-s->aty[j], 0.0) * s->lb_fin[j] + fmin(-s->aty[j], 0.0) * s->ub_fin[j];
    double sps = 0.0, sds = 0.0;
    for (int i = 0; i < m; ++i) sps += s->pslack[i];
    for (int j = 0; j < n; ++j) sds += s->dslack[j];
    s->dray_obj = (sps + sds) / (w->cbr * w->ovr);
    for (int i = 0; i < m; ++i) {
        double pv = s->ax[i];
        s->pslack[i] = (fmax(0.0, -pv) * (double)isfinite(w->cl[i]) + fmax(0.0, pv) * (double)isfinite(w->cu[i])) *
                       w->con_rescale[i];
    }
    for (int j = 0; j < n; ++j) {
        double dv = -s->aty[j];
        s->dslack[j] = (fmax(0.0, dv) * (double)!isfinite(w->lb[j]) - fmin(0.0, dv) * (double)!isfinite(w->ub[j])) *
                       w->var_rescale[j];
    }
    s->max_pray = inf_norm(m, s->pslack);
    if (s->ptype != P_LP && w->qtype != PDHCG_NON_Q) {
        obj_product(s, s->dx);
        double qn = inf_norm(n, s->qx);
        s->max_pray = fmax(s->max_pray, qn / w->ovr);
    }
    double dsn = inf_norm(n, s->dslack);
    s->max_dray = dsn;
    double sc = fmax(drn, dsn);
    if (sc > 0.0) {
        s->max_dray /= sc;
        s->dray_obj /= sc;
    } else {
        s->max_dray = 0.0;
        s->dray_obj = 0.0;
    }
}

/* reference src/utils.cu:219-289 */
static void check_termination(orc_solver_t *s) {
    const termination_criteria_t *c = &s->params.termination_criteria;
    if (s->rel_du < c->eps_feasible_relative && s->rel_pr < c->eps_feasible_relative &&
        s->rel_gap < c->eps_optimal_relative) {
        s->reason = TERMINATION_REASON_OPTIMAL;
        return;
    }
    if (s->has_cones) { /* ray certificates are box-only (reference src/utils.cu:263-278) */
        if (s->total_count >= c->iteration_limit) s->reason = TERMINATION_REASON_ITERATION_LIMIT;
        else if (s->cum_time >= c->time_sec_limit) s->reason = TERMINATION_REASON_TIME_LIMIT;
        return;
    }
    if (s->dray_obj > 0.0 && s->max_dray / s->dray_obj <= c->eps_infeasible) {
        s->reason = TERMINATION_REASON_PRIMAL_INFEASIBLE;
        return;
    }
    if (s->pray_lin_obj < 0.0 && s->max_pray / (-s->pray_lin_obj) <= c->eps_infeasible) {
        s->reason = TERMINATION_REASON_DUAL_INFEASIBLE;
        return;
    }
    if (s->total_count >= c->iteration_limit) {
        s->reason = TERMINATION_REASON_ITERATION_LIMIT;
        return;
    }
    if (s->cum_time >= c->time_sec_limit) s->reason = TERMINATION_REASON_TIME_LIMIT;
}

static double elapsed(const orc_solver_t *s) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - s->t_start.tv_sec) + 1e-9 * (double)(t.tv_nsec - s->t_start.tv_nsec);
}

/* Body of the main loop, reference src/solver.cu:121-169. */
int orc_solver_run(orc_solver_t *s, int iters) {
    const pdhg_parameters_t *p = &s->params;
    int ef = p->termination_evaluation_frequency;
    for (int done = 0; done < iters; ++done) {
        if (s->reason != TERMINATION_REASON_UNSPECIFIED) break;
        if (!(s->total_count < p->termination_criteria.iteration_limit)) {
            s->reason = TERMINATION_REASON_ITERATION_LIMIT;
            residual(s);
            break;
        }
        int tc = s->total_count;
        if (s->is_major || tc == 0 || tc % print_freq(tc) == 0) {
            residual(s);
            if (!s->has_cones && !s->has_affine && s->is_major && tc < 3 * ef) infeasibility(s);
            s->cum_time = elapsed(s);
            check_termination(s);
            if (s->reason != TERMINATION_REASON_UNSPECIFIED) break;
        }
        if (s->is_major || tc == 0) {
            s->do_restart = should_restart(s);
            if (s->do_restart) restart(s);
        }
        s->is_major = ((tc + 1) % ef) == 0;
        pdhg_step(s);
        if (s->is_major || s->do_restart) {
            fixed_point_error(s);
            if (s->do_restart) {
                s->init_fpe = s->fpe;
                s->do_restart = 0;
            }
        }
        halpern(s, p->reflection_coefficient);
        s->inner_count++;
        s->total_count++;
    }
    return (int)s->reason;
}

int orc_solver_total_count(const orc_solver_t *s) { return s->total_count; }
int orc_solver_inner_count(const orc_solver_t *s) { return s->inner_total; }

void orc_solver_scalings(const orc_solver_t *s, double *cr, double *vr, double *cbr, double *ovr, double *step,
                         double *pw) {
    if (cr) memcpy(cr, s->w->con_rescale, (size_t)s->m * sizeof(double));
    if (vr) memcpy(vr, s->w->var_rescale, (size_t)s->n * sizeof(double));
    if (cbr) *cbr = s->w->cbr;
    if (ovr) *ovr = s->w->ovr;
    if (step) *step = s->step_size;
    if (pw) *pw = s->primal_weight;
}

void orc_solver_vectors(const orc_solver_t *s, double *xp, double *x0, double *dx, double *yp, double *y0) {
    if (xp) memcpy(xp, s->xp, (size_t)s->n * sizeof(double));
    if (x0) memcpy(x0, s->x0, (size_t)s->n * sizeof(double));
    if (dx) memcpy(dx, s->dx, (size_t)s->n * sizeof(double));
    if (yp) memcpy(yp, s->yp, (size_t)s->m * sizeof(double));
    if (y0) memcpy(y0, s->y0, (size_t)s->m * sizeof(double));
}

void orc_solver_loop_state(const orc_solver_t *s, double *out) {
    out[0] = s->inner_tol;
    out[1] = s->fpe;
    out[2] = s->rel_pr;
    out[3] = s->rel_du;
    out[4] = s->tol_scale;
}

/* reference src/pdhg_core_op.cu:1431-1496 */
static pdhcg_result_t *make_result(orc_solver_t *s) {
    orc_scaled *w = s->w;
    int n = s->n, m = s->m;
    pdhcg_result_t *r = (pdhcg_result_t *)calloc(1, sizeof(pdhcg_result_t));
    spmv(&s->At, s->yp, s->aty);
    obj_product(s, s->xp);
    r->primal_solution = (double *)xm((size_t)n * sizeof(double));
    r->dual_solution = (double *)xm((size_t)m * sizeof(double));
    r->reduced_cost = (double *)xm((size_t)n * sizeof(double));
    for (int j = 0; j < n; ++j) {
        double gr = w->c[j];
        if (w->qtype != PDHCG_NON_Q) gr += s->qx[j];
        double rc = (gr - s->aty[j]) * w->var_rescale[j] / w->ovr;
        if (!isfinite(w->lb[j])) rc = fmin(rc, 0.0);
        if (!isfinite(w->ub[j])) rc = fmax(rc, 0.0);
        r->reduced_cost[j] = rc;
        r->primal_solution[j] = s->xp[j] / w->var_rescale[j] / w->cbr;
    }
    for (int i = 0; i < m; ++i) r->dual_solution[i] = s->yp[i] / w->con_rescale[i] / w->ovr;
    r->num_variables = s->orig->num_variables;
    r->num_constraints = s->orig->num_constraints;
    r->num_nonzeros = s->orig->constraint_matrix_num_nonzeros;
    r->total_count = s->total_count;
    r->total_inner_count = s->inner_total;
    r->rescaling_time_sec = w->rescale_time;
    r->cumulative_time_sec = s->cum_time;
    r->relative_primal_residual = s->rel_pr;
    r->relative_dual_residual = s->rel_du;
    r->absolute_primal_residual = s->abs_pr;
    r->absolute_dual_residual = s->abs_du;
    r->primal_objective_value = s->pobj;
    r->dual_objective_value = s->dobj;
    r->objective_gap = s->gap;
    r->relative_objective_gap = s->rel_gap;
    r->max_primal_ray_infeasibility = s->max_pray;
    r->max_dual_ray_infeasibility = s->max_dray;
    r->primal_ray_linear_objective = s->pray_lin_obj;
    r->dual_ray_objective = s->dray_obj;
    r->termination_reason = s->reason;
    return r;
}

void orc_solver_free(orc_solver_t *s) {
    if (!s) return;
    orc_scaled_free(s->w);
    orc_csr_free(&s->At);
    orc_csr_free(&s->Rt);
    free_dummy(s->dummy);
    double *arr[] = {s->x0, s->x, s->xp, s->xr, s->y0, s->y, s->yp, s->yr, s->aty, s->ax, s->dslack, s->pslack,
                     s->pres, s->dres, s->dx, s->dy, s->qx, s->rx, s->rxbuf, s->cl_fin, s->cu_fin, s->lb_fin,
                     s->ub_fin, s->grad, s->dirn, s->pbuf, s->diag_h, s->mdiag, s->minv,
                     s->proj_warm, s->res_warm, s->eff, s->snap, s->pgp, s->aff_d, s->aff_pw, s->aff_rw,
                     s->aproj, s->amem, s->acomp};
    for (size_t i = 0; i < sizeof(arr) / sizeof(arr[0]); ++i) free(arr[i]);
    free(s);
}

pdhcg_result_t *orc_solve_qp_problem(const qp_problem_t *prob, const pdhg_parameters_t *params) {
    if (!prob) return NULL;
    if (orc_validate_fixed_cone_sections(prob) != 0) return NULL;
    orc_solver_t *s = orc_solver_create(prob, params);
    if (!s) return NULL;
    while (s->reason == TERMINATION_REASON_UNSPECIFIED) {
        if (!(s->total_count < s->params.termination_criteria.iteration_limit)) {
            s->reason = TERMINATION_REASON_ITERATION_LIMIT;
            residual(s);
            break;
        }
        orc_solver_run(s, 1 << 30);
    }
    pdhcg_result_t *r = make_result(s);
    orc_solver_free(s);
    return r;
}

void orc_result_free(pdhcg_result_t *r) {
    if (!r) return;
    free(r->primal_solution); free(r->dual_solution); free(r->reduced_cost);
    free(r);
}
