/*
 * ORACLE (test infrastructure, see oracle.h): diagonal problem scaling,
 * restated from reference src/preconditioner.c.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define SC_EPS 1e-12
#define LOG_MIN_ABS 1e-300
#define LOG_LIMIT 69.07755278982137
#define PSD_LOG_LIMIT (0.5 * LOG_LIMIT)
#define TAPER 8

static void *xm(size_t n) { void *p = malloc(n ? n : 1); if (!p) abort(); return p; }

void orc_csr_free(orc_csr *a) {
    free(a->rp); free(a->ci); free(a->v);
    memset(a, 0, sizeof(*a));
}

static void csr_copy(orc_csr *d, const CsrComponent *s, int rows, int cols, int nnz) {
    d->rows = rows; d->cols = cols; d->nnz = nnz;
    d->rp = (int *)xm((size_t)(rows + 1) * sizeof(int));
    d->ci = (int *)xm((size_t)nnz * sizeof(int));
    d->v = (double *)xm((size_t)nnz * sizeof(double));
    if (s && s->row_ptr) memcpy(d->rp, s->row_ptr, (size_t)(rows + 1) * sizeof(int));
    else memset(d->rp, 0, (size_t)(rows + 1) * sizeof(int));
    if (nnz > 0) {
        memcpy(d->ci, s->col_ind, (size_t)nnz * sizeof(int));
        memcpy(d->v, s->val, (size_t)nnz * sizeof(double));
    }
}

static double *dup(const double *s, int n) {
    double *d = (double *)xm((size_t)n * sizeof(double));
    memcpy(d, s, (size_t)n * sizeof(double));
    return d;
}

/* reference src/preconditioner.c:47-99 */
static void psd_log_project(double *ls, int start, int order, double *f) {
    memset(f, 0, (size_t)order * sizeof(double));
    int slot = start;
    for (int c = 0; c < order; ++c)
        for (int r = c; r < order; ++r, ++slot) {
            double v = ls[slot];
            if (isnan(v)) v = 0.0;
            else if (v > LOG_LIMIT) v = LOG_LIMIT;
            else if (v < -LOG_LIMIT) v = -LOG_LIMIT;
            if (r == c) f[r] += 2.0 * v;
            else { f[r] += v; f[c] += v; }
        }
    double sum = 0.0;
    for (int i = 0; i < order; ++i) sum += f[i];
    double corr = sum / (2.0 * order + 2.0), diag = order + 2.0;
    for (int i = 0; i < order; ++i) {
        f[i] = (f[i] - corr) / diag;
        f[i] = fmax(-PSD_LOG_LIMIT, fmin(PSD_LOG_LIMIT, f[i]));
    }
    slot = start;
    for (int c = 0; c < order; ++c)
        for (int r = c; r < order; ++r, ++slot) ls[slot] = f[r] + f[c];
}

static int max_psd_order(const cone_blocks_t *b) {
    int mx = 0;
    for (int k = 0; k < b->num_cones; ++k)
        if (b->type[k] == CONE_PSD && b->v_dim[k] > mx) mx = b->v_dim[k];
    return mx;
}

/* reference src/preconditioner.c:108-144; phase 0 = Ruiz, 1 = Pock-Chambolle */
static void cone_adjust(double *s, const cone_blocks_t *b, int phase, bool preserve) {
    int wo = preserve ? 0 : max_psd_order(b);
    double *f = wo > 0 ? (double *)xm((size_t)wo * sizeof(double)) : NULL;
    for (int k = 0; k < b->num_cones; ++k) {
        if (!preserve) {
            if (b->type[k] == CONE_PSD) {
                int st = b->start_idx[k], L = orc_cone_block_length(b, k);
                for (int i = st; i < st + L; ++i) s[i] = log(fmax(s[i], LOG_MIN_ABS));
                psd_log_project(s, st, b->v_dim[k], f);
                for (int i = st; i < st + L; ++i) s[i] = exp(s[i]);
            }
            continue;
        }
        int st = b->start_idx[k], L = orc_cone_block_length(b, k);
        double mx = 0.0, sq = 0.0;
        for (int i = st; i < st + L; ++i) { mx = fmax(mx, s[i]); sq += s[i] * s[i]; }
        double rms = sqrt(sq / (double)L);
        double bs = phase == 0 ? (L <= TAPER ? mx : rms) : (L <= TAPER ? rms : sqrt(mx * rms));
        for (int i = st; i < st + L; ++i) s[i] = bs;
    }
    free(f);
}

/* reference src/preconditioner.c:256-309 */
static void scale_by(orc_scaled *w, const double *cr, const double *vr) {
    for (int i = 0; i < w->n; ++i) { w->c[i] /= vr[i]; w->ub[i] *= vr[i]; w->lb[i] *= vr[i]; }
    for (int i = 0; i < w->m; ++i) { w->cl[i] /= cr[i]; w->cu[i] /= cr[i]; w->g[i] /= cr[i]; }
    for (int r = 0; r < w->m; ++r)
        for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) w->A.v[e] /= (cr[r] * vr[w->A.ci[e]]);
    for (int r = 0; r < w->n; ++r)
        for (int e = w->Q.rp[r]; e < w->Q.rp[r + 1]; ++e) w->Q.v[e] /= (vr[r] * vr[w->Q.ci[e]]);
    if (w->R.nnz > 0)
        for (int r = 0; r < w->k; ++r)
            for (int e = w->R.rp[r]; e < w->R.rp[r + 1]; ++e) w->R.v[e] /= vr[w->R.ci[e]];
}

/* reference src/preconditioner.c:334-495 */
static void curtis_reid(orc_scaled *w, const qp_problem_t *src, int iters, bool preserve, double *cum_c,
                        double *cum_v) {
    int m = w->m, n = w->n;
    double *cr = (double *)xm((size_t)m * sizeof(double)), *vr = (double *)xm((size_t)n * sizeof(double));
    for (int i = 0; i < m; ++i) cr[i] = 1.0;
    for (int j = 0; j < n; ++j) vr[j] = 1.0;
    int wo = 0;
    if (!preserve) {
        wo = max_psd_order(&src->cones);
        int a = max_psd_order(&src->affine_cones);
        if (a > wo) wo = a;
    }
    double *f = wo > 0 ? (double *)xm((size_t)wo * sizeof(double)) : NULL;
    if (iters > 0 && m > 0 && n > 0 && w->A.nnz > 0) {
        double *rl = (double *)calloc((size_t)m, sizeof(double)), *cl = (double *)calloc((size_t)n, sizeof(double));
        double *rs = (double *)calloc((size_t)m, sizeof(double)), *cs = (double *)calloc((size_t)n, sizeof(double));
        double *csum = (double *)calloc((size_t)n, sizeof(double));
        int *cc = (int *)calloc((size_t)n, sizeof(int));
        for (int r = 0; r < m; ++r)
            for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) {
                double la = log(fmax(fabs(w->A.v[e]), LOG_MIN_ABS));
                rs[r] += la; cs[w->A.ci[e]] += la; cc[w->A.ci[e]]++;
            }
        for (int it = 0; it < iters; ++it) {
            for (int r = 0; r < m; ++r) {
                int b = w->A.rp[r], e2 = w->A.rp[r + 1];
                double s = rs[r];
                for (int e = b; e < e2; ++e) s -= cl[w->A.ci[e]];
                rl[r] = e2 > b ? s / (double)(e2 - b) : 0.0;
            }
            for (int k = 0; k < src->affine_cones.num_cones; ++k) {
                if (!preserve) {
                    if (src->affine_cones.type[k] == CONE_PSD)
                        psd_log_project(rl, src->affine_cones.start_idx[k], src->affine_cones.v_dim[k], f);
                    continue;
                }
                int st = src->affine_cones.start_idx[k], L = orc_cone_block_length(&src->affine_cones, k);
                double bsum = 0.0;
                int bc = 0;
                for (int r = st; r < st + L; ++r) {
                    bsum += rs[r];
                    bc += w->A.rp[r + 1] - w->A.rp[r];
                    for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) bsum -= cl[w->A.ci[e]];
                }
                double bl = bc > 0 ? bsum / (double)bc : 0.0;
                for (int r = st; r < st + L; ++r) rl[r] = bl;
            }
            memcpy(csum, cs, (size_t)n * sizeof(double));
            for (int r = 0; r < m; ++r)
                for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) csum[w->A.ci[e]] -= rl[r];
            for (int j = 0; j < n; ++j) cl[j] = cc[j] > 0 ? csum[j] / (double)cc[j] : 0.0;
            for (int k = 0; k < src->cones.num_cones; ++k) {
                if (!preserve) {
                    if (src->cones.type[k] == CONE_PSD)
                        psd_log_project(cl, src->cones.start_idx[k], src->cones.v_dim[k], f);
                    continue;
                }
                int st = src->cones.start_idx[k], L = orc_cone_block_length(&src->cones, k);
                double bsum = 0.0;
                int bc = 0;
                for (int j = st; j < st + L; ++j) { bsum += csum[j]; bc += cc[j]; }
                double bl = bc > 0 ? bsum / (double)bc : 0.0;
                for (int j = st; j < st + L; ++j) cl[j] = bl;
            }
        }
        for (int r = 0; r < m; ++r) {
            double v = rl[r];
            cr[r] = isnan(v) ? 1.0 : exp(v > LOG_LIMIT ? LOG_LIMIT : (v < -LOG_LIMIT ? -LOG_LIMIT : v));
        }
        for (int j = 0; j < n; ++j) {
            double v = cl[j];
            vr[j] = isnan(v) ? 1.0 : exp(v > LOG_LIMIT ? LOG_LIMIT : (v < -LOG_LIMIT ? -LOG_LIMIT : v));
        }
        free(rl); free(cl); free(rs); free(cs); free(csum); free(cc);
    }
    scale_by(w, cr, vr);
    for (int i = 0; i < m; ++i) cum_c[i] *= cr[i];
    for (int j = 0; j < n; ++j) cum_v[j] *= vr[j];
    free(cr); free(vr); free(f);
}

/* reference src/preconditioner.c:497-556 */
static void ruiz(orc_scaled *w, const qp_problem_t *src, int iters, bool preserve, double *cum_c, double *cum_v) {
    int m = w->m, n = w->n;
    double *cr = (double *)xm((size_t)m * sizeof(double)), *vr = (double *)xm((size_t)n * sizeof(double));
    for (int it = 0; it < iters; ++it) {
        for (int j = 0; j < n; ++j) vr[j] = 0.0;
        for (int i = 0; i < m; ++i) cr[i] = 0.0;
        for (int r = 0; r < m; ++r)
            for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) {
                double a = fabs(w->A.v[e]);
                int c = w->A.ci[e];
                if (a > vr[c]) vr[c] = a;
                if (a > cr[r]) cr[r] = a;
            }
        for (int j = 0; j < n; ++j) vr[j] = vr[j] < SC_EPS ? 1.0 : sqrt(vr[j]);
        for (int i = 0; i < m; ++i) cr[i] = cr[i] < SC_EPS ? 1.0 : sqrt(cr[i]);
        cone_adjust(vr, &src->cones, 0, preserve);
        cone_adjust(cr, &src->affine_cones, 0, preserve);
        scale_by(w, cr, vr);
        for (int j = 0; j < n; ++j) cum_v[j] *= vr[j];
        for (int i = 0; i < m; ++i) cum_c[i] *= cr[i];
    }
    free(cr); free(vr);
}

/* reference src/preconditioner.c:558-600 */
static void pock_chambolle(orc_scaled *w, const qp_problem_t *src, double alpha, bool preserve, double *cum_c,
                           double *cum_v) {
    int m = w->m, n = w->n;
    double *cr = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    double *vr = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int r = 0; r < m; ++r)
        for (int e = w->A.rp[r]; e < w->A.rp[r + 1]; ++e) {
            double a = fabs(w->A.v[e]);
            vr[w->A.ci[e]] += pow(a, 2.0 - alpha);
            cr[r] += pow(a, alpha);
        }
    for (int j = 0; j < n; ++j) vr[j] = vr[j] < SC_EPS ? 1.0 : sqrt(vr[j]);
    for (int i = 0; i < m; ++i) cr[i] = cr[i] < SC_EPS ? 1.0 : sqrt(cr[i]);
    cone_adjust(vr, &src->cones, 1, preserve);
    cone_adjust(cr, &src->affine_cones, 1, preserve);
    scale_by(w, cr, vr);
    for (int j = 0; j < n; ++j) cum_v[j] *= vr[j];
    for (int i = 0; i < m; ++i) cum_c[i] *= cr[i];
    free(cr); free(vr);
}

/* reference src/preconditioner.c:602-652 */
static void bound_obj(orc_scaled *w) {
    double b2 = 0.0;
    for (int i = 0; i < w->m; ++i) {
        if (isfinite(w->cl[i]) && w->cl[i] != w->cu[i]) b2 += w->cl[i] * w->cl[i];
        if (isfinite(w->cu[i])) b2 += w->cu[i] * w->cu[i];
    }
    for (int i = 0; i < w->m; ++i) b2 += w->g[i] * w->g[i];
    double c2 = 0.0;
    for (int j = 0; j < w->n; ++j) c2 += w->c[j] * w->c[j];
    w->cbr = 1.0 / (sqrt(b2) + 1.0);
    w->ovr = 1.0 / (sqrt(c2) + 1.0);
    for (int i = 0; i < w->m; ++i) { w->cl[i] *= w->cbr; w->cu[i] *= w->cbr; w->g[i] *= w->cbr; }
    for (int j = 0; j < w->n; ++j) { w->lb[j] *= w->cbr; w->ub[j] *= w->cbr; w->c[j] *= w->ovr; }
    for (int e = 0; e < w->Q.nnz; ++e) w->Q.v[e] *= w->ovr / w->cbr;
    if (w->R.nnz > 0) {
        double f = sqrt(w->ovr / w->cbr);
        for (int e = 0; e < w->R.nnz; ++e) w->R.v[e] *= f;
    }
}

/* Q type: reference src/utils.cu:834-885 */
static quad_obj_type_t detect_qtype(const orc_scaled *w) {
    int has_s = w->Q.nnz > 0, has_r = w->R.nnz > 0;
    if (!has_s && !has_r) return PDHCG_NON_Q;
    if (has_r) return has_s ? PDHCG_LOW_RANK_PLUS_SPARSE_Q : PDHCG_LOW_RANK_Q;
    for (int i = 0; i < w->n; ++i)
        for (int e = w->Q.rp[i]; e < w->Q.rp[i + 1]; ++e)
            if (w->Q.ci[e] != i) return PDHCG_SPARSE_Q;
    return PDHCG_DIAG_Q;
}

/* reference src/preconditioner.c:654-713 and preprocess_qp_problem :715-830 */
orc_scaled *orc_rescale(const pdhg_parameters_t *params, const qp_problem_t *p) {
    clock_t t0 = clock();
    orc_scaled *w = (orc_scaled *)calloc(1, sizeof(orc_scaled));
    w->n = p->num_variables;
    w->m = p->num_constraints;
    w->k = p->num_rank_lowrank_obj;
    csr_copy(&w->A, p->constraint_matrix, w->m, w->n, p->constraint_matrix_num_nonzeros);
    csr_copy(&w->Q, p->objective_sparse_matrix, w->n, w->n, p->objective_sparse_matrix_num_nonzeros);
    csr_copy(&w->R, p->objective_lowrank_matrix, w->k, w->n, p->objective_lowrank_matrix_num_nonzeros);
    w->lb = dup(p->variable_lower_bound, w->n);
    w->ub = dup(p->variable_upper_bound, w->n);
    w->c = dup(p->objective_vector, w->n);
    w->cl = dup(p->constraint_lower_bound, w->m);
    w->cu = dup(p->constraint_upper_bound, w->m);
    w->g = dup(p->affine_cone_offset, w->m);
    w->obj_const = p->objective_constant;
    w->con_rescale = (double *)xm((size_t)w->m * sizeof(double));
    w->var_rescale = (double *)xm((size_t)w->n * sizeof(double));
    for (int i = 0; i < w->m; ++i) w->con_rescale[i] = 1.0;
    for (int j = 0; j < w->n; ++j) w->var_rescale[j] = 1.0;
    bool preserve = params->use_cone_preserving_scaling;
    if (params->curtis_reid_iterations > 0)
        curtis_reid(w, p, params->curtis_reid_iterations, preserve, w->con_rescale, w->var_rescale);
    if (params->l_inf_ruiz_iterations > 0)
        ruiz(w, p, params->l_inf_ruiz_iterations, preserve, w->con_rescale, w->var_rescale);
    if (params->has_pock_chambolle_alpha)
        pock_chambolle(w, p, params->pock_chambolle_alpha, preserve, w->con_rescale, w->var_rescale);
    if (params->bound_objective_rescaling) bound_obj(w);
    else { w->cbr = 1.0; w->ovr = 1.0; }
    /* reference src/preconditioner.c:311-328 */
    if (p->cones.is_fixed)
        for (int j = 0; j < w->n; ++j)
            if (p->cones.is_fixed[j]) {
                double v = p->primal_start ? p->primal_start[j] : 0.0;
                w->lb[j] = w->ub[j] = v * w->var_rescale[j] * w->cbr;
            }
    /* D kind: reference src/preconditioner.c:742-791 */
    const CsrComponent *D = p->objective_lowrank_middle_matrix;
    int dnnz = p->objective_lowrank_middle_matrix_num_nonzeros, k = w->k;
    if (D && D->row_ptr && dnnz > 0 && k > 0) {
        int diag_only = 1;
        for (int r = 0; r < k && diag_only; ++r)
            for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e)
                if (D->col_ind[e] != r) { diag_only = 0; break; }
        if (diag_only) {
            w->dkind = 1;
            w->ddiag = (double *)calloc((size_t)k, sizeof(double));
            for (int r = 0; r < k; ++r)
                for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e) w->ddiag[r] = D->val[e];
        } else {
            w->dkind = 2;
            w->ddense = (double *)calloc((size_t)k * k, sizeof(double));
            for (int r = 0; r < k; ++r)
                for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e)
                    w->ddense[(size_t)r * k + D->col_ind[e]] = D->val[e];
        }
    }
    w->qtype = detect_qtype(w);
    if (w->qtype == PDHCG_DIAG_Q) {
        w->qdiag = (double *)calloc((size_t)(w->n > 0 ? w->n : 1), sizeof(double));
        for (int i = 0; i < w->n; ++i)
            for (int e = w->Q.rp[i]; e < w->Q.rp[i + 1]; ++e)
                if (w->Q.ci[e] < w->n) w->qdiag[w->Q.ci[e]] = w->Q.v[e] + 1e-12;
    }
    w->rescale_time = (double)(clock() - t0) / CLOCKS_PER_SEC;
    return w;
}

void orc_scaled_free(orc_scaled *w) {
    if (!w) return;
    orc_csr_free(&w->A); orc_csr_free(&w->Q); orc_csr_free(&w->R);
    free(w->lb); free(w->ub); free(w->c); free(w->cl); free(w->cu); free(w->g);
    free(w->ddiag); free(w->ddense); free(w->qdiag); free(w->con_rescale); free(w->var_rescale);
    free(w);
}
