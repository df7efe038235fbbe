// Diagonal equilibration of the working problem, FP64 on the host.
//
// Algorithm and constants follow reference src/preconditioner.c:
//   Curtis-Reid log least squares (:334-495), L-inf Ruiz (:497-556),
//   Pock-Chambolle with exponent alpha (:558-600), bound/objective scaling
//   (:602-652), cone-preserving block merging (:108-144), PSD diagonal
//   congruence in log space (:47-99), fixed-slot pinning (:311-328), Q type and
//   D kind detection (:715-830, src/utils.cu:834-885).
// Every scaling pass divides the stored values by the same factors in the same
// order as the reference, so the scaled data match it bit for bit.
#include <chrono>
#include <cmath>
#include <cstring>

#include "host_scale.h"

namespace pdhcg {

HostCSR transpose(const HostCSR &a) {
    HostCSR t;
    t.rows = a.cols;
    t.cols = a.rows;
    t.rp.assign((size_t)t.rows + 1, 0);
    t.ci.resize(a.ci.size());
    t.v.resize(a.v.size());
    for (int c : a.ci) t.rp[(size_t)c + 1]++;
    for (int j = 0; j < t.rows; ++j) t.rp[(size_t)j + 1] += t.rp[j];
    std::vector<int> cursor(t.rp.begin(), t.rp.end() - 1);
    for (int r = 0; r < a.rows; ++r)
        for (int e = a.rp[r]; e < a.rp[(size_t)r + 1]; ++e) {
            int pos = cursor[a.ci[e]]++;
            t.ci[pos] = r;
            t.v[pos] = a.v[e];
        }
    return t;
}

namespace {

constexpr double kScaleEps = 1e-12;
constexpr double kLogMinAbs = 1e-300;
constexpr double kLogLimit = 69.07755278982137;
constexpr double kPsdLogLimit = 0.5 * kLogLimit;
constexpr int kTaperLength = 8;

enum class Phase { Ruiz, PockChambolle };

HostCSR csr_from(const CsrComponent *c, int rows, int cols, int nnz) {
    HostCSR h;
    h.rows = rows;
    h.cols = cols;
    h.rp.assign((size_t)rows + 1, 0);
    if (c && c->row_ptr) std::memcpy(h.rp.data(), c->row_ptr, ((size_t)rows + 1) * sizeof(int));
    if (nnz > 0) {
        h.ci.assign(c->col_ind, c->col_ind + nnz);
        h.v.assign(c->val, c->val + nnz);
    }
    return h;
}

int max_psd_order(const cone_blocks_t *b) {
    int mx = 0;
    for (int k = 0; b && k < b->num_cones; ++k)
        if (b->type[k] == CONE_PSD && b->v_dim[k] > mx) mx = b->v_dim[k];
    return mx;
}

double clamp_log(double v) {
    if (std::isnan(v)) return 0.0;
    return v > kLogLimit ? kLogLimit : (v < -kLogLimit ? -kLogLimit : v);
}

// Replaces the log scalings of one svec block by the nearest diagonal
// congruence d_r + d_c (least squares over the block, reference :47-89).
void psd_congruence_log(double *ls, int start, int order, std::vector<double> &f) {
    f.assign((size_t)order, 0.0);
    int slot = start;
    for (int c = 0; c < order; ++c)
        for (int r = c; r < order; ++r, ++slot) {
            double v = clamp_log(ls[slot]);
            if (r == c) f[r] += 2.0 * v;
            else {
                f[r] += v;
                f[c] += v;
            }
        }
    double sum = 0.0;
    for (int i = 0; i < order; ++i) sum += f[i];
    const double corr = sum / (2.0 * order + 2.0), diag = order + 2.0;
    for (int i = 0; i < order; ++i) f[i] = std::fmax(-kPsdLogLimit, std::fmin(kPsdLogLimit, (f[i] - corr) / diag));
    slot = start;
    for (int c = 0; c < order; ++c)
        for (int r = c; r < order; ++r, ++slot) ls[slot] = f[r] + f[c];
}

// Makes the scaling uniform over each cone block (or congruent over PSD
// blocks when blocks are not preserved), reference :108-144.
void merge_cone_scaling(std::vector<double> &s, const cone_blocks_t *b, Phase phase, bool preserve) {
    if (!b) return;
    std::vector<double> f;
    for (int k = 0; k < b->num_cones; ++k) {
        const int st = b->start_idx[k], len = cone_block_slots(b, k);
        if (!preserve) {
            if (b->type[k] == CONE_PSD) {
                for (int i = st; i < st + len; ++i) s[i] = std::log(std::fmax(s[i], kLogMinAbs));
                psd_congruence_log(s.data(), st, b->v_dim[k], f);
                for (int i = st; i < st + len; ++i) s[i] = std::exp(s[i]);
            }
            continue;
        }
        double mx = 0.0, sq = 0.0;
        for (int i = st; i < st + len; ++i) {
            mx = std::fmax(mx, s[i]);
            sq += s[i] * s[i];
        }
        const double rms = std::sqrt(sq / (double)len);
        const bool small = len <= kTaperLength;
        const double v = phase == Phase::Ruiz ? (small ? mx : rms) : (small ? rms : std::sqrt(mx * rms));
        for (int i = st; i < st + len; ++i) s[i] = v;
    }
}

struct Scaler {
    ScaledProblem &w;
    const cone_blocks_t *cones, *acones;
    MatrixPasses *dev;  // entry-wise passes of Ruiz, Pock-Chambolle and bound/objective (null: here)

    // Divides the problem by row factors cr and column factors vr and
    // accumulates them (reference :256-309 and the cum_* updates).
    void apply(const std::vector<double> &cr, const std::vector<double> &vr) {
        for (int j = 0; j < w.n; ++j) {
            w.c[j] /= vr[j];
            w.ub[j] *= vr[j];
            w.lb[j] *= vr[j];
        }
        for (int i = 0; i < w.m; ++i) {
            w.cl[i] /= cr[i];
            w.cu[i] /= cr[i];
            w.g[i] /= cr[i];
        }
        for (int j = 0; j < w.n; ++j) w.var_rescale[j] *= vr[j];
        for (int i = 0; i < w.m; ++i) w.con_rescale[i] *= cr[i];
        if (dev) {
            dev->divide(cr, vr);
            return;
        }
        for (int r = 0; r < w.m; ++r)
            for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) w.A.v[e] /= (cr[r] * vr[w.A.ci[e]]);
        for (int r = 0; r < w.n; ++r)
            for (int e = w.Q.rp[r]; e < w.Q.rp[(size_t)r + 1]; ++e) w.Q.v[e] /= (vr[r] * vr[w.Q.ci[e]]);
        if (w.R.nnz() > 0)
            for (int r = 0; r < w.k; ++r)
                for (int e = w.R.rp[r]; e < w.R.rp[(size_t)r + 1]; ++e) w.R.v[e] /= vr[w.R.ci[e]];
    }

    static double root_or_one(double v) { return v < kScaleEps ? 1.0 : std::sqrt(v); }

    void ruiz(int iters, bool preserve) {
        std::vector<double> cr((size_t)w.m), vr((size_t)w.n);
        for (int it = 0; it < iters; ++it) {
            std::fill(cr.begin(), cr.end(), 0.0);
            std::fill(vr.begin(), vr.end(), 0.0);
            if (dev) dev->abs_max(cr, vr);
            else
                for (int r = 0; r < w.m; ++r)
                    for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) {
                        const double a = std::fabs(w.A.v[e]);
                        const int c = w.A.ci[e];
                        if (a > vr[c]) vr[c] = a;
                        if (a > cr[r]) cr[r] = a;
                    }
            for (double &v : vr) v = root_or_one(v);
            for (double &v : cr) v = root_or_one(v);
            merge_cone_scaling(vr, cones, Phase::Ruiz, preserve);
            merge_cone_scaling(cr, acones, Phase::Ruiz, preserve);
            apply(cr, vr);
        }
    }

    void pock_chambolle(double alpha, bool preserve) {
        std::vector<double> cr((size_t)w.m, 0.0), vr((size_t)w.n, 0.0);
        if (dev) dev->pow_sums(alpha, cr, vr);
        else
            for (int r = 0; r < w.m; ++r)
                for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) {
                    const double a = std::fabs(w.A.v[e]);
                    vr[w.A.ci[e]] += std::pow(a, 2.0 - alpha);
                    cr[r] += std::pow(a, alpha);
                }
        for (double &v : vr) v = root_or_one(v);
        for (double &v : cr) v = root_or_one(v);
        merge_cone_scaling(vr, cones, Phase::PockChambolle, preserve);
        merge_cone_scaling(cr, acones, Phase::PockChambolle, preserve);
        apply(cr, vr);
    }

    void curtis_reid(int iters, bool preserve) {
        const int m = w.m, n = w.n;
        std::vector<double> cr((size_t)m, 1.0), vr((size_t)n, 1.0);
        std::vector<double> f;
        if (iters > 0 && m > 0 && n > 0 && w.A.nnz() > 0) {
            std::vector<double> rl(m, 0.0), cl(n, 0.0), rs(m, 0.0), cs(n, 0.0), csum(n, 0.0);
            std::vector<int> cc(n, 0);
            for (int r = 0; r < m; ++r)
                for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) {
                    const double la = std::log(std::fmax(std::fabs(w.A.v[e]), kLogMinAbs));
                    rs[r] += la;
                    cs[w.A.ci[e]] += la;
                    cc[w.A.ci[e]]++;
                }
            for (int it = 0; it < iters; ++it) {
                for (int r = 0; r < m; ++r) {
                    const int b = w.A.rp[r], e2 = w.A.rp[(size_t)r + 1];
                    double s = rs[r];
                    for (int e = b; e < e2; ++e) s -= cl[w.A.ci[e]];
                    rl[r] = e2 > b ? s / (double)(e2 - b) : 0.0;
                }
                for (int k = 0; acones && k < acones->num_cones; ++k) {
                    if (!preserve) {
                        if (acones->type[k] == CONE_PSD)
                            psd_congruence_log(rl.data(), acones->start_idx[k], acones->v_dim[k], f);
                        continue;
                    }
                    const int st = acones->start_idx[k], len = cone_block_slots(acones, k);
                    double bs = 0.0;
                    int bc = 0;
                    for (int r = st; r < st + len; ++r) {
                        bs += rs[r];
                        bc += w.A.rp[(size_t)r + 1] - w.A.rp[r];
                        for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) bs -= cl[w.A.ci[e]];
                    }
                    const double v = bc > 0 ? bs / (double)bc : 0.0;
                    for (int r = st; r < st + len; ++r) rl[r] = v;
                }
                csum = cs;
                for (int r = 0; r < m; ++r)
                    for (int e = w.A.rp[r]; e < w.A.rp[(size_t)r + 1]; ++e) csum[w.A.ci[e]] -= rl[r];
                for (int j = 0; j < n; ++j) cl[j] = cc[j] > 0 ? csum[j] / (double)cc[j] : 0.0;
                for (int k = 0; cones && k < cones->num_cones; ++k) {
                    if (!preserve) {
                        if (cones->type[k] == CONE_PSD)
                            psd_congruence_log(cl.data(), cones->start_idx[k], cones->v_dim[k], f);
                        continue;
                    }
                    const int st = cones->start_idx[k], len = cone_block_slots(cones, k);
                    double bs = 0.0;
                    int bc = 0;
                    for (int j = st; j < st + len; ++j) {
                        bs += csum[j];
                        bc += cc[j];
                    }
                    const double v = bc > 0 ? bs / (double)bc : 0.0;
                    for (int j = st; j < st + len; ++j) cl[j] = v;
                }
            }
            auto expc = [](double v) { return std::isnan(v) ? 1.0 : std::exp(clamp_log(v)); };
            for (int r = 0; r < m; ++r) cr[r] = expc(rl[r]);
            for (int j = 0; j < n; ++j) vr[j] = expc(cl[j]);
        }
        apply(cr, vr);
    }

    void bound_objective() {
        double b2 = 0.0, c2 = 0.0;
        for (int i = 0; i < w.m; ++i) {
            if (std::isfinite(w.cl[i]) && w.cl[i] != w.cu[i]) b2 += w.cl[i] * w.cl[i];
            if (std::isfinite(w.cu[i])) b2 += w.cu[i] * w.cu[i];
        }
        for (int i = 0; i < w.m; ++i) b2 += w.g[i] * w.g[i];
        for (int j = 0; j < w.n; ++j) c2 += w.c[j] * w.c[j];
        w.cbr = 1.0 / (std::sqrt(b2) + 1.0);
        w.ovr = 1.0 / (std::sqrt(c2) + 1.0);
        for (int i = 0; i < w.m; ++i) {
            w.cl[i] *= w.cbr;
            w.cu[i] *= w.cbr;
            w.g[i] *= w.cbr;
        }
        for (int j = 0; j < w.n; ++j) {
            w.lb[j] *= w.cbr;
            w.ub[j] *= w.cbr;
            w.c[j] *= w.ovr;
        }
        if (dev) {
            dev->multiply_qr(w.ovr / w.cbr, std::sqrt(w.ovr / w.cbr));
            return;
        }
        for (double &v : w.Q.v) v *= w.ovr / w.cbr;
        if (w.R.nnz() > 0) {
            const double f = std::sqrt(w.ovr / w.cbr);
            for (double &v : w.R.v) v *= f;
        }
    }
};

quad_obj_type_t classify_q(const ScaledProblem &w) {
    const bool has_s = w.Q.nnz() > 0, has_r = w.R.nnz() > 0;
    if (!has_s && !has_r) return PDHCG_NON_Q;
    if (has_r) return has_s ? PDHCG_LOW_RANK_PLUS_SPARSE_Q : PDHCG_LOW_RANK_Q;
    for (int i = 0; i < w.n; ++i)
        for (int e = w.Q.rp[i]; e < w.Q.rp[(size_t)i + 1]; ++e)
            if (w.Q.ci[e] != i) return PDHCG_SPARSE_Q;
    return PDHCG_DIAG_Q;
}

}  // namespace

ScaledProblem scale_problem(const pdhg_parameters_t &params, const qp_problem_t &p, MatrixPasses *dev) {
    const auto t0 = std::chrono::steady_clock::now();
    ScaledProblem w;
    w.n = p.num_variables;
    w.k = p.num_rank_lowrank_obj;
    w.cones = &p.cones;
    w.affine_cones = &p.affine_cones;
    const bool dummy_row = p.num_constraints == 0 || p.constraint_matrix == nullptr;
    if (dummy_row) {
        // one free row x_0 in (-inf, inf) (reference src/utils.cu:82-134)
        w.m = 1;
        w.A.rows = 1;
        w.A.cols = w.n;
        w.A.rp = {0, 1};
        w.A.ci = {0};
        w.A.v = {1.0};
        w.cl = {-INFINITY};
        w.cu = {INFINITY};
        w.g = {0.0};
        w.dual_start = {0.0};
    } else {
        w.m = p.num_constraints;
        w.A = csr_from(p.constraint_matrix, w.m, w.n, p.constraint_matrix_num_nonzeros);
        w.cl.assign(p.constraint_lower_bound, p.constraint_lower_bound + w.m);
        w.cu.assign(p.constraint_upper_bound, p.constraint_upper_bound + w.m);
        w.g.assign(p.affine_cone_offset, p.affine_cone_offset + w.m);
        if (p.dual_start) w.dual_start.assign(p.dual_start, p.dual_start + w.m);
    }
    w.Q = csr_from(p.objective_sparse_matrix, w.n, w.n, p.objective_sparse_matrix_num_nonzeros);
    w.R = csr_from(p.objective_lowrank_matrix, w.k, w.n, p.objective_lowrank_matrix_num_nonzeros);
    w.lb.assign(p.variable_lower_bound, p.variable_lower_bound + w.n);
    w.ub.assign(p.variable_upper_bound, p.variable_upper_bound + w.n);
    w.c.assign(p.objective_vector, p.objective_vector + w.n);
    w.obj_const = p.objective_constant;
    if (p.primal_start) w.primal_start.assign(p.primal_start, p.primal_start + w.n);

    // norms of the unscaled working problem (reference src/solver_state.cu:854-934)
    const bool linf = params.optimality_norm == NORM_TYPE_L_INF;
    double acc = 0.0;
    for (double c : w.c) {
        if (linf) {
            if (std::fabs(c) > acc) acc = std::fabs(c);
        } else acc += c * c;
    }
    w.obj_norm = linf ? acc : std::sqrt(acc);
    acc = 0.0;
    for (int i = 0; i < w.m; ++i) {
        const double lo = w.cl[i], hi = w.cu[i];
        if (linf) {
            if (std::isfinite(lo) && lo != hi && std::fabs(lo) > acc) acc = std::fabs(lo);
            if (std::isfinite(hi) && std::fabs(hi) > acc) acc = std::fabs(hi);
        } else {
            if (std::isfinite(lo) && lo != hi) acc += lo * lo;
            if (std::isfinite(hi)) acc += hi * hi;
        }
    }
    for (int i = 0; i < w.m; ++i) acc = linf ? std::fmax(acc, std::fabs(w.g[i])) : acc + w.g[i] * w.g[i];
    w.bound_norm = linf ? acc : std::sqrt(acc);

    w.con_rescale.assign((size_t)w.m, 1.0);
    w.var_rescale.assign((size_t)w.n, 1.0);
    Scaler s{w, &p.cones, &p.affine_cones, nullptr};
    const bool preserve = params.use_cone_preserving_scaling;
    if (params.curtis_reid_iterations > 0) s.curtis_reid(params.curtis_reid_iterations, preserve);
    if (dev) {
        dev->load(w);
        s.dev = dev;
    }
    if (params.l_inf_ruiz_iterations > 0) s.ruiz(params.l_inf_ruiz_iterations, preserve);
    if (params.has_pock_chambolle_alpha) s.pock_chambolle(params.pock_chambolle_alpha, preserve);
    if (params.bound_objective_rescaling) s.bound_objective();
    else w.cbr = w.ovr = 1.0;
    if (dev) dev->finish(w);
    if (p.cones.is_fixed)
        for (int j = 0; j < w.n; ++j)
            if (p.cones.is_fixed[j]) {
                const double v = p.primal_start ? p.primal_start[j] : 0.0;
                w.lb[j] = w.ub[j] = v * w.var_rescale[j] * w.cbr;
            }

    // middle matrix D: diagonal when every stored entry sits on the diagonal
    const CsrComponent *D = p.objective_lowrank_middle_matrix;
    const int dnnz = p.objective_lowrank_middle_matrix_num_nonzeros;
    if (D && D->row_ptr && dnnz > 0 && w.k > 0) {
        bool diag_only = true;
        for (int r = 0; r < w.k && diag_only; ++r)
            for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e)
                if (D->col_ind[e] != r) {
                    diag_only = false;
                    break;
                }
        if (diag_only) {
            w.dkind = MiddleKind::Diag;
            w.ddiag.assign((size_t)w.k, 0.0);
            for (int r = 0; r < w.k; ++r)
                for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e) w.ddiag[r] = D->val[e];
        } else {
            w.dkind = MiddleKind::Dense;
            w.ddense.assign((size_t)w.k * w.k, 0.0);
            for (int r = 0; r < w.k; ++r)
                for (int e = D->row_ptr[r]; e < D->row_ptr[r + 1]; ++e)
                    w.ddense[(size_t)r * w.k + D->col_ind[e]] = D->val[e];
        }
    }
    w.qtype = classify_q(w);
    if (w.qtype == PDHCG_DIAG_Q) {
        w.qdiag.assign((size_t)w.n, 0.0);
        for (int i = 0; i < w.n; ++i)
            for (int e = w.Q.rp[i]; e < w.Q.rp[(size_t)i + 1]; ++e)
                if (w.Q.ci[e] < w.n) w.qdiag[w.Q.ci[e]] = w.Q.v[e] + 1e-12;
    }
    w.rescale_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return w;
}

}  // namespace pdhcg
