/*
 * ORACLE — FP64 CPU restatement of the reference PDHCG solve path.
 *
 * This is test infrastructure.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the
 * timed CPU baseline).  The product library libpdhcg_amd.so never links,
 * loads or calls anything here.
 *
 * Every function cites the reference file:line it restates
 * (Lhongpei/PDHCG v0.3.0).  The public types are those of include/pdhcg_types.h,
 * so problems and results can be compared field by field with the product.
 *
 * Pinning (what checks each part against something other than itself):
 *   orc_create_qp_problem / CSR conversions : reference known answers
 *       (test/test_problem_validation.c), brute-force dense round trips, and
 *       bit-exact agreement of two independent implementations.
 *   orc_rescale                             : closed forms on tiny matrices
 *       (Ruiz/Pock-Chambolle fixed points computed by hand in tests).
 *   orc_solve_qp_problem                    : reference end-to-end known answers
 *       (test/test_c_api_working.c:72-277, test/test_pdhcg_python.py:36-137),
 *       independent KKT recomputation from the returned (x, y) in numpy,
 *       and scipy's QP/LP solutions of the same random instances.
 *   Trajectory quantities (iteration counts, primal weight sequence, BB inner
 *   counts) are pinned only by the restatement itself: parity unpinned at
 *   trajectory level, pinned at tolerance level.
 */
#pragma once

#include "../include/pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Threads of the row-parallel products (OpenMP; results do not depend on
 * the count); 0 restores the default. */
void orc_set_threads(int threads);
int orc_get_threads(void);

/* ---- problem assembly (reference src/pdhcg.c, src/cone_utils.c, src/utils.cu:580-804) ---- */
qp_problem_t *orc_create_qp_problem(const double *objective_c,
                                    const matrix_desc_t *Q_desc,
                                    const matrix_desc_t *R_desc,
                                    const matrix_desc_t *D_desc,
                                    const matrix_desc_t *A_desc,
                                    const double *con_lb,
                                    const double *con_ub,
                                    const double *var_lb,
                                    const double *var_ub,
                                    const double *objective_constant,
                                    int num_var_cones,
                                    const cone_spec_t *var_cones,
                                    const matrix_desc_t *affine_cone_matrix_desc,
                                    const double *affine_cone_offset,
                                    int num_affine_cones,
                                    const cone_spec_t *affine_cones);
void orc_qp_problem_free(qp_problem_t *prob);
void orc_set_start_values(qp_problem_t *prob, const double *primal, const double *dual);
int orc_set_cone_fixed(qp_problem_t *prob, int cone_idx, int slot, double value);
int orc_validate_fixed_cone_sections(const qp_problem_t *problem);
void orc_set_default_parameters(pdhg_parameters_t *params);

/* CSR transpose with rows of the result in increasing source-row order
 * (the output of cusparseCsr2cscEx2 used at reference src/solver_state.cu:735-777). */
void orc_csr_transpose(int rows, int cols, const int *row_ptr, const int *col_ind, const double *val,
                       int *t_row_ptr, int *t_col_ind, double *t_val);

/* ---- solve (reference src/solver.cu:35-210) ---- */
pdhcg_result_t *orc_solve_qp_problem(const qp_problem_t *prob, const pdhg_parameters_t *params);
void orc_result_free(pdhcg_result_t *result);

/* Stepping interface for the timed CPU baseline: set up, run k iterations of
 * the main loop body, read back.  Same loop as orc_solve_qp_problem. */
typedef struct orc_solver orc_solver_t;
orc_solver_t *orc_solver_create(const qp_problem_t *prob, const pdhg_parameters_t *params);
int orc_solver_run(orc_solver_t *s, int iterations); /* returns termination reason (0 = still running) */
int orc_solver_total_count(const orc_solver_t *s);
int orc_solver_inner_count(const orc_solver_t *s);
void orc_solver_free(orc_solver_t *s);

/* Scaled-space quantities exposed for tests: after orc_solver_create,
 * copies the diagonal scalings and the step size. */
void orc_solver_scalings(const orc_solver_t *s, double *con_rescale, double *var_rescale, double *cbr, double *ovr,
                         double *step_size, double *primal_weight);

/* Loop state for the trajectory tests, in the scaled space: the stored PDHG
 * iterate and the restart anchor (n, m entries; x_p, x_0, the last delta
 * x written by the fixed-point error or a restart, y_p, y_0; any pointer may
 * be NULL), and [inner tolerance, fixed-point error, relative primal and
 * dual residual of the last evaluation, Jacobi tolerance scale]. */
void orc_solver_vectors(const orc_solver_t *s, double *xp, double *x0, double *dx, double *yp, double *y0);
void orc_solver_loop_state(const orc_solver_t *s, double *out);

/* The restart rule (reference src/utils.cu:291-322) as a pure function. */
int orc_restart_decision(const restart_parameters_t *rp, int ef, int total_count, int inner_count, double fpe,
                         double init_fpe, double last_trial_fpe);

/* ---- cone projections on the scaled space, one block at a time ----
 * x holds the block's slots, d the block's variable scaling (the projection
 * is onto D K; d = 1 gives the Euclidean projection onto K), alpha the power
 * cone exponent.  Returns 0, or -1 for an unknown type. */
int orc_project_block(int type, double *x, int v_dim, const double *d, double alpha);
/* Dual-cone residual (rc - P_{K*}(rc)) scaled by d, written to out. */
int orc_dual_residual_block(int type, const double *rc, int v_dim, const double *d, double alpha, double *out);
/* Fixed sections (orc_section.c): the block's fixed slots (fixed[i] != 0)
 * keep their value and the free ones are projected onto the section of D K in
 * the metric diag(metric d^2) (metric NULL: 1).  SOC, RSOC, EXP and POWER;
 * returns -1 for other types. */
int orc_section_project(int type, double *x, int v_dim, const double *d, const double *metric, const char *fixed,
                        double alpha);
/* Residual of a block with fixed slots: (x - P(x - rc)) d on the free slots,
 * 0 on the fixed ones, P the section projection in the metric d^2. */
void orc_section_dual_residual(int type, const double *x, const double *rc, int v_dim, const double *d,
                               const char *fixed, double alpha, double *out);
int orc_section_has_fixed(const char *fixed, int len);

/* Permutations and partition cuts (orc_permute.c; reference src/permute.cu,
 * src/partition_utils.c, distributed/distributed_utils.cu:179-216).  method:
 * permute_method_t values.  Return 1 on success, 0 on rejection. */
qp_problem_t *orc_qp_problem_copy(const qp_problem_t *p);
int orc_permute_problem(qp_problem_t *p, const int *row_perm, const int *col_perm);
qp_problem_t *orc_permute_problem_copy(const qp_problem_t *p, const int *row_perm, const int *col_perm);
int orc_validate_cone_permutation(const qp_problem_t *p, const int *col_perm);
int orc_validate_affine_row_permutation(const qp_problem_t *p, const int *row_perm);
void orc_generate_cone_aware_permutation(const qp_problem_t *p, int method, int block_size, int *perm);
void orc_generate_affine_row_permutation(const qp_problem_t *p, int method, int block_size, int *perm);
void orc_repermute_solution(pdhcg_result_t *r, const int *row_perm, const int *col_perm);
void orc_balanced_cuts(const int *weights, int total, int parts, int *cuts);
int orc_optimize_partition_cuts(int total, int parts, const int *forbidden_starts, const int *forbidden_ends,
                                int num_forbidden, int *cuts);

#ifdef __cplusplus
}
#endif

