/*
 * PDHCG C interface implemented by libpdhcg_amd.so (MI355X / gfx950).
 *
 * Each entry point keeps the name, signature, argument meaning and error
 * convention of the reference declaration it replaces
 * (reference include/pdhcg.h:22-70).  Errors are reported as in the
 * reference: NULL (or a nonzero int) plus one line on stderr.  Unlike the
 * reference, device errors never terminate the process; they also return
 * NULL.
 */
#pragma once

#include "pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Builds a problem from optional matrix descriptors (NULL Q means 0, NULL R
 * means no low-rank term, NULL D means identity).  A supplies the scalar rows
 * con_lb <= A x <= con_ub; F (affine_cone_matrix_desc) supplies rows with
 * F x + affine_cone_offset in the product of affine_cones.  Every array is
 * deep-copied.  Replaces reference src/pdhcg.c:365-508
 * (declaration include/pdhcg.h:31-46). */
qp_problem_t *create_qp_problem(const double *objective_c,
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

/* Warm start; dual has rows(A) + rows(F) entries in [dual_A, dual_F] order.
 * A primal start that changes a fixed cone slot is rejected with a message
 * and leaves the problem unchanged.  Replaces reference src/pdhcg.c:569-618
 * (include/pdhcg.h:49). */
void set_start_values(qp_problem_t *prob, const double *primal, const double *dual);

/* Fixes one slot of a variable cone block; nonzero on a bad index, a PSD
 * block or a non-finite value.  Replaces reference src/pdhcg.c:620-666
 * (include/pdhcg.h:51). */
int set_cone_fixed(qp_problem_t *prob, int cone_idx, int slot, double value);

/* Quadratic-constraint reformulation.  Problems with quadratic constraints
 * are outside this library's solve path: the call reports that on stderr and
 * returns NULL (reference include/pdhcg.h:53, src/qcqp_transform.c:96). */
qp_problem_t *qcqp_to_socp_qp(const qp_problem_t *orig_qcqp, cone_type_t default_type);

/* Solves on the current HIP device; params NULL means defaults.  The inputs
 * are not modified.  The result is heap-allocated and released by
 * pdhcg_result_free.  Replaces reference src/pdhcg.c:894-922
 * (include/pdhcg.h:56). */
pdhcg_result_t *solve_qp_problem(const qp_problem_t *prob, const pdhg_parameters_t *params);

/* Multi-process solve.  This build is single-GPU, so it behaves as the
 * reference's non-distributed build does: a message and NULL
 * (reference include/pdhcg.h:59, distributed/distributed_ops_stub.c:78-84). */
pdhcg_result_t *solve_qp_problem_distributed(const pdhg_parameters_t *params, const qp_problem_t *original_problem);

/* Fills params with the defaults (reference src/utils.cu:324-368,
 * include/pdhcg.h:62). */
void set_default_parameters(pdhg_parameters_t *params);

/* reference src/pdhcg.c:510-521 (include/pdhcg.h:64). */
void pdhcg_result_free(pdhcg_result_t *results);

/* reference src/pdhcg.c:532-567 (include/pdhcg.h:66). */
void qp_problem_free(qp_problem_t *prob);

#ifdef __cplusplus
}
#endif
