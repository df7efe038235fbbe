/*
 * Public types of the PDHCG C interface, as exported by libpdhcg_amd.so.
 *
 * This header declares the same binary interface as the reference's
 * include/pdhcg_types.h:28-300 (Lhongpei/PDHCG v0.3.0): identical type
 * names, enum values, struct members, member order and member types, so a
 * program compiled against the reference header links against this library
 * unchanged.  The text is written for this project; only the interface is
 * shared.
 */
#pragma once

#include <stdbool.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Why the iteration stopped (reference include/pdhcg_types.h:28-39). */
typedef enum {
    TERMINATION_REASON_UNSPECIFIED,
    TERMINATION_REASON_OPTIMAL,
    TERMINATION_REASON_PRIMAL_INFEASIBLE,
    TERMINATION_REASON_DUAL_INFEASIBLE,
    TERMINATION_REASON_INFEASIBLE_OR_UNBOUNDED,
    TERMINATION_REASON_TIME_LIMIT,
    TERMINATION_REASON_ITERATION_LIMIT,
    TERMINATION_REASON_USER_INTERRUPT,
    TERMINATION_REASON_FEAS_POLISH_SUCCESS
} termination_reason_t;

/* Norm used by the KKT residuals (reference :41-45). */
typedef enum { NORM_TYPE_L2 = 0, NORM_TYPE_L_INF = 1 } norm_type_t;

/* Host CSR triple owned by a qp_problem_t (reference :47-52). */
typedef struct {
    int *row_ptr;
    int *col_ind;
    double *val;
} CsrComponent;

/* Structure of the quadratic objective (reference :54-61). */
typedef enum {
    PDHCG_SPARSE_Q,
    PDHCG_DIAG_Q,
    PDHCG_LOW_RANK_PLUS_SPARSE_Q,
    PDHCG_LOW_RANK_Q,
    PDHCG_NON_Q
} quad_obj_type_t;

/* Cone kinds; the numeric values are part of the ABI (reference :63-71).
 *   ROTATED_SOC  : slots (v_1..v_k, s, t), 2 s t >= |v|^2, s,t >= 0
 *   STANDARD_SOC : slots (v_1..v_k, w, z), |(v, w)|_2 <= z
 *   EXPONENTIAL  : (x, y, z), closure of {y > 0, y exp(x/y) <= z}
 *   POWER        : (x, y, z), x^a y^(1-a) >= |z|, x, y >= 0
 *   PSD          : svec of a symmetric order-k matrix, lower triangle in
 *                  column-major order, off-diagonal entries times sqrt(2). */
typedef enum {
    CONE_ROTATED_SOC = 0,
    CONE_STANDARD_SOC = 1,
    CONE_EXPONENTIAL = 2,
    CONE_POWER = 3,
    CONE_PSD = 4,
    NUM_CONE_TYPES = 5
} cone_type_t;

/* Structure-of-arrays description of cone blocks (reference :73-82). */
typedef struct {
    int num_cones;
    int *start_idx;
    int *v_dim;          /* SOC/RSOC vector dimension; PSD matrix order; 1 for EXP/POWER */
    cone_type_t *type;
    double *power_alpha; /* alpha in (0,1) for POWER blocks, NULL when there are none */
    int fixed_mask_size; /* length of is_fixed, 0 when no mask is stored */
    char *is_fixed;
} cone_blocks_t;

/* One cone block as given by the caller (reference :84-91). */
typedef struct {
    cone_type_t type;
    int start_idx;        /* variable index, or row of F for affine cones */
    int v_dim;
    double power_alpha;
    const char *is_fixed; /* per-slot fixed flags, variable non-PSD cones only */
} cone_spec_t;

/* Assembled problem: min c'x + 1/2 x'(Q + R'DR)x + const
 *                     s.t. l <= A x <= u, F x + g in K_aff, x in [lb, ub] and K
 * with [A; F] stored as one CSR (reference :93-134). */
typedef struct {
    int num_variables;
    int num_constraints;
    int num_rank_lowrank_obj;
    double *variable_lower_bound;
    double *variable_upper_bound;
    double *objective_vector;
    double objective_constant;

    CsrComponent *constraint_matrix;
    int constraint_matrix_num_nonzeros;

    CsrComponent *objective_sparse_matrix;
    int objective_sparse_matrix_num_nonzeros;

    CsrComponent *objective_lowrank_matrix;
    int objective_lowrank_matrix_num_nonzeros;

    CsrComponent *objective_lowrank_middle_matrix;
    int objective_lowrank_middle_matrix_num_nonzeros;

    double *constraint_lower_bound;
    double *constraint_upper_bound;

    double *affine_cone_offset;
    cone_blocks_t affine_cones;

    int num_quadratic_constraints;
    int *quadratic_constraint_row_indices;
    CsrComponent **quadratic_constraint_matrices;
    int *quadratic_constraint_matrix_num_nonzeros;

    cone_blocks_t cones;
    int num_original_variables;

    double *primal_start;
    double *dual_start;
} qp_problem_t;

/* Adaptive restart and PID primal-weight controls (reference :136-145). */
typedef struct {
    double artificial_restart_threshold;
    double sufficient_reduction_for_restart;
    double necessary_reduction_for_restart;
    double k_p;
    double k_i;
    double k_d;
    double i_smooth;
} restart_parameters_t;

typedef struct {
    double eps_optimal_relative;
    double eps_feasible_relative;
    double eps_feas_polish_relative;
    double eps_infeasible;
    double time_sec_limit;
    int iteration_limit;
} termination_criteria_t;

typedef struct {
    int iteration_limit;
    double initial_tolerance;
    double min_tolerance;
} inner_solver_parameters_t;

typedef enum {
    UNIFORM_PARTITION,
    NNZ_BALANCE_PARTITION,
} partition_method_t;

typedef enum {
    NO_PERMUTATION,
    FULL_RANDOM_PERMUTATION,
    BLOCK_RANDOM_PERMUTATION,
} permute_method_t;

typedef struct {
    int row_dims;
    int col_dims;
    bool decided;
} grid_size_t;

/* Solver parameters (reference :183-208). */
typedef struct {
    int curtis_reid_iterations;
    int l_inf_ruiz_iterations;
    bool has_pock_chambolle_alpha;
    double pock_chambolle_alpha;
    bool bound_objective_rescaling;
    bool use_cone_preserving_scaling;
    int verbose;
    int termination_evaluation_frequency;
    int sv_max_iter;
    double sv_tol;
    termination_criteria_t termination_criteria;
    restart_parameters_t restart_params;
    double reflection_coefficient;
    bool feasibility_polishing;
    norm_type_t optimality_norm;
    inner_solver_parameters_t inner_solver_parameters;
    bool presolve;
    bool diag_jacobi_precond;
    cone_type_t default_cone_type;
    partition_method_t partition_method;
    permute_method_t permute_method;
    grid_size_t grid_size;
    int permute_block_size;
} pdhg_parameters_t;

/* Solve result; arrays are malloc'd and released by pdhcg_result_free
 * (reference :210-248). */
typedef struct {
    int num_variables;
    int num_constraints;
    int num_nonzeros;

    int num_reduced_variables;
    int num_reduced_constraints;
    int num_reduced_nonzeros;

    double *primal_solution;
    double *dual_solution;
    double *reduced_cost;

    int total_count;
    int total_inner_count;
    double rescaling_time_sec;
    double cumulative_time_sec;

    double absolute_primal_residual;
    double relative_primal_residual;
    double absolute_dual_residual;
    double relative_dual_residual;
    double primal_objective_value;
    double dual_objective_value;
    double objective_gap;
    double relative_objective_gap;
    double max_primal_ray_infeasibility;
    double max_dual_ray_infeasibility;
    double primal_ray_linear_objective;
    double dual_ray_objective;
    termination_reason_t termination_reason;
    double feasibility_polishing_time;
    int feasibility_iteration;

    double presolve_time;
    int presolve_status;
} pdhcg_result_t;

/* Input matrix formats (reference :250-257). */
typedef enum { matrix_dense = 0, matrix_csr = 1, matrix_csc = 2, matrix_coo = 3 } matrix_format_t;

/* Input matrix descriptor; the library deep-copies what it points at
 * (reference :259-300). */
typedef struct {
    int m;
    int n;
    matrix_format_t fmt;
    double zero_tolerance; /* |a| <= tolerance is dropped (dense: default 1e-12) */
    union MatrixData {
        struct MatrixDense {
            const double *A; /* row-major m*n */
        } dense;
        struct MatrixCSR {
            int nnz;
            const int *row_ptr;
            const int *col_ind;
            const double *vals;
        } csr;
        struct MatrixCSC {
            int nnz;
            const int *col_ptr;
            const int *row_ind;
            const double *vals;
        } csc;
        struct MatrixCOO {
            int nnz;
            const int *row_ind;
            const int *col_ind;
            const double *vals;
        } coo;
    } data;
} matrix_desc_t;

#ifdef __cplusplus
}
#endif
