/*
 * Problem permutation and partition cuts: the integer preparation that
 * precedes splitting one problem over several devices.  The names,
 * arguments and error behaviour are those of the reference's internal
 * interfaces, so the reference's distributed driver
 * (distributed/distributed_solver.cu:491-533) can call these in place of its
 * own.  All host code; nothing here touches a device.
 */
#pragma once

#include <stdbool.h>

#include "pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Permutes a problem in place: new variable i is old variable col_perm[i],
 * new row i is old row row_perm[i]; matrices, bounds, start values, cone
 * starts and the fixed mask follow.  Returns false (message on stderr) when
 * either permutation is not a permutation or splits / reorders a cone block.
 * Replaces reference src/permute.cu:143-224 (internal/permute.h:31). */
bool permute_problem(qp_problem_t *qp, int *row_perm, int *col_perm);

/* Deep copy, then permute_problem; NULL on failure.  Free with
 * qp_problem_free.  Replaces reference src/permute.cu:420-433 (internal/permute.h:33). */
qp_problem_t *permute_problem_return_new(const qp_problem_t *qp, int *row_perm, int *col_perm);

/* Random permutation of the variables (rows) that moves whole cone blocks
 * (affine cone blocks) and free runs of block_size (1 for
 * FULL_RANDOM_PERMUTATION), shuffled with rand() % (i + 1) from the last
 * unit down.  Seed with srand().  Replaces reference src/permute.cu:335-418
 * (internal/permute.h:35-39). */
void generate_cone_aware_permutation(const qp_problem_t *qp, permute_method_t method, int block_size, int *perm);
void generate_affine_cone_aware_row_permutation(const qp_problem_t *qp, permute_method_t method, int block_size,
                                                int *perm);

/* True when perm is a permutation keeping every (affine) cone block
 * contiguous and in order.  Replaces reference src/permute.cu:264-316
 * (internal/permute.h:40-41). */
bool validate_cone_permutation(const qp_problem_t *qp, const int *col_perm);
bool validate_affine_cone_row_permutation(const qp_problem_t *qp, const int *row_perm);

/* Maps a result of the permuted problem back to the original order.
 * Replaces reference src/permute.cu:483-494 (internal/permute.h:42). */
void repermute_solution(pdhcg_result_t *result, int *row_perm, int *col_perm);

/* cuts[0..P] (malloc'd, caller frees): cut p is one past the first index
 * where the running weight reaches p * total / P.  Replaces reference
 * distributed/distributed_utils.cu:179-216. */
int *get_balanced_cuts(const int *weights, int total_dim, int num_partitions);

/* Moves cuts[1..P-1] to boundaries outside the forbidden spans
 * [starts[i], ends[i]] (a cut at b separates index b-1 from b), minimising
 * the sum of squared displacements; false when no legal set exists.
 * Replaces reference src/partition_utils.c:88-248 (internal/partition_utils.h:25-30). */
bool optimize_partition_cuts(int total_dimension, int num_partitions, const int *forbidden_starts,
                             const int *forbidden_ends, int num_forbidden_intervals, int *cuts);

#ifdef __cplusplus
}
#endif
