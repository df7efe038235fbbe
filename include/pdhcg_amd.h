/*
 * Additional entry points of libpdhcg_amd.so (MI355X).  None of these exist
 * in the reference interface; they are new symbols that leave the pdhcg.h ABI
 * untouched.  They expose what the reference keeps internal (the solver state
 * of reference src/solver.cu:35-210, stepped in chunks of iterations) so that
 * benchmarks and tests can time the iteration without setup, and they select
 * the arithmetic precision out of band (the parameter struct is ABI).
 *
 * The precision and graph-window switches are process-wide and unsynchronised:
 * set them before solves start, not while another thread is solving.  Each
 * solve owns its HIP stream and device buffers; a cancel request applies to
 * every running solve.
 */
#pragma once

#include "pdhcg_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the device iteration.  FP64: matrix values, vectors and
 * SpMV accumulation in double (the reference's arithmetic).  MIXED: matrix
 * values, vectors and SpMV accumulation in float; every reduction, the
 * restart test, the primal weight and the KKT evaluation in double. */
#define PDHCG_PRECISION_FP64 0
#define PDHCG_PRECISION_MIXED 1

/* Precision used by solve_qp_problem.  Defaults to FP64, or to MIXED when
 * the environment variable PDHCG_PRECISION is "mixed" when first queried. */
void pdhcg_set_precision(int precision);
int pdhcg_get_precision(void);

/* Graph replay of iteration windows (default off; the environment variable
 * PDHCG_GRAPH=1 turns it on when first queried).  For LP and diagonal-Q
 * problems without cones, each run of iterations between two KKT evaluations
 * (no restart, no fixed-point error) is replayed from an instantiated
 * hipGraph of its SpMV launches instead of being launched kernel by kernel;
 * the arithmetic is the same, so results are bit-identical either way.  The
 * setting applies to solvers created afterwards; windows are not used while
 * kernel timing is enabled.  pdhcg_solver_graph_stats reports the windows
 * replayed so far and the iterations they covered.  Measured on MI355X the
 * replay is 3-7% slower than the launch-by-launch loop, which already runs
 * at the device's dispatch rate; hence the default. */
void pdhcg_set_graph_window(int enable);
int pdhcg_get_graph_window(void);

/* Sets or clears the cancellation flag checked at every KKT evaluation
 * (termination reason USER_INTERRUPT), as the reference's SIGINT handler does
 * through its global flag (reference src/pdhcg.c:31). */
void pdhcg_request_cancel(int flag);

/* Solver handle: scaling, device upload and step-size estimation happen in
 * create; run performs up to `iterations` passes of the main loop body
 * (evaluations, restarts and termination included) and returns the
 * termination reason, 0 while still running, or -1 on a device error.
 * The problem must outlive the handle. */
typedef struct pdhcg_solver pdhcg_solver_t;
pdhcg_solver_t *pdhcg_solver_create(const qp_problem_t *prob, const pdhg_parameters_t *params, int precision);
int pdhcg_solver_run(pdhcg_solver_t *solver, int iterations);
int pdhcg_solver_counts(pdhcg_solver_t *solver, int *total_iterations, int *inner_iterations);
int pdhcg_solver_synchronize(pdhcg_solver_t *solver);
pdhcg_result_t *pdhcg_solver_result(pdhcg_solver_t *solver);
void pdhcg_solver_free(pdhcg_solver_t *solver);

/* Kernel timing with HIP events recorded on the solver's stream.
 * pdhcg_solver_timing(s, p) with p >= 1 clears and enables recording of every
 * p-th launch of each instrumented kernel (an event pair leaves the GPU idle
 * for a few microseconds, so p > 1 keeps the overhead off the iteration);
 * p = 0 disables it.  pdhcg_solver_kernel_time then returns the summed
 * duration and the count of the recorded launches of one of the
 * instrumented kernels:
 *   0  R x per-chunk partial sums (dense low-rank R), ungated launches
 *   1  R'u plus the diagonal of Q (dense low-rank R), ungated launches
 *   2  A'y SpMV of the PDHG step (fused primal epilogue, or stored for BB)
 *   3  A x_bar SpMV of the PDHG step with the fused dual epilogue */
int pdhcg_solver_timing(pdhcg_solver_t *solver, int enable);
int pdhcg_solver_graph_stats(pdhcg_solver_t *solver, int *windows, int *iterations);
int pdhcg_solver_kernel_time(pdhcg_solver_t *solver, int kernel, double *total_ms, int *launches);

/* The device cone operators the solver applies to its iterates, on host
 * vectors of n entries (the reference applies them internally:
 * src/cone_dispatch.cu:38-57 projection, :88-115 dual residual).
 * d is the per-slot scaling (NULL: all ones); the projection is onto
 * diag(d) K for every block of `cones`, in place on x.  The dual residual
 * writes (rc - P_{K*}(rc)) * d on the cone slots of out and leaves the other
 * slots of out as they are.  Blocks with fixed slots (cones->is_fixed, of at
 * least n entries) are projected onto their fixed section: the fixed slots
 * keep their value and the free ones move to the nearest point, in the metric
 * d^2, that puts the block in the cone (reference
 * src/kernels/pdhcg_soc_cone_kernels.cu:23-260 and its siblings); the dual
 * residual rejects such blocks (their residual needs a primal point).
 * Return 0, or -1 with a message on stderr. */
int pdhcg_project_cones(const cone_blocks_t *cones, const double *d, double *x, int n);
int pdhcg_cone_dual_residual(const cone_blocks_t *cones, const double *d, const double *rc, double *out, int n);

/* Projection of `count` points, stored one after another in x (count * n
 * doubles), in order and in place, through one device cone set: each PSD
 * eigensolve starts from the eigenvectors the previous point's projection
 * left for that block, as the solver's iterations do (reference: the same
 * projection per point, src/kernels/pdhcg_psd_cone_kernels.cu:415-486, with
 * a fresh eigensolve each time).  The results are the projections of the
 * points (the nearest point is unique); only the rounding and the number of
 * Jacobi sweeps depend on the previous point.  Return 0, or -1 with a
 * message on stderr. */
int pdhcg_project_cones_seq(const cone_blocks_t *cones, const double *d, double *x, int n, int count);

/* pdhcg_project_cones without device allocation; it has no counterpart in the
 * reference.  It is for callers that cannot call hipMalloc / hipFree: code
 * under HIP graph capture elsewhere in the process, tight serving loops,
 * harnesses that forbid device allocation.  Same arguments, kernels and
 * results (bit-identical) as pdhcg_project_cones, but every device buffer of
 * the call (the point, the scaling, the block lists and their warm starts, the
 * PSD and long-section work space) is carved from one fixed 32 MiB array of
 * the library in each device's global memory, held until the process ends.
 * A call whose buffers exceed it (about n <= 1.9M slots with few blocks) is
 * refused: -1 and a message, x untouched.  Calls from several host threads
 * are serialized by a process-wide lock (they would otherwise race on the
 * one array); the array keeps the last call's contents, and each call
 * clears what must start at zero.  Works on the device current at the call.
 * Returns 0, or -1 with a message on stderr. */
int pdhcg_project_cones_noalloc(const cone_blocks_t *cones, const double *d, double *x, int n);

/* Batched solve; the reference solves one problem per call and has no
 * counterpart.  `count` independent problems are advanced together on the
 * device, one workgroup per problem, each running the same restarted Halpern
 * PDHG loop as solve_qp_problem (FP64) with its own step size, primal weight,
 * restarts and termination.  results[i] receives problem i's result (free it
 * with pdhcg_result_free), or NULL when problem i has something the batch does
 * not handle (message on stderr): variable cone blocks, a sparse Q with
 * off-diagonal entries or of type PDHCG_SPARSE_Q, quadratic constraints,
 * affine PSD cones, a low-rank term above rank 256.  Affine second-order
 * cones of any length are taken (above 64 vector slots the problem's
 * workgroup projects them one at a time).  params (NULL: defaults) apply to every
 * problem; the time limit is checked between windows of iterations.  With
 * pdhcg_set_precision(PDHCG_PRECISION_MIXED) the problems first iterate with
 * FP32 iterates and matrix values (FP64 reductions and scalars) until each
 * terminates or stops improving at the FP32 rounding floor; the iterates are
 * then widened and every problem is evaluated, and if need be finished, in
 * FP64, so every status and residual is an FP64 one.  Returns 0, or -1 for
 * invalid arguments or a device error. */
int pdhcg_solve_batch(const qp_problem_t *const *problems, int count, const pdhg_parameters_t *params,
                      pdhcg_result_t **results);

#ifdef __cplusplus
}
#endif
