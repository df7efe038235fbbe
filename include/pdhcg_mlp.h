/*
 * Learned warm start for a parametric problem family: a multilayer
 * perceptron in -> H -> H -> out (GELU after the two hidden layers) that maps
 * the parameters of an instance to a starting point (primal and dual values)
 * handed to set_start_values (reference src/pdhcg.c:569-618), which scales it
 * into the iteration space at initialisation
 * (reference src/solver_state.cu:809-829).  The reference has no learned
 * warm start; none of these symbols has a counterpart there.
 *
 * Arithmetic: FP32 master weights and Adam moments; the matrix products take
 * BF16 operands (round to nearest even) on the MI355X matrix cores with FP32
 * accumulation; bias, GELU, loss and optimizer in FP32.  All arrays are
 * row-major host arrays; every call synchronises before it returns.
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pdhcg_mlp pdhcg_mlp_t;

/* New network with He-normal weights (seeded) and zero biases; hidden must
 * be a multiple of 16.  NULL (message on stderr) on invalid sizes or a
 * device error. */
pdhcg_mlp_t *pdhcg_mlp_create(int in_dim, int hidden, int out_dim, unsigned seed);
void pdhcg_mlp_free(pdhcg_mlp_t *mlp);

/* Parameters in the order W1 (in x H), b1, W2 (H x H), b2, W3 (H x out), b3;
 * num_params gives the length of that vector. */
int pdhcg_mlp_num_params(const pdhcg_mlp_t *mlp);
int pdhcg_mlp_get_params(pdhcg_mlp_t *mlp, float *params);
int pdhcg_mlp_set_params(pdhcg_mlp_t *mlp, const float *params);

/* One forward and backward pass of the mean squared error
 * sum (Y - T)^2 / (rows * out) over X (rows x in) and T (rows x out), rows a
 * multiple of 16: the loss into *loss and, when grads is not NULL, the
 * gradient in the parameter order above.  The weights do not change. */
int pdhcg_mlp_loss_grad(pdhcg_mlp_t *mlp, const float *X, const float *T, int rows, double *loss, float *grads);

/* Adam (lr, betas 0.9 / 0.999, eps 1e-8) on minibatches of `batch` rows (a
 * multiple of 16), `epochs` passes over `count` samples in a seeded shuffled
 * order (a trailing partial minibatch is skipped); *final_loss receives the
 * mean minibatch loss of the last epoch.  Optimizer state persists across
 * calls. */
int pdhcg_mlp_train(pdhcg_mlp_t *mlp, const float *X, const float *T, int count, int epochs, int batch, float lr,
                    unsigned seed, double *final_loss);

/* Y (count x out) = network(X (count x in)). */
int pdhcg_mlp_predict(pdhcg_mlp_t *mlp, const float *X, int count, float *Y);

/* C (M x N) = op(A) op(B) through the BF16 matrix-core product used by the
 * network (operands rounded to BF16, FP32 accumulation); op is the transpose
 * of the row-major array when ta / tb is nonzero.  M, N, K positive multiples
 * of 16.  For tests of the kernel. */
int pdhcg_mfma_gemm(int ta, int tb, int M, int N, int K, const float *A, const float *B, float *C);

#ifdef __cplusplus
}
#endif
