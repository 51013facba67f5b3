/*
 * esl_mlp.h -- surrogate network of the stationary-distribution pipeline.
 *
 * No counterpart in the reference.  A fully connected network maps the 2N
 * log-rates of a rate set to a distribution over the terminal-SCC states
 * (softmax output); it is trained on stationary vectors the pipeline computes
 * and serves as the warm start of the terminal-kernel solves
 * (esl_solve_opts.warm_pi), which remain exact: the Krylov solve continues
 * from the prediction until the FP64 residual meets its tolerance.
 *
 * All products run on the matrix cores (BF16 operands, FP32 accumulation);
 * parameters, activations and Adam moments are FP32.  Device pointers are
 * row-major [rows][features] FP32.
 */
#ifndef ESL_MLP_H
#define ESL_MLP_H

#include "esl.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct esl_mlp esl_mlp;

/* n_in -> hidden (x hidden_layers, GELU) -> n_out (softmax); seeded He init. */
esl_status esl_mlp_create(esl_ctx* ctx, int n_in, int hidden, int hidden_layers, int n_out, uint64_t seed,
						  esl_mlp** out);
void esl_mlp_destroy(esl_mlp* m);
/* Parameter vector layout: for each layer l, W_l (dims[l] x dims[l+1], row-major) then b_l. */
int64_t esl_mlp_num_params(const esl_mlp* m);
esl_status esl_mlp_get_params(esl_ctx* ctx, const esl_mlp* m, float* host_out);
esl_status esl_mlp_set_params(esl_ctx* ctx, esl_mlp* m, const float* host_in);
/* probabilities [rows][n_out] of inputs [rows][n_in] */
esl_status esl_mlp_forward(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, float* d_probs);
/* the prediction as a warm start: probabilities of inputs [rows][n_in], widened to
 * FP64 [rows][n_out] -- the layout of esl_solve_opts.warm_pi when n_out = n_t */
esl_status esl_mlp_predict_warm(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, double* d_warm);
/* gradients of the mean cross-entropy against targets [rows][n_out] (rows sum to 1),
 * copied to host_grads (num_params floats) when it is not NULL */
esl_status esl_mlp_gradients(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, const float* d_target,
							 float* host_grads, float* loss_out);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean cross-entropy */
esl_status esl_mlp_train_step(esl_ctx* ctx, esl_mlp* m, int rows, const float* d_x, const float* d_target, float lr,
							  float* loss_out);
/* `steps` Adam steps on consecutive minibatches of `rows` rows (step j reads rows
 * [j rows, (j + 1) rows) of d_x and d_target), the learning rate geometric from
 * lr_first to lr_last; enqueued back to back with one host synchronisation at
 * the end.  mean_loss_out (may be NULL): the mean loss over all the steps. */
esl_status esl_mlp_train_steps(esl_ctx* ctx, esl_mlp* m, int steps, int rows, const float* d_x,
							   const float* d_target, float lr_first, float lr_last, float* mean_loss_out);

#ifdef __cplusplus
}
#endif

#endif /* ESL_MLP_H */
