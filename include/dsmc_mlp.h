/* dsmc_mlp.h -- C ABI of the coordinate-MLP surrogate (no counterpart in the
 * reference, which has no learned component).
 *
 * A fully connected network widths[0] -> widths[1] -> ... -> widths[L] with
 * SiLU after every layer but the last, trained and run on the GPU by
 * hand-written BF16 MFMA kernels (FP32 accumulate): BF16 weights and
 * activations, FP32 master weights, Adam.  Its use here: predict per-cell
 * (log n, ux, uy, log T) at a new flow parameter point so a DSMC run can
 * start from local Maxwellians (dsmc_init_fields) instead of freestream.
 *
 * Parameters travel as one flat FP32 array: for each layer l the weights
 * W_l[out][in] (row-major) followed by the biases b_l[out].
 * Limits: widths[0] <= 32 inputs, widths[L] <= 32 outputs, hidden widths
 * any positive size (padded to multiples of 32 on the device).
 * All host arrays are borrowed for the duration of the call.
 */
#ifndef DSMC_MLP_H
#define DSMC_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsmc_mlp_s *dsmc_mlp;

/* Message of the last failed dsmc_mlp_* call on this thread.  Status codes:
 * 0 OK, -1 invalid argument, -2 HIP runtime failure. */
const char *dsmc_mlp_last_error(void);

/* nlayers weight layers, widths[nlayers + 1]; max_batch bounds every batch. */
int dsmc_mlp_create(const int *widths, int nlayers, int max_batch, dsmc_mlp *out);
int dsmc_mlp_destroy(dsmc_mlp m);
int dsmc_mlp_nparams(dsmc_mlp m, int64_t *n);
/* Sets the FP32 master parameters (and their BF16 copies); resets Adam. */
int dsmc_mlp_set_params(dsmc_mlp m, const float *params);
int dsmc_mlp_get_params(dsmc_mlp m, float *params);

/* Inference on n rows x[n][widths[0]] -> y[n][widths[L]] (host arrays). */
int dsmc_mlp_forward(dsmc_mlp m, const float *x, int64_t n, float *y);
/* One forward and backward pass of the mean squared error over a batch of
 * n rows (8 <= n <= max_batch, n a multiple of 8): loss = mean over rows and outputs of (y - t)^2;
 * grad = d loss / d params, flat like the parameters.  No update. */
int dsmc_mlp_gradient(dsmc_mlp m, const float *x, const float *t, int n, float *loss, float *grad);
/* One Adam step (bias-corrected) with the given gradient, or with the one the
 * last training step left on the device when grad is NULL. */
int dsmc_mlp_adam(dsmc_mlp m, const float *grad, float lr, float beta1, float beta2, float eps);

/* Training set resident in device memory: x[n][widths[0]], t[n][widths[L]]. */
int dsmc_mlp_set_data(dsmc_mlp m, const float *x, const float *t, int64_t n);
/* steps minibatch Adam steps of `batch` rows (a multiple of 8) drawn with replacement from the
 * resident set (Philox on (seed, step)); loss_out[steps] optional. */
int dsmc_mlp_train(dsmc_mlp m, int steps, int batch, float lr, float beta1, float beta2, float eps,
                   uint64_t seed, float *loss_out);
/* Mean time (ms) of the last train call's steps and the floating-point work
 * of one step (forward + backward) at its batch size. */
int dsmc_mlp_train_stats(dsmc_mlp m, double *ms_per_step, double *flop_per_step);

#ifdef __cplusplus
}
#endif
#endif
