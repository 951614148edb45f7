/*
 * channel_closure.h -- C-ABI of the learned subgrid-stress closure that
 * rides on the MI355X channel step: a 46 -> 64 -> 64 -> 64 -> 6 ReLU MLP with
 * BF16 weights and activations and FP32 accumulation on the matrix cores.
 *
 * The reference has no counterpart (it is a DNS code without subgrid
 * model).  Inputs are box-filtered velocity-gradient features, outputs the
 * six subgrid stresses tau_xx, tau_yy, tau_zz, tau_xy, tau_xz, tau_yz (the
 * feature and target definitions live with the dataset builder,
 * channel_amd/closure.py).
 *
 * Conventions as in channel_hip.h: int status (0 ok, > 0 HIP error, < 0
 * CHAN_E_*), chan_last_error() for the message, plain host pointers, one host
 * thread per handle.
 *
 * Parameters travel in the canonical FP32 layout W1[46][64], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], W4[64][6], b4[6]
 * (CHAN_MLP_NPARAM floats; W[in][out]).  The library keeps FP32 master
 * weights and BF16 copies for the kernels.
 */
#ifndef CHANNEL_CLOSURE_H
#define CHANNEL_CLOSURE_H

#include <stdint.h>

#include "channel_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CHAN_MLP_IN 46
#define CHAN_MLP_OUT 6
#define CHAN_MLP_NPARAM 11718

typedef struct chan_mlp chan_mlp;

/* Allocates device storage for up to `capacity` resident samples. */
int chan_mlp_create(int64_t capacity, chan_mlp **out);
int chan_mlp_destroy(chan_mlp *m);

int chan_mlp_set_params(chan_mlp *m, const float *params);
int chan_mlp_get_params(chan_mlp *m, float *params);

/* Uploads n <= capacity samples: features X[n][46] (rounded to BF16 on the
 * device) and targets T[n][6] (FP32).  Resets the Adam moments. */
int chan_mlp_set_data(chan_mlp *m, const float *X, const float *T, int64_t n);

/* Inference on host rows X[n][46] -> Y[n][6]; n <= capacity. */
int chan_mlp_forward(chan_mlp *m, const float *X, int64_t n, float *Y);

/* Inference that allocates nothing: device rows dX[n][46] (FP32) ->
 * dY[n][6] (FP32), both device pointers of the model's GPU, which must be
 * current.  Same kernels and results as chan_mlp_forward; the BF16 inputs
 * are staged in a static device array of CHAN_MLP_NOALLOC_MAX rows (8 MiB,
 * held by every GPU the library runs on until the process ends) instead of
 * device memory allocated per call.  n > CHAN_MLP_NOALLOC_MAX is rejected
 * with CHAN_E_ARG: split the call.  Runs on the model's stream and returns
 * when it is done.  All calls on one GPU share that array: calls that
 * overlap on one GPU (two host threads, two models) race.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). */
#define CHAN_MLP_NOALLOC_MAX 65536
int chan_mlp_forward_noalloc(chan_mlp *m, const float *dX, int64_t n, float *dY);

/* Loss sum((y - t)^2) / (6 count) over resident samples [first, first+count)
 * and its gradient in the canonical layout (FP32 accumulation of BF16
 * operands; the stored activations and dz are BF16). */
int chan_mlp_grad(chan_mlp *m, int64_t first, int64_t count, float *grad, double *loss);

/* `steps` Adam steps (beta1 0.9, beta2 0.999, eps 1e-8) over consecutive
 * minibatches of `batch` resident samples, wrapping around; the loss of each
 * step before its update goes to losses[steps] (may be NULL).  Returns the
 * average duration of one step in *ms_per_step (may be NULL). */
int chan_mlp_train(chan_mlp *m, int32_t steps, int32_t batch, float lr, float *losses, double *ms_per_step);

/* Timed inference over the resident samples: `reps` forward passes, average
 * milliseconds per pass in *ms. */
int chan_mlp_bench_forward(chan_mlp *m, int32_t reps, double *ms);

/* ---- the closure inside a coarse (LES) channel step, on the device ----
 * Samples are the points (iy, x, z) of rows iy = 0..ny of the context's
 * dealiased physical grid 2nxd x nzd, in that order: n = (ny+1) 2nxd nzd
 * (chan_les_samples).  Features, 46 per sample: the nine velocity gradients
 * du_c/dx_d (index 3c + d; d = x, y, z; x and z derivatives spectral, y
 * derivatives second-order differences on the stretched rows) at rows iy,
 * iy-2, iy-1, iy+1, iy+2 (clamped to 0..ny), then y+ = min(y-ymin, ymax-y)
 * u_tau / nu with u_tau from the mean wall shear of the (0,0) mode at both
 * walls.  Stresses (the MLP's outputs): tau_xx, tau_yy, tau_zz, tau_xy,
 * tau_xz, tau_yz. */
int chan_les_samples(const chan_ctx *c, int64_t *n);
/* Features X[n][46] (FP32, host) of the context's resident field. */
int chan_les_features(chan_ctx *c, float *X);
/* Sets the context's static body force (the F0 of chan_set_body_force) to
 * -div(tau), tau = MLP(m) of the normalised features of the resident field:
 * inputs (X - xmean) / xscale, stresses tau = Y * tscale + tmean, with
 * norm = [46 xmean][46 xscale][6 tmean][6 tscale].  Everything stays on the
 * device; the next chan_step / chan_buildrhs applies the force with the
 * wall treatment of chan_set_body_force. */
int chan_les_closure_force(chan_ctx *c, chan_mlp *m, const float *norm);

#ifdef __cplusplus
}
#endif
#endif
