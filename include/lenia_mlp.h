/*
 * lenia_mlp.h -- C-ABI of the parameter-proposal network (in liblenia_hip.so).
 *
 * An MLP 81 -> 256 -> 256 -> 256 -> 3 with GELU (tanh form) hidden activations that
 * scores Flow Lenia parameter sets: output 0 is the survival logit, 1 the final mass
 * concentration, 2 the centre-of-mass drift.  Trained in bf16 on the matrix cores with
 * fp32 master weights, Adam (beta1 0.9, beta2 0.999, eps 1e-8), loss = BCE(out0) +
 * MSE(out1) + MSE(out2), mean over the batch.
 *
 * No reference counterpart: the reference has no learned component.  The engine the
 * network learns from is the step behind lenia_hip.h.
 *
 * Parameters are fp32 in one padded array (lm_param_count floats): W1 [128][256], b1
 * [256], W2 [256][256], b2, W3 [256][256], b3, W4 [256][64], b4 [64], row-major with
 * W[in][out]; rows 81..127 of W1 and columns 3..63 of W4 / b4 must be zero.
 */
#ifndef LENIA_MLP_H
#define LENIA_MLP_H

#include <stddef.h>
#include <stdint.h>

#include "lenia_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LM_IN 81
#define LM_IN_PAD 128
#define LM_HIDDEN 256
#define LM_OUT 3
#define LM_OUT_PAD 64

size_t lm_param_count(void);
/* max_batch: rows per training step and per inference chunk (rounded up to 64). */
lh_status lm_create(int device, int max_batch, void **mlp);
lh_status lm_destroy(void *mlp);
lh_status lm_set_params(void *mlp, const float *src, size_t count);
lh_status lm_get_params(void *mlp, float *dst, size_t count);

/* Training set resident in HBM: x [n][81] features, y [n][3] targets (survival 0/1,
 * concentration, drift). */
lh_status lm_set_data(void *mlp, const float *x, const float *y, int n);
/* n_batches Adam steps on the rows order[b*batch .. b*batch+batch) of the resident
 * set; *mean_loss (optional) receives the mean of the batch losses. */
lh_status lm_train(void *mlp, const int *order, int n_batches, int batch, float lr, float *mean_loss);
/* Loss and gradient (lm_param_count floats) of one batch of host rows, no update. */
lh_status lm_gradients(void *mlp, const float *x, const float *y, int n, float *grad, float *loss);
/* Outputs [n][3] of host rows x [n][81]. */
lh_status lm_forward(void *mlp, const float *x, int n, float *out);
/* Scores n parameter sets drawn on the GPU from (seed, index) -- the same draws as
 * lenia_amd/proposal.py -- in chunks of max_batch; logits[n] = output 0. */
lh_status lm_score_sampled(void *mlp, uint64_t seed, long long n, float *logits);
/* Features [n][81] of sampled sets (the sampler's output, for checking it). */
lh_status lm_sample_features(void *mlp, uint64_t seed, long long first, int n, float *x);

#ifdef __cplusplus
}
#endif
#endif
