/*
 * clspmv_mlp.h -- the learned format selector's network (libclspmv.so).
 *
 * No counterpart in the reference (clSpMV chooses formats from measured
 * tables only).  Callers: the format selector, which maps a matrix's
 * structural features to the best format/tile class measured by the
 * autotuner.  A fully connected network in -> hidden -> hidden -> classes
 * with ReLU, softmax cross-entropy loss and Adam, FP32, trained and run on
 * the GPU by this library's own kernels.  Parameters are one flat array:
 * W1 (in x hidden, row-major), b1, W2 (hidden x hidden), b2,
 * W3 (hidden x classes), b3.  All data pointers are host pointers; the
 * network copies them to the device.
 */
#ifndef CLSPMV_MLP_H_
#define CLSPMV_MLP_H_

#include <stdint.h>

#include "clspmv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clspmv_mlp_s* clspmv_mlp_t;

/* He-uniform initial weights from `seed`, zero biases. */
int clspmv_mlp_create(clspmv_ctx_t ctx, int in, int hidden, int classes, uint64_t seed, clspmv_mlp_t* out);
int clspmv_mlp_free(clspmv_mlp_t m);
long long clspmv_mlp_num_params(clspmv_mlp_t m);
int clspmv_mlp_get_params(clspmv_mlp_t m, float* params);
/* Also resets the Adam moments and step count (a new optimisation). */
int clspmv_mlp_set_params(clspmv_mlp_t m, const float* params);
/* Loss (mean cross-entropy) and its gradient with respect to every
 * parameter, for n samples x (n x in) with labels in [0, classes). */
int clspmv_mlp_loss_grad(clspmv_mlp_t m, const float* x, const int* labels, int n, float* loss, float* grads);
/* `steps` full-batch Adam steps (lr, beta1, beta2, eps, weight decay 0);
 * writes the final loss.  The update is Kingma & Ba's Algorithm 1:
 * m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * p -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).  The moments and the
 * step number t persist in the handle, so k calls of s steps take the same
 * steps as one call of k*s. */
int clspmv_mlp_train(clspmv_mlp_t m, const float* x, const int* labels, int n, int steps, float lr, float beta1,
                     float beta2, float eps, float* loss);
/* Logits (n x classes) and argmax class of each sample. */
int clspmv_mlp_predict(clspmv_mlp_t m, const float* x, int n, float* logits, int* classes);

#ifdef __cplusplus
}
#endif
#endif
