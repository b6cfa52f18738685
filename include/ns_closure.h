/*
 * ns_closure.h -- learned subgrid closure for coarse spectral runs.
 *
 * No reference counterpart: the reference has no closure model.  A 3x3
 * convolutional network of four hidden layers of 32 channels with GELU and a
 * linear output layer (1 -> 32 -> 32 -> 32 -> 32 -> 1 channels, periodic
 * padding, 28,353 parameters) maps the coarse vorticity to the
 * subgrid vorticity tendency.  It runs and trains on the GPU with bf16 MFMA
 * and FP32 accumulation; FP32 master parameters.
 *
 * Parameters are flat FP32, per layer the weights [cout][3*ky+kx][cin] and
 * then the biases [cout].  Fields are [batch][n][n] FP32 row-major.
 * Functions return 0 on success, non-zero on failure (nsd_closure_error).
 */
#ifndef NS_CLOSURE_H
#define NS_CLOSURE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct NsdClosure NsdClosure;

/* network for n x n fields (power of two, 16..4096), batches up to max_batch;
 * He-uniform initial weights from `seed`, zero biases */
NsdClosure *nsd_closure_create(int n, int max_batch, unsigned long long seed, char *err,
                               size_t errlen);
void nsd_closure_destroy(NsdClosure *c);
const char *nsd_closure_error(const NsdClosure *c);
int nsd_closure_param_count(const NsdClosure *c);
/* copy the FP32 parameters out of (to_device = 0) or into the network */
int nsd_closure_params(NsdClosure *c, float *params, int to_device);
/* y = net(x) for `batch` fields (host arrays) */
int nsd_closure_forward(NsdClosure *c, const float *x, int batch, float *y);
/* loss = mean (net(x) - t)^2 over all points; grad = d loss / d params */
int nsd_closure_gradient(NsdClosure *c, const float *x, const float *t, int batch, double *loss,
                         float *grad);

/* the network sees in_scale * vorticity and its output times out_scale is
 * the tendency; nsd_closure_train sets both from its data set (unit mean
 * squares), nsd_closure_set_scales for weights trained elsewhere */
int nsd_closure_set_scales(NsdClosure *c, double in_scale, double out_scale);
int nsd_closure_get_scales(const NsdClosure *c, double *in_scale, double *out_scale);
int nsd_closure_size(const NsdClosure *c);

/* Training data in device memory: `count` pairs (coarse vorticity, subgrid
 * tendency), each an n x n FP32 field.  nsd_closure_dataset_slot returns
 * the device addresses of pair `index` (filled by nsd_subgrid_sample of
 * ns_device.h); nsd_closure_dataset_fill copies host fields in. */
int nsd_closure_dataset_reserve(NsdClosure *c, int count);
int nsd_closure_dataset_slot(NsdClosure *c, int index, float **x, float **t);
int nsd_closure_dataset_fill(NsdClosure *c, int index, const float *x, const float *t);
int nsd_closure_dataset_read(NsdClosure *c, int index, float *x, float *t);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on the mean squared error over 90 % of
 * the `count` pairs in shuffled minibatches, the learning rate decaying from
 * lr to lr / 20 along a cosine over all steps; every tenth pair is held out
 * for the validation loss.  Losses are in the scaled units (1 = predicting
 * zero). */
int nsd_closure_train(NsdClosure *c, int count, int epochs, int batch, double lr,
                      unsigned long long seed, double *loss_first, double *loss_last,
                      double *loss_valid);
/* tendency y of the vorticity x, both device FP32 [n][n], queued on the HIP
 * stream `stream` (capturable into a graph; not re-entrant per network) */
int nsd_closure_eval_device(NsdClosure *c, const float *x, float *y, void *stream);
/* the same for `batch` fields [batch][n][n] (batch <= the network's max_batch) */
int nsd_closure_eval_batch(NsdClosure *c, const float *x, float *y, int batch, void *stream);
int nsd_closure_max_batch(const NsdClosure *c);

#ifdef __cplusplus
}
#endif

#endif /* NS_CLOSURE_H */
