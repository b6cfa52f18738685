/*
 * ns_closure.h -- learned subgrid closure for coarse spectral runs.
 *
 * No reference counterpart: the reference has no closure model.  A 4-layer
 * 3x3 convolutional network (1 -> 32 -> 32 -> 32 -> 1 channels, GELU,
 * periodic padding, 19,105 parameters) maps the coarse vorticity to the
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

#ifdef __cplusplus
}
#endif

#endif /* NS_CLOSURE_H */
