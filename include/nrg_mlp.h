/*
 * nrg_mlp.h -- BF16 MFMA multilayer perceptron for the learned chemistry
 * surrogate (gfx950), with its own training and inference kernels.
 *
 * The reference has no counterpart: NRG integrates every cell with SLATEC.
 * This network maps a cell's state and time step to the increment of one
 * split step; nrg_chem_surrogate.h gates its answers and routes rejected
 * cells back to the FP64 integrator of nrg_chem.h.
 *
 * Network: n_in -> [n_hidden x n_layers, GELU (tanh form)] -> n_out, weights
 * and activations in bfloat16 on MFMA (v_mfma_f32_32x32x16_bf16), FP32
 * accumulation, FP32 master weights and Adam moments.  Internally the input
 * is padded to 32 features and the output to 32 columns with zeros.
 *
 * Parameters are exchanged as one FP32 vector, layer by layer:
 *   W_l (fan_in x fan_out, row-major: W_l[k*fan_out + n]), then b_l (fan_out).
 * Inputs x are row-major n x n_in FP32; outputs y are row-major n x n_out FP32.
 * All device pointers are HBM pointers; every call synchronises the
 * handle's stream before returning.
 */
#ifndef NRG_MLP_H
#define NRG_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nrg_mlp nrg_mlp_t;

/* Returns NRG_CHEM_* status codes (nrg_chem.h); nrg_chem_last_error() has details. */
int nrg_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                   int32_t device, nrg_mlp_t **out);
void nrg_mlp_destroy(nrg_mlp_t *h);
int64_t nrg_mlp_num_params(const nrg_mlp_t *h);

/* Host FP32 parameter vector in the order above (also resets Adam's moments). */
int nrg_mlp_set_params(nrg_mlp_t *h, const float *params);
int nrg_mlp_get_params(nrg_mlp_t *h, float *params);

/* Inference with the fused kernel: y = f(x) for n rows. */
int nrg_mlp_forward_dev(nrg_mlp_t *h, int64_t n, const float *x, float *y);

/*
 * Loss and gradients of L = sum_rows sum_c w_c (y_c - t_c)^2 / n over a batch
 * (w: n_out FP32 weights, NULL = 1).  grad receives the FP32 gradient vector
 * in the parameter order (device pointer); loss (device, 1 float) may be NULL.
 * y_out (device, n x n_out, nullable) receives the training-path forward.
 */
int nrg_mlp_grad_dev(nrg_mlp_t *h, int64_t n, const float *x, const float *t, const float *w,
                     float *grad, float *loss, float *y_out);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; loss may be NULL. */
int nrg_mlp_train_step_dev(nrg_mlp_t *h, int64_t n, const float *x, const float *t,
                           const float *w, float lr, float *loss);

/* HIP stream of the handle (a hipStream_t). */
void *nrg_mlp_stream(nrg_mlp_t *h);

#ifdef __cplusplus
}
#endif
#endif /* NRG_MLP_H */
