// mlp_internal.h -- what the surrogate step needs from the perceptron handle.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/nrg_mlp.h"

namespace nrg {

typedef __bf16 bf16;

constexpr int MLP_IN_PAD = 32;       // input features, zero padded
constexpr int MLP_OUT_PAD = 32;      // output columns, zero padded
constexpr int MLP_HIDDEN_MAX = 256;
constexpr int MLP_MAX_LAYERS = 8;    // weight matrices (hidden layers + 1)

// Fused inference on packed bf16 inputs X (n x MLP_IN_PAD) into Y (n x n_out
// FP32), enqueued on `stream` (no synchronisation).
#define NRG_HIDDEN __attribute__((visibility("hidden")))
NRG_HIDDEN int mlp_infer_bf16(nrg_mlp_t *h, long long n, const bf16 *X, float *Y, hipStream_t stream);
NRG_HIDDEN hipStream_t mlp_stream(nrg_mlp_t *h);
NRG_HIDDEN int mlp_n_out(const nrg_mlp_t *h);
NRG_HIDDEN int mlp_n_in(const nrg_mlp_t *h);

}  // namespace nrg
