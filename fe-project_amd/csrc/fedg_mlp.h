// Per-column MLP (fedg_mlp.hip): dense ReLU network, BF16 matrix-core
// products with FP32 accumulation, FP32 master weights, Adam.
#pragma once
#include <hip/hip_runtime.h>

#include "fedg_internal.h"

namespace fedg {

struct MlpImpl;

int mlp_create(int n_in, int n_hidden, int n_hidden_layers, int n_out, unsigned seed, MlpImpl** out);
void mlp_destroy(MlpImpl* p);
// Y [B x n_out] = net(X [B x n_in]); device fp32, row-major
int mlp_forward(MlpImpl* p, const float* X, float* Y, int B);
// mean squared error against T and its gradient w.r.t. every parameter
int mlp_loss_grad(MlpImpl* p, const float* X, const float* T, int B, float* loss);
int mlp_adam(MlpImpl* p, float lr, float b1, float b2, float eps);
long mlp_param_count(const MlpImpl* p);
// parameters and gradients: W of every layer (N x K row-major), then every bias
int mlp_get(MlpImpl* p, float* params_host, float* grads_host);
int mlp_set(MlpImpl* p, const float* params_host);
int mlp_dims(const MlpImpl* p, int* n_in, int* n_out);
hipStream_t mlp_stream(const MlpImpl* p);

}  // namespace fedg
