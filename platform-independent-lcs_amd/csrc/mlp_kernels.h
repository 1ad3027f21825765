// Launch interface of the MLP cost-model kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

constexpr int kMlpMaxH = 64;   // hidden width limit (per-thread arrays)
constexpr int kMlpMaxO = 32;   // output width limit (one predicted cost per candidate schedule)

struct MlpShape {
    int32_t I, H, O;
};

int mlp_param_count(const MlpShape &S);
// A1, A2 may be null (inference); Y may be null (training keeps activations only).
int lcsm_forward(const MlpShape &S, const float *P, const float *X, float *A1, float *A2, float *Y, int N,
                 hipStream_t s);
// Masked MSE: loss = sum(M * (Y - T)^2) * inv_count; G = dloss/dP.
int lcsm_backward(const MlpShape &S, const float *P, const float *X, const float *A1, const float *A2,
                  const float *Y, const float *T, const float *M, float inv_count, float *D3, float *D2, float *D1,
                  float *loss_terms, float *loss_out, float *G, int N, hipStream_t s);
int lcsm_adam(float *P, const float *G, float *Mo, float *Ve, int count, float lr, int t, hipStream_t s);
int lcsm_argmin(const MlpShape &S, const float *Y, int N, int32_t *choice, hipStream_t s);
