// A small fp32 MLP (in -> H -> H -> out, ReLU) that predicts the cost of
// each combing schedule from a pair's shape; trained on the timings the
// autotuner records.  Hand-written forward, backward and Adam kernels; no
// library GEMM (the matrices are 64 wide and the batches a few thousand
// rows, so one thread per sample or per weight is the whole job).
//
// Parameter layout (floats): W1[H][I] b1[H] W2[H][H] b2[H] W3[O][H] b3[O].
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_kernels.h"

namespace lcsm {

__device__ __forceinline__ void offsets(const MlpShape &S, int &w1, int &b1, int &w2, int &b2, int &w3, int &b3)
{
    w1 = 0;
    b1 = w1 + S.H * S.I;
    w2 = b1 + S.H;
    b2 = w2 + S.H * S.H;
    w3 = b2 + S.H;
    b3 = w3 + S.O * S.H;
}

// One thread per sample: activations A1, A2 (after ReLU) and outputs Y.
__global__ void mlp_forward_kernel(MlpShape S, const float *P, const float *X, float *A1, float *A2, float *Y, int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int w1, b1, w2, b2, w3, b3;
    offsets(S, w1, b1, w2, b2, w3, b3);
    float h1[kMlpMaxH], h2[kMlpMaxH];
    const float *x = X + (size_t)n * S.I;
    for (int j = 0; j < S.H; ++j) {
        float acc = P[b1 + j];
        for (int k = 0; k < S.I; ++k) acc += P[w1 + j * S.I + k] * x[k];
        h1[j] = acc > 0.f ? acc : 0.f;
    }
    for (int j = 0; j < S.H; ++j) {
        float acc = P[b2 + j];
        for (int i = 0; i < S.H; ++i) acc += P[w2 + j * S.H + i] * h1[i];
        h2[j] = acc > 0.f ? acc : 0.f;
    }
    for (int o = 0; o < S.O; ++o) {
        float acc = P[b3 + o];
        for (int j = 0; j < S.H; ++j) acc += P[w3 + o * S.H + j] * h2[j];
        if (Y) Y[(size_t)n * S.O + o] = acc;
    }
    if (A1)
        for (int j = 0; j < S.H; ++j) {
            A1[(size_t)n * S.H + j] = h1[j];
            A2[(size_t)n * S.H + j] = h2[j];
        }
}

// One thread per sample: output error of the masked mean squared error and
// the errors back-propagated to both hidden layers; per-sample loss terms.
__global__ void mlp_delta_kernel(MlpShape S, const float *P, const float *A1, const float *A2, const float *Y,
                                 const float *T, const float *M, float inv_count, float *D3, float *D2, float *D1,
                                 float *loss_terms, int N)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int w1, b1, w2, b2, w3, b3;
    offsets(S, w1, b1, w2, b2, w3, b3);
    float d3[kMlpMaxO], d2[kMlpMaxH];
    float lt = 0.f;
    for (int o = 0; o < S.O; ++o) {
        const float e = Y[(size_t)n * S.O + o] - T[(size_t)n * S.O + o];
        const float mk = M[(size_t)n * S.O + o];
        lt += mk * e * e;
        d3[o] = 2.f * mk * e * inv_count;
        D3[(size_t)n * S.O + o] = d3[o];
    }
    loss_terms[n] = lt * inv_count;
    for (int j = 0; j < S.H; ++j) {
        float acc = 0.f;
        for (int o = 0; o < S.O; ++o) acc += d3[o] * P[w3 + o * S.H + j];
        d2[j] = A2[(size_t)n * S.H + j] > 0.f ? acc : 0.f;
        D2[(size_t)n * S.H + j] = d2[j];
    }
    for (int i = 0; i < S.H; ++i) {
        float acc = 0.f;
        for (int j = 0; j < S.H; ++j) acc += d2[j] * P[w2 + j * S.H + i];
        D1[(size_t)n * S.H + i] = A1[(size_t)n * S.H + i] > 0.f ? acc : 0.f;
    }
}

// One thread per parameter: the gradient summed over samples in a fixed
// order (deterministic).
__global__ void mlp_grad_kernel(MlpShape S, const float *X, const float *A1, const float *A2, const float *D3,
                                const float *D2, const float *D1, float *G, int N)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int w1, b1, w2, b2, w3, b3;
    offsets(S, w1, b1, w2, b2, w3, b3);
    const int total = b3 + S.O;
    if (p >= total) return;
    float acc = 0.f;
    if (p < b1) {
        const int j = (p - w1) / S.I, k = (p - w1) % S.I;
        for (int n = 0; n < N; ++n) acc += D1[(size_t)n * S.H + j] * X[(size_t)n * S.I + k];
    } else if (p < w2) {
        const int j = p - b1;
        for (int n = 0; n < N; ++n) acc += D1[(size_t)n * S.H + j];
    } else if (p < b2) {
        const int j = (p - w2) / S.H, i = (p - w2) % S.H;
        for (int n = 0; n < N; ++n) acc += D2[(size_t)n * S.H + j] * A1[(size_t)n * S.H + i];
    } else if (p < w3) {
        const int j = p - b2;
        for (int n = 0; n < N; ++n) acc += D2[(size_t)n * S.H + j];
    } else if (p < b3) {
        const int o = (p - w3) / S.H, j = (p - w3) % S.H;
        for (int n = 0; n < N; ++n) acc += D3[(size_t)n * S.O + o] * A2[(size_t)n * S.H + j];
    } else {
        const int o = p - b3;
        for (int n = 0; n < N; ++n) acc += D3[(size_t)n * S.O + o];
    }
    G[p] = acc;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction at step t.
__global__ void mlp_adam_kernel(float *P, const float *G, float *Mo, float *Ve, int count, float lr, int t)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = G[p];
    const float m = b1 * Mo[p] + (1.f - b1) * g;
    const float v = b2 * Ve[p] + (1.f - b2) * g * g;
    Mo[p] = m;
    Ve[p] = v;
    const float mh = m / (1.f - powf(b1, (float)t));
    const float vh = v / (1.f - powf(b2, (float)t));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

// Sum of the per-sample loss terms into out[0] (one block, fixed order).
__global__ void mlp_sum_kernel(const float *terms, int N, float *out)
{
    __shared__ float s[256];
    float acc = 0.f;
    for (int n = threadIdx.x; n < N; n += 256) acc += terms[n];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// One thread per row of X: the index of the smallest prediction, i.e. the
// schedule the model expects to be cheapest.
__global__ void mlp_argmin_kernel(MlpShape S, const float *Y, int N, int32_t *choice)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    int best = 0;
    float bv = Y[(size_t)n * S.O];
    for (int o = 1; o < S.O; ++o) {
        const float v = Y[(size_t)n * S.O + o];
        if (v < bv) { bv = v; best = o; }
    }
    choice[n] = best;
}

}  // namespace lcsm

using namespace lcsm;

int mlp_param_count(const MlpShape &S) { return S.H * S.I + S.H + S.H * S.H + S.H + S.O * S.H + S.O; }

static inline int blocks_for(int n) { return (n + 255) / 256 > 0 ? (n + 255) / 256 : 1; }

int lcsm_forward(const MlpShape &S, const float *P, const float *X, float *A1, float *A2, float *Y, int N,
                 hipStream_t s)
{
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks_for(N)), dim3(256), 0, s, S, P, X, A1, A2, Y, N);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsm_backward(const MlpShape &S, const float *P, const float *X, const float *A1, const float *A2,
                  const float *Y, const float *T, const float *M, float inv_count, float *D3, float *D2, float *D1,
                  float *loss_terms, float *loss_out, float *G, int N, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_delta_kernel, dim3(blocks_for(N)), dim3(256), 0, s, S, P, A1, A2, Y, T, M, inv_count, D3,
                       D2, D1, loss_terms, N);
    hipLaunchKernelGGL(mlp_sum_kernel, dim3(1), dim3(256), 0, s, loss_terms, N, loss_out);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks_for(mlp_param_count(S))), dim3(256), 0, s, S, X, A1, A2, D3, D2,
                       D1, G, N);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsm_adam(float *P, const float *G, float *Mo, float *Ve, int count, float lr, int t, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks_for(count)), dim3(256), 0, s, P, G, Mo, Ve, count, lr, t);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsm_argmin(const MlpShape &S, const float *Y, int N, int32_t *choice, hipStream_t s)
{
    hipLaunchKernelGGL(mlp_argmin_kernel, dim3(blocks_for(N)), dim3(256), 0, s, S, Y, N, choice);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
