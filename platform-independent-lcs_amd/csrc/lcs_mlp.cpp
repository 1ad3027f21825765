// C ABI of the learned schedule cost model (declared in include/lcs_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <random>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"
#include "mlp_kernels.h"

using namespace lcs;

struct lcs_mlp {
    lcs_ctx *ctx = nullptr;
    MlpShape S{};
    int nparam = 0;
    int adam_t = 0;
    DevBuf params, adam_m, adam_v, grads, work;
};

namespace {
int reserve_work(lcs_mlp *M, int N, float **X, float **T, float **Mk, float **A1, float **A2, float **Y, float **D3,
                 float **D2, float **D1, float **lt, float **loss)
{
    const MlpShape &S = M->S;
    const size_t words = (size_t)N * (S.I + 2 * S.O + 4 * S.H + 2 * S.O) + (size_t)N + 16;
    int rc = M->work.reserve(M->ctx, words * 4);
    if (rc) return rc;
    float *p = (float *)M->work.ptr;
    *X = p; p += (size_t)N * S.I;
    *T = p; p += (size_t)N * S.O;
    *Mk = p; p += (size_t)N * S.O;
    *A1 = p; p += (size_t)N * S.H;
    *A2 = p; p += (size_t)N * S.H;
    *Y = p; p += (size_t)N * S.O;
    *D3 = p; p += (size_t)N * S.O;
    *D2 = p; p += (size_t)N * S.H;
    *D1 = p; p += (size_t)N * S.H;
    *lt = p; p += N;
    *loss = p;
    return LCS_OK;
}
}  // namespace

namespace lcs {
const lcs_ctx *mlp_ctx(const lcs_mlp *M) { return M ? M->ctx : nullptr; }

int mlp_shape(const lcs_mlp *M, int *n_in, int *hidden, int *n_out)
{
    if (!M) return LCS_EINVAL;
    if (n_in) *n_in = M->S.I;
    if (hidden) *hidden = M->S.H;
    if (n_out) *n_out = M->S.O;
    return LCS_OK;
}
}  // namespace lcs

extern "C" {

int lcs_mlp_create(lcs_ctx *ctx, int32_t n_in, int32_t hidden, int32_t n_out, uint64_t seed, lcs_mlp **out)
{
    if (!ctx || !out || n_in < 1 || n_in > 64 || hidden < 1 || hidden > kMlpMaxH || n_out < 1 || n_out > kMlpMaxO)
        return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    lcs_mlp *M = new (std::nothrow) lcs_mlp();
    if (!M) return LCS_ENOMEM;
    M->ctx = ctx;
    M->S = MlpShape{n_in, hidden, n_out};
    M->nparam = mlp_param_count(M->S);
    const size_t bytes = (size_t)M->nparam * 4;
    int rc;
    if ((rc = M->params.reserve(ctx, bytes)) || (rc = M->adam_m.reserve(ctx, bytes)) ||
        (rc = M->adam_v.reserve(ctx, bytes)) || (rc = M->grads.reserve(ctx, bytes))) {
        lcs_mlp_destroy(M);
        return rc;
    }
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases
    std::vector<float> P((size_t)M->nparam, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int rows, int fan_in) {
        const double lim = 1.0 / std::sqrt((double)fan_in);
        for (size_t i = 0; i < (size_t)rows * fan_in; ++i)
            P[off + i] = (float)(((double)(rng() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * lim);
    };
    const MlpShape &S = M->S;
    fill(0, S.H, S.I);
    fill((size_t)S.H * S.I + S.H, S.H, S.H);
    fill((size_t)S.H * S.I + S.H + (size_t)S.H * S.H + S.H, S.O, S.H);
    std::vector<float> zeros((size_t)M->nparam, 0.f);
    if (hipMemcpy(M->params.ptr, P.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->adam_m.ptr, zeros.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(M->adam_v.ptr, zeros.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        lcs_mlp_destroy(M);
        return LCS_EIO;
    }
    *out = M;
    return LCS_OK;
}

int lcs_mlp_destroy(lcs_mlp *M)
{
    if (!M) return LCS_EINVAL;
    if (M->ctx) detach_model(M->ctx, M);   // a context the model was attached to forgets it
    M->params.release(M->ctx);
    M->adam_m.release(M->ctx);
    M->adam_v.release(M->ctx);
    M->grads.release(M->ctx);
    M->work.release(M->ctx);
    delete M;
    return LCS_OK;
}

int lcs_mlp_param_count(lcs_mlp *M, int32_t *count)
{
    if (!M || !count) return LCS_EINVAL;
    *count = M->nparam;
    return LCS_OK;
}

int lcs_mlp_get_params(lcs_mlp *M, float *out)
{
    if (!M || !out) return LCS_EINVAL;
    return hipMemcpy(out, M->params.ptr, (size_t)M->nparam * 4, hipMemcpyDeviceToHost) == hipSuccess ? LCS_OK : LCS_EIO;
}

int lcs_mlp_set_params(lcs_mlp *M, const float *in)
{
    if (!M || !in) return LCS_EINVAL;
    M->adam_t = 0;
    return hipMemcpy(M->params.ptr, in, (size_t)M->nparam * 4, hipMemcpyHostToDevice) == hipSuccess ? LCS_OK : LCS_EIO;
}

int lcs_mlp_predict(lcs_mlp *M, const float *X, int32_t n, float *Y)
{
    if (!M || n < 0 || (n && (!X || !Y))) return LCS_EINVAL;
    if (n == 0) return LCS_OK;
    float *dX, *dT, *dM, *A1, *A2, *dY, *D3, *D2, *D1, *lt, *loss;
    int rc = reserve_work(M, n, &dX, &dT, &dM, &A1, &A2, &dY, &D3, &D2, &D1, &lt, &loss);
    if (rc) return rc;
    hipStream_t s = M->ctx->stream;
    if (hipMemcpyAsync(dX, X, (size_t)n * M->S.I * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if ((rc = lcsm_forward(M->S, (const float *)M->params.ptr, dX, nullptr, nullptr, dY, n, s))) return rc;
    if (hipMemcpyAsync(Y, dY, (size_t)n * M->S.O * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    return hipStreamSynchronize(s) == hipSuccess ? LCS_OK : LCS_EIO;
}

int lcs_mlp_loss_grad(lcs_mlp *M, const float *X, const float *T, const float *Mk, int32_t n, float *loss_out,
                      float *grad_out)
{
    if (!M || n <= 0 || !X || !T || !Mk) return LCS_EINVAL;
    float *dX, *dT, *dM, *A1, *A2, *dY, *D3, *D2, *D1, *lt, *loss;
    int rc = reserve_work(M, n, &dX, &dT, &dM, &A1, &A2, &dY, &D3, &D2, &D1, &lt, &loss);
    if (rc) return rc;
    hipStream_t s = M->ctx->stream;
    const MlpShape &S = M->S;
    double cnt = 0;
    for (size_t i = 0; i < (size_t)n * S.O; ++i) cnt += Mk[i];
    const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    if (hipMemcpyAsync(dX, X, (size_t)n * S.I * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dT, T, (size_t)n * S.O * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dM, Mk, (size_t)n * S.O * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return LCS_EIO;
    const float *P = (const float *)M->params.ptr;
    float *G = (float *)M->grads.ptr;
    if ((rc = lcsm_forward(S, P, dX, A1, A2, dY, n, s))) return rc;
    if ((rc = lcsm_backward(S, P, dX, A1, A2, dY, dT, dM, inv, D3, D2, D1, lt, loss, G, n, s))) return rc;
    if (loss_out && hipMemcpyAsync(loss_out, loss, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (grad_out && hipMemcpyAsync(grad_out, G, (size_t)M->nparam * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return LCS_EIO;
    return hipStreamSynchronize(s) == hipSuccess ? LCS_OK : LCS_EIO;
}

int lcs_mlp_train(lcs_mlp *M, const float *X, const float *T, const float *Mk, int32_t n, int32_t steps, float lr,
                  float *final_loss)
{
    if (!M || n <= 0 || steps < 0 || !X || !T || !Mk || !(lr > 0.f)) return LCS_EINVAL;
    float *dX, *dT, *dM, *A1, *A2, *dY, *D3, *D2, *D1, *lt, *loss;
    int rc = reserve_work(M, n, &dX, &dT, &dM, &A1, &A2, &dY, &D3, &D2, &D1, &lt, &loss);
    if (rc) return rc;
    hipStream_t s = M->ctx->stream;
    const MlpShape &S = M->S;
    double cnt = 0;
    for (size_t i = 0; i < (size_t)n * S.O; ++i) cnt += Mk[i];
    const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    if (hipMemcpyAsync(dX, X, (size_t)n * S.I * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dT, T, (size_t)n * S.O * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(dM, Mk, (size_t)n * S.O * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return LCS_EIO;
    float *P = (float *)M->params.ptr, *G = (float *)M->grads.ptr;
    for (int it = 0; it < steps; ++it) {
        if ((rc = lcsm_forward(S, P, dX, A1, A2, dY, n, s))) return rc;
        if ((rc = lcsm_backward(S, P, dX, A1, A2, dY, dT, dM, inv, D3, D2, D1, lt, loss, G, n, s))) return rc;
        if ((rc = lcsm_adam(P, G, (float *)M->adam_m.ptr, (float *)M->adam_v.ptr, M->nparam, lr, ++M->adam_t, s)))
            return rc;
    }
    // loss after the last update
    if ((rc = lcsm_forward(S, P, dX, A1, A2, dY, n, s))) return rc;
    if ((rc = lcsm_backward(S, P, dX, A1, A2, dY, dT, dM, inv, D3, D2, D1, lt, loss, G, n, s))) return rc;
    if (final_loss && hipMemcpyAsync(final_loss, loss, 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    return hipStreamSynchronize(s) == hipSuccess ? LCS_OK : LCS_EIO;
}

int lcs_mlp_choose(lcs_mlp *M, const float *X, int32_t n, int32_t *choice)
{
    if (!M || n < 0 || (n && (!X || !choice))) return LCS_EINVAL;
    if (n == 0) return LCS_OK;
    float *dX, *dT, *dM, *A1, *A2, *dY, *D3, *D2, *D1, *lt, *loss;
    int rc = reserve_work(M, n, &dX, &dT, &dM, &A1, &A2, &dY, &D3, &D2, &D1, &lt, &loss);
    if (rc) return rc;
    hipStream_t s = M->ctx->stream;
    int32_t *dC = (int32_t *)dT;  // T is unused for inference
    if (hipMemcpyAsync(dX, X, (size_t)n * M->S.I * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if ((rc = lcsm_forward(M->S, (const float *)M->params.ptr, dX, nullptr, nullptr, dY, n, s))) return rc;
    if ((rc = lcsm_argmin(M->S, dY, n, dC, s))) return rc;
    if (hipMemcpyAsync(choice, dC, (size_t)n * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    return hipStreamSynchronize(s) == hipSuccess ? LCS_OK : LCS_EIO;
}

}  // extern "C"
