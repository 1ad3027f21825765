// lcs_braid_multiply: staggered_multiply<RowGlue> of the reference
// (src/braid_multiplication.hpp:273-406) on the GPU, host pointers.
#include <hip/hip_runtime.h>

#include <vector>

#include "braid_kernels.h"
#include "lcs_hip.h"
#include "lcs_internal.h"

using namespace lcs;

namespace {
bool is_permutation(const int32_t *x, int32_t n)
{
    std::vector<unsigned char> seen((size_t)n, 0);
    for (int32_t i = 0; i < n; ++i) {
        if (x[i] < 0 || x[i] >= n || seen[(size_t)x[i]]) return false;
        seen[(size_t)x[i]] = 1;
    }
    return true;
}
}  // namespace

extern "C" int lcs_braid_multiply(lcs_ctx *ctx, const int32_t *p_row_to_col, int32_t p_size,
                                  const int32_t *q_row_to_col, int32_t q_size, int32_t k, int32_t row_glue,
                                  int32_t *out_row_to_col)
{
    if (!ctx || p_size < 0 || q_size < 0 || k < 0 || k > p_size || k > q_size) return LCS_EINVAL;
    if ((p_size && !p_row_to_col) || (q_size && !q_row_to_col)) return LCS_EINVAL;
    const int32_t out_size = p_size + q_size - k;
    if (out_size && !out_row_to_col) return LCS_EINVAL;
    if (!is_permutation(p_row_to_col, p_size) || !is_permutation(q_row_to_col, q_size)) return LCS_EINVAL;
    if (k > kMaxCoreK) return LCS_ERANGE;
    if (out_size == 0) return LCS_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    const size_t ws_words = lcsb_multiply_ws_words(p_size, q_size, k);
    const size_t words = (size_t)p_size + q_size + out_size + ws_words + 16;
    DevBuf buf;
    int rc = buf.reserve(ctx, words * 4);
    if (rc) return rc;
    int32_t *dp = (int32_t *)buf.ptr, *dq = dp + p_size, *dout = dq + q_size, *dws = dout + out_size;
    hipStream_t s = ctx->stream;
    bool ok = (!p_size || hipMemcpyAsync(dp, p_row_to_col, (size_t)p_size * 4, hipMemcpyHostToDevice, s) == hipSuccess) &&
              (!q_size || hipMemcpyAsync(dq, q_row_to_col, (size_t)q_size * 4, hipMemcpyHostToDevice, s) == hipSuccess);
    if (!ok) { buf.release(ctx); return LCS_EIO; }
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) { buf.release(ctx); return LCS_EIO; }
    rc = lcsb_multiply_device(dp, p_size, dq, q_size, k, row_glue ? 1 : 0, dout, dws, s);
    if (!rc) {
        (void)hipEventRecord(ctx->ev1, s);
        ok = hipMemcpyAsync(out_row_to_col, dout, (size_t)out_size * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        rc = ok ? LCS_OK : LCS_EIO;
        float ms = 0.f;
        if (ok && hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_ms = ms;
    }
    buf.release(ctx);
    return rc;
}
