// Semi-local score read-out on the GPU (declared in include/lcs_hip.h).
#include <hip/hip_runtime.h>

#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"

int lcsq_launch_dominance(const int32_t *perm, int32_t size, const int32_t *qr, const int32_t *qc, int32_t nq,
                          int32_t *out, hipStream_t s);

using namespace lcs;

namespace {
int run_dominance(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size, const std::vector<int32_t> &qr,
                  const std::vector<int32_t> &qc, int32_t *out)
{
    const int32_t nq = (int32_t)qr.size();
    DevBuf buf;
    int rc = buf.reserve(ctx, ((size_t)size + 3 * (size_t)nq + 4) * 4);
    if (rc) return rc;
    int32_t *dp = (int32_t *)buf.ptr, *dr = dp + size, *dc = dr + nq, *dout = dc + nq;
    hipStream_t s = ctx->stream;
    bool ok = (!size || hipMemcpyAsync(dp, row_to_col, (size_t)size * 4, hipMemcpyHostToDevice, s) == hipSuccess) &&
              hipMemcpyAsync(dr, qr.data(), (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dc, qc.data(), (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) ok = lcsq_launch_dominance(dp, size, dr, dc, nq, dout, s) == 0;
    if (ok) ok = hipMemcpyAsync(out, dout, (size_t)nq * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
                 hipStreamSynchronize(s) == hipSuccess;
    buf.release(ctx);
    return ok ? LCS_OK : LCS_EIO;
}
}  // namespace

extern "C" int lcs_permutation_dominance(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size, const int32_t *qr,
                                         const int32_t *qc, int32_t nq, int32_t *out)
{
    if (!ctx || size < 0 || nq < 0 || (size && !row_to_col) || (nq && (!qr || !qc || !out))) return LCS_EINVAL;
    if (nq == 0) return LCS_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    std::vector<int32_t> r(qr, qr + nq), c(qc, qc + nq);
    for (int32_t q = 0; q < nq; ++q)
        if (r[(size_t)q] < 0) r[(size_t)q] = 0;
    return run_dominance(ctx, row_to_col, size, r, c, out);
}

extern "C" int lcs_semilocal_string_substring(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n,
                                              const int32_t *qi, const int32_t *qj, int32_t nq, int32_t *out)
{
    if (!ctx || m < 0 || n < 0 || nq < 0 || (m + n && !row_to_col) || (nq && (!qi || !qj || !out))) return LCS_EINVAL;
    if (nq == 0) return LCS_OK;
    for (int32_t q = 0; q < nq; ++q)
        if (qi[q] < 0 || qi[q] > qj[q] || qj[q] > n) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    std::vector<int32_t> r((size_t)nq), c(qj, qj + nq);
    for (int32_t q = 0; q < nq; ++q) r[(size_t)q] = m + qi[q];
    int rc = run_dominance(ctx, row_to_col, m + n, r, c, out);
    if (rc) return rc;
    for (int32_t q = 0; q < nq; ++q) out[q] = (qj[q] - qi[q]) - out[q];
    return LCS_OK;
}
