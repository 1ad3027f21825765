// lcs_comb_fixup: the reference's residual fixup (Lcs_Semi_Fixup,
// src/lcs_residual_fixup.hpp:74-122) as a GPU operation on the band chain
// with explicit initial strands.
#include <hip/hip_runtime.h>

#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"

using namespace lcs;

extern "C" int lcs_comb_fixup(lcs_ctx *ctx, const int32_t *a_prepared, int32_t m, const int32_t *b, int32_t n,
                              int32_t *h, int32_t *v, int32_t i0, int32_t isize, int32_t j0, int32_t jsize)
{
    if (!ctx || m < 0 || n < 0 || i0 < 0 || j0 < 0 || isize < 0 || jsize < 0) return LCS_EINVAL;
    if ((long long)i0 + isize > m || (long long)j0 + jsize > n) return LCS_EINVAL;
    if (isize == 0 || jsize == 0) return LCS_OK;  // the reference returns early too
    if (!a_prepared || !b || !h || !v) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    // Device row r of the sub-rectangle is the reference's row i0+isize-1-r
    // (rows are stored reversed in a_prepared).
    std::vector<int32_t> a_rows((size_t)isize);
    for (int32_t r = 0; r < isize; ++r) a_rows[(size_t)r] = a_prepared[i0 + isize - 1 - r];
    int rc;
    const size_t bytes = ((size_t)isize * 3 + (size_t)jsize * 3) * 4;
    DevBuf tmp;
    if ((rc = tmp.reserve(ctx, bytes))) return rc;
    int32_t *da = (int32_t *)tmp.ptr, *dhin = da + isize, *dhout = dhin + isize;
    int32_t *db = dhout + isize, *dvin = db + jsize, *dvout = dvin + jsize;
    hipStream_t s = ctx->stream;
    bool ok = hipMemcpyAsync(da, a_rows.data(), (size_t)isize * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dhin, h + i0, (size_t)isize * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(db, b + j0, (size_t)jsize * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dvin, v + j0, (size_t)jsize * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (!ok) { tmp.release(ctx); return LCS_EIO; }
    rc = enqueue_chain(ctx, da, isize, db, jsize, nullptr, dhout, dvout, s, dhin, dvin);
    if (!rc) {
        ok = hipMemcpyAsync(h + i0, dhout, (size_t)isize * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(v + j0, dvout, (size_t)jsize * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
        rc = ok ? check_abort(ctx) : LCS_EIO;
    }
    tmp.release(ctx);
    return rc;
}
