// Semi-local score read-out on the GPU (declared in include/lcs_hip.h): the
// four quadrant queries and the device-resident dominance index.
#include <hip/hip_runtime.h>

#include <climits>
#include <new>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"
#include "query_kernels.h"

using namespace lcs;

struct lcs_dominance_index {
    lcs_ctx *ctx = nullptr;
    int32_t size = 0;
    int levels = 0;
    DevBuf mem;   // RankWord words[levels * W], then uint32 zeros[levels]
    DevBuf io;    // query staging for the host entry point
    lcsq::RankWord *words() const { return (lcsq::RankWord *)mem.ptr; }
    uint32_t *zeros() const { return (uint32_t *)((char *)mem.ptr + words_bytes()); }
    size_t words_bytes() const { return (size_t)levels * (((size_t)size >> 6) + 1) * sizeof(lcsq::RankWord); }
};

namespace {

int bit_length(int32_t x)
{
    int l = 0;
    while (l < 31 && ((int64_t)1 << l) <= x) ++l;
    return l;
}

bool valid_kind(int32_t kind) { return kind >= LCS_SEMI_STRING_SUBSTRING && kind <= LCS_SEMI_DOMINANCE; }

// Host-side validation of a query set (kind, m, n against size).
int check_queries(int32_t m, int32_t n, int32_t size, int32_t kind, const int32_t *q0, const int32_t *q1, int32_t nq)
{
    if (!valid_kind(kind) || m < 0 || n < 0 || (int64_t)m + n != size) return LCS_EINVAL;
    lcsq::Window w;
    for (int32_t q = 0; q < nq; ++q)
        if (!lcsq::query_window(kind, m, n, q0[q], q1[q], w)) return LCS_EINVAL;
    return LCS_OK;
}

int build_index(lcs_ctx *ctx, const int32_t *src, bool src_on_host, int32_t size, lcs_dominance_index **out)
{
    auto *idx = new (std::nothrow) lcs_dominance_index;
    if (!idx) return LCS_ENOMEM;
    idx->ctx = ctx;
    idx->size = size;
    idx->levels = size > 0 ? bit_length(size) : 0;
    hipStream_t s = ctx->stream;
    int rc = LCS_OK;
    if (size > 0) {
        const size_t W = ((size_t)size >> 6) + 1, nb = (W + 63) / 64;
        rc = idx->mem.reserve(ctx, idx->words_bytes() + (size_t)idx->levels * 4);
        DevBuf scratch;
        if (!rc) rc = scratch.reserve(ctx, (2 * (size_t)size + nb + 1 + (src_on_host ? (size_t)size : 0)) * 4);
        if (!rc) {
            int32_t *sc = (int32_t *)scratch.ptr;
            int32_t *bad = sc + 2 * (size_t)size + nb;
            const int32_t *perm = src;
            bool ok = hipMemsetAsync(bad, 0, 4, s) == hipSuccess;
            if (ok && src_on_host) {
                int32_t *staged = bad + 1;
                ok = hipMemcpyAsync(staged, src, (size_t)size * 4, hipMemcpyHostToDevice, s) == hipSuccess;
                perm = staged;
            }
            if (ok) ok = lcsq_launch_index_build(perm, size, idx->levels, idx->words(), idx->zeros(), sc, s) == 0;
            int32_t flag = 0;
            if (ok)
                ok = hipMemcpyAsync(&flag, bad, 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
                     hipStreamSynchronize(s) == hipSuccess;
            rc = !ok ? LCS_EIO : flag ? LCS_EINVAL : LCS_OK;
        }
        scratch.release(ctx);
    }
    if (rc) {
        lcs_dominance_index_destroy(idx);
        return rc;
    }
    *out = idx;
    return LCS_OK;
}

// Host queries through an index: stage, launch, copy back, synchronise.
int query_index_host(lcs_dominance_index *idx, int32_t m, int32_t n, int32_t kind, const int32_t *q0,
                     const int32_t *q1, int32_t nq, int32_t *out)
{
    lcs_ctx *ctx = idx->ctx;
    int rc = idx->io.reserve(ctx, 3 * (size_t)nq * 4);
    if (rc) return rc;
    int32_t *d0 = (int32_t *)idx->io.ptr, *d1 = d0 + nq, *dout = d1 + nq;
    hipStream_t s = ctx->stream;
    bool ok = hipMemcpyAsync(d0, q0, (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d1, q1, (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok)
        ok = lcsq_launch_index_query(idx->words(), idx->zeros(), idx->levels, idx->size, kind, m, n, d0, d1, nq, dout,
                                     s) == 0;
    if (ok)
        ok = hipMemcpyAsync(out, dout, (size_t)nq * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    return ok ? LCS_OK : LCS_EIO;
}

// Host queries by a direct count over the permutation (no set-up).
int query_scan_host(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind, const int32_t *q0,
                    const int32_t *q1, int32_t nq, int32_t *out)
{
    const int32_t size = m + n;
    DevBuf buf;
    int rc = buf.reserve(ctx, ((size_t)size + 3 * (size_t)nq + 4) * 4);
    if (rc) return rc;
    int32_t *dp = (int32_t *)buf.ptr, *d0 = dp + size, *d1 = d0 + nq, *dout = d1 + nq;
    hipStream_t s = ctx->stream;
    bool ok = (!size || hipMemcpyAsync(dp, row_to_col, (size_t)size * 4, hipMemcpyHostToDevice, s) == hipSuccess) &&
              hipMemcpyAsync(d0, q0, (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d1, q1, (size_t)nq * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (ok) ok = lcsq_launch_scan(dp, size, kind, m, n, d0, d1, nq, dout, s) == 0;
    if (ok)
        ok = hipMemcpyAsync(out, dout, (size_t)nq * 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    buf.release(ctx);
    return ok ? LCS_OK : LCS_EIO;
}

// A direct count reads every row per query (about 30 us per query at 2^21
// rows on MI355X); the index costs 0.3-0.6 ms to build and then answers
// billions of queries per second.  Past ~2^25 scanned rows the index wins.
bool prefer_index(int32_t size, int32_t nq) { return (double)size * nq > (double)(1 << 25); }

int semilocal_host(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind, const int32_t *q0,
                   const int32_t *q1, int32_t nq, int32_t *out)
{
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    const int32_t size = m + n;
    if (!prefer_index(size, nq)) return query_scan_host(ctx, row_to_col, m, n, kind, q0, q1, nq, out);
    lcs_dominance_index *idx = nullptr;
    int rc = build_index(ctx, row_to_col, true, size, &idx);
    if (rc == LCS_EINVAL)   // values outside [-1, size): the direct count takes any int32
        return query_scan_host(ctx, row_to_col, m, n, kind, q0, q1, nq, out);
    if (rc) return rc;
    rc = query_index_host(idx, m, n, kind, q0, q1, nq, out);
    lcs_dominance_index_destroy(idx);
    return rc;
}

}  // namespace

extern "C" int lcs_semilocal_query(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind,
                                   const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out)
{
    if (!ctx || m < 0 || n < 0 || (int64_t)m + n > INT32_MAX || nq < 0 || (m + n && !row_to_col) ||
        (nq && (!q0 || !q1 || !out)))
        return LCS_EINVAL;
    int rc = check_queries(m, n, m + n, kind, q0, q1, nq);
    if (rc) return rc;
    if (nq == 0) return LCS_OK;
    return semilocal_host(ctx, row_to_col, m, n, kind, q0, q1, nq, out);
}

extern "C" int lcs_permutation_dominance(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size, const int32_t *qr,
                                         const int32_t *qc, int32_t nq, int32_t *out)
{
    if (!ctx || size < 0 || nq < 0 || (size && !row_to_col) || (nq && (!qr || !qc || !out))) return LCS_EINVAL;
    if (nq == 0) return LCS_OK;
    return semilocal_host(ctx, row_to_col, size, 0, LCS_SEMI_DOMINANCE, qr, qc, nq, out);
}

extern "C" int lcs_semilocal_string_substring(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n,
                                              const int32_t *qi, const int32_t *qj, int32_t nq, int32_t *out)
{
    return lcs_semilocal_query(ctx, row_to_col, m, n, LCS_SEMI_STRING_SUBSTRING, qi, qj, nq, out);
}

extern "C" int lcs_dominance_index_build(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size,
                                         lcs_dominance_index **out)
{
    if (!ctx || !out || size < 0 || (size && !row_to_col)) return LCS_EINVAL;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    return build_index(ctx, row_to_col, true, size, out);
}

extern "C" int lcs_dominance_index_build_device(lcs_ctx *ctx, const int32_t *d_row_to_col, int32_t size,
                                                lcs_dominance_index **out)
{
    if (!ctx || !out || size < 0 || (size && !d_row_to_col)) return LCS_EINVAL;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    return build_index(ctx, d_row_to_col, false, size, out);
}

extern "C" int lcs_dominance_index_query(lcs_dominance_index *idx, int32_t m, int32_t n, int32_t kind,
                                         const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out)
{
    if (!idx || nq < 0 || (nq && (!q0 || !q1 || !out))) return LCS_EINVAL;
    int rc = check_queries(m, n, idx->size, kind, q0, q1, nq);
    if (rc) return rc;
    if (nq == 0) return LCS_OK;
    if (hipSetDevice(idx->ctx->device) != hipSuccess) return LCS_EIO;
    return query_index_host(idx, m, n, kind, q0, q1, nq, out);
}

extern "C" int lcs_dominance_index_query_device(lcs_dominance_index *idx, int32_t m, int32_t n, int32_t kind,
                                                const int32_t *d_q0, const int32_t *d_q1, int32_t nq, int32_t *d_out)
{
    if (!idx || nq < 0 || (nq && (!d_q0 || !d_q1 || !d_out))) return LCS_EINVAL;
    if (!valid_kind(kind) || m < 0 || n < 0 || (int64_t)m + n != idx->size) return LCS_EINVAL;
    if (nq == 0) return LCS_OK;
    if (hipSetDevice(idx->ctx->device) != hipSuccess) return LCS_EIO;
    return lcsq_launch_index_query(idx->words(), idx->zeros(), idx->levels, idx->size, kind, m, n, d_q0, d_q1, nq,
                                   d_out, idx->ctx->stream) == 0
               ? LCS_OK
               : LCS_EIO;
}

extern "C" int64_t lcs_dominance_index_bytes(const lcs_dominance_index *idx)
{
    return idx ? (int64_t)idx->mem.bytes : 0;
}

extern "C" void lcs_dominance_index_destroy(lcs_dominance_index *idx)
{
    if (!idx) return;
    if (idx->mem.ptr || idx->io.ptr) {
        (void)hipSetDevice(idx->ctx->device);
        (void)hipStreamSynchronize(idx->ctx->stream);
    }
    idx->mem.release(idx->ctx);
    idx->io.release(idx->ctx);
    delete idx;
}
