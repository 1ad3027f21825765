// C ABI of the engine: contexts, single-pair combing, permutation helpers.
// Declarations and the reference interfaces each entry replaces are in
// include/lcs_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "lcs_hip.h"
#include "comb_kernels.h"
#include "lcs_internal.h"

namespace lcs {

static int hip_status(hipError_t e) { return e == hipSuccess ? LCS_OK : LCS_EIO; }

int DevBuf::reserve(lcs_ctx *ctx, size_t want)
{
    if (want <= bytes) return LCS_OK;
    size_t grow = std::max(want, bytes + bytes / 2);
    if (ctx->held + grow - bytes > ctx->budget) grow = want;
    if (ctx->held + grow - bytes > ctx->budget) return LCS_ENOMEM;
    if (ptr) {
        (void)hipFree(ptr);
        ctx->held -= bytes;
        ptr = nullptr;
        bytes = 0;
    }
    void *p = nullptr;
    if (hipMalloc(&p, grow) != hipSuccess) {
        (void)hipGetLastError();
        return LCS_ENOMEM;
    }
    ptr = p;
    bytes = grow;
    ctx->held += grow;
    return LCS_OK;
}

void DevBuf::release(lcs_ctx *ctx)
{
    if (ptr) {
        (void)hipFree(ptr);
        ctx->held -= bytes;
    }
    ptr = nullptr;
    bytes = 0;
}

// Default tile height: the largest power of two <= 16 that still gives every
// SIMD of the device at least one band of 64*tile_m rows.
int default_tile_m(const lcs_ctx *ctx, int32_t m)
{
    const long long simds = 4LL * ctx->num_cus;
    int tm = 16;
    while (tm > 1 && (long long)m < simds * 64LL * tm) tm >>= 1;
    return tm;
}

static bool valid_tile_m(int tm) { return tm == 1 || tm == 2 || tm == 4 || tm == 8 || tm == 16 || tm == 32; }

// Enqueue one single-pair chain on `stream` with device pointers.
int enqueue_chain(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                  const lcs_schedule *sched, int32_t *d_h, int32_t *d_v, hipStream_t stream,
                  const int32_t *d_h_in, const int32_t *d_v_in)
{
    if (m == 0 || n == 0) {
        if (d_h_in || d_v_in) {
            if (m && hipMemcpyAsync(d_h, d_h_in, (size_t)m * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess) return LCS_EIO;
            if (n && hipMemcpyAsync(d_v, d_v_in, (size_t)n * 4, hipMemcpyDeviceToDevice, stream) != hipSuccess) return LCS_EIO;
            return LCS_OK;
        }
        return lcsk_launch_identity(d_h, m, d_v, n, stream);
    }
    int tm = (sched && sched->tile_m) ? sched->tile_m : default_tile_m(ctx, m);
    if (!valid_tile_m(tm)) return LCS_EINVAL;
    int wps = (sched && sched->waves_per_simd) ? sched->waves_per_simd : 1;
    if (wps < 1 || wps > 8) return LCS_EINVAL;
    const long long rows_per_band = 64LL * tm;
    const long long nbands = ((long long)m + rows_per_band - 1) / rows_per_band;
    int rc = ctx->ring.reserve(ctx, (2ull * (size_t)n + 64) * sizeof(unsigned long long));
    if (rc) return rc;
    ChainArgs A{};
    A.a = d_a;
    A.b = d_b;
    A.m = m;
    A.n = n;
    A.nbands = (int32_t)nbands;
    A.h_out = d_h;
    A.v_out = d_v;
    A.h_in = d_h_in;
    A.v_in = d_v_in;
    A.E = (unsigned long long *)ctx->ring.ptr;
    A.ticket = ctx->ctrl;
    A.abort_flag = ctx->ctrl + 1;
    A.timeout_ticks = ctx->timeout_ticks;
    A.trace = ctx->band_trace;
    A.trace_cap = ctx->band_trace_cap;
    const int waves_per_block = 4;
    long long waves = std::min<long long>(nbands, 4LL * ctx->num_cus * wps);
    int blocks = (int)((waves + waves_per_block - 1) / waves_per_block);
    return lcsk_launch_chain(A, tm, blocks, 64 * waves_per_block, stream);
}

int check_abort(lcs_ctx *ctx)
{
    unsigned flag = 0;
    if (hipMemcpy(&flag, ctx->ctrl + 1, sizeof(flag), hipMemcpyDeviceToHost) != hipSuccess) return LCS_EIO;
    return flag ? LCS_ETIMEDOUT : LCS_OK;
}

}  // namespace lcs

using namespace lcs;

extern "C" {

int lcs_ctx_create(int device, size_t hbm_budget, lcs_ctx **out)
{
    if (!out) return LCS_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return LCS_EIO;
    }
    if (device < 0 || device >= ndev) return LCS_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LCS_EIO;
    lcs_ctx *ctx = new (std::nothrow) lcs_ctx();
    if (!ctx) return LCS_ENOMEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return LCS_EIO; }
    ctx->num_cus = prop.multiProcessorCount;
    ctx->name = prop.name;
    ctx->arch = prop.gcnArchName;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { delete ctx; return LCS_EIO; }
    ctx->budget = hbm_budget ? hbm_budget : free_b / 4 * 3;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return LCS_EIO; }
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        lcs_ctx_destroy(ctx);
        return LCS_EIO;
    }
    void *p = nullptr;
    if (hipMalloc(&p, 256) != hipSuccess) { lcs_ctx_destroy(ctx); return LCS_ENOMEM; }
    ctx->ctrl = (unsigned *)p;
    if (hipMemset(ctx->ctrl, 0, 256) != hipSuccess) { lcs_ctx_destroy(ctx); return LCS_EIO; }
    *out = ctx;
    return LCS_OK;
}

int lcs_ctx_destroy(lcs_ctx *ctx)
{
    if (!ctx) return LCS_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    destroy_extras(ctx);
    ctx->ring.release(ctx);
    ctx->in_a.release(ctx);
    ctx->in_b.release(ctx);
    ctx->out_h.release(ctx);
    ctx->out_v.release(ctx);
    if (ctx->ctrl) (void)hipFree(ctx->ctrl);
    if (ctx->band_trace) (void)hipFree(ctx->band_trace);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LCS_OK;
}

int lcs_ctx_device_name(lcs_ctx *ctx, char *buf, size_t len)
{
    if (!ctx || !buf || len == 0) return LCS_EINVAL;
    std::snprintf(buf, len, "%s (%s)", ctx->name.c_str(), ctx->arch.c_str());
    return LCS_OK;
}

int lcs_ctx_info(lcs_ctx *ctx, int32_t *num_cus, double *last_kernel_ms)
{
    if (!ctx) return LCS_EINVAL;
    if (ctx->pending_timing && hipEventQuery(ctx->ev1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_ms = ms;
        ctx->pending_timing = false;
    }
    if (num_cus) *num_cus = ctx->num_cus;
    if (last_kernel_ms) *last_kernel_ms = ctx->last_kernel_ms;
    return LCS_OK;
}

int lcs_comb(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
             const lcs_schedule *sched, int32_t *h_out, int32_t *v_out, int32_t *llcs_out)
{
    if (!ctx || m < 0 || n < 0) return LCS_EINVAL;
    if ((m > 0 && (!a || !h_out)) || (n > 0 && (!b || !v_out))) return LCS_EINVAL;
    if ((long long)m + n > 0x7fffffffLL) return LCS_ERANGE;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    int rc;
    if ((rc = ctx->in_a.reserve(ctx, (size_t)std::max(m, 1) * 4))) return rc;
    if ((rc = ctx->in_b.reserve(ctx, (size_t)std::max(n, 1) * 4))) return rc;
    if ((rc = ctx->out_h.reserve(ctx, (size_t)std::max(m, 1) * 4))) return rc;
    if ((rc = ctx->out_v.reserve(ctx, (size_t)std::max(n, 1) * 4))) return rc;
    int32_t *da = (int32_t *)ctx->in_a.ptr, *db = (int32_t *)ctx->in_b.ptr;
    int32_t *dh = (int32_t *)ctx->out_h.ptr, *dv = (int32_t *)ctx->out_v.ptr;
    hipStream_t s = ctx->stream;
    if (m && hipMemcpyAsync(da, a, (size_t)m * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (n && hipMemcpyAsync(db, b, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    if ((rc = enqueue_chain(ctx, da, m, db, n, sched, dh, dv, s))) return rc;
    if (hipEventRecord(ctx->ev1, s) != hipSuccess) return LCS_EIO;
    if (m && hipMemcpyAsync(h_out, dh, (size_t)m * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (n && hipMemcpyAsync(v_out, dv, (size_t)n * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_ms = ms;
    if ((rc = check_abort(ctx))) return rc;
    if (llcs_out) {
        int32_t c = 0;
        for (int32_t j = 0; j < n; ++j) c += v_out[j] < m;
        *llcs_out = c;
    }
    return LCS_OK;
}

int lcs_comb_device(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                    const lcs_schedule *sched, int32_t *d_h_out, int32_t *d_v_out, void *stream)
{
    if (!ctx || m < 0 || n < 0) return LCS_EINVAL;
    if ((m > 0 && (!d_a || !d_h_out)) || (n > 0 && (!d_b || !d_v_out))) return LCS_EINVAL;
    if ((long long)m + n > 0x7fffffffLL) return LCS_ERANGE;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    int rc = enqueue_chain(ctx, d_a, m, d_b, n, sched, d_h_out, d_v_out, s);
    if (rc) return rc;
    if (hipEventRecord(ctx->ev1, s) != hipSuccess) return LCS_EIO;
    ctx->pending_timing = true;
    return LCS_OK;
}

int lcs_debug_band_trace(lcs_ctx *ctx, int32_t max_bands, uint64_t *host_out)
{
    // max_bands > 0: (re)arm tracing of up to max_bands bands; 0: disarm;
    // host_out != NULL: copy the last trace (2 words per band) out.
    if (!ctx || max_bands < 0) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    if (host_out) {
        if (!ctx->band_trace || max_bands > ctx->band_trace_cap) return LCS_EINVAL;
        if (hipMemcpy(host_out, ctx->band_trace, (size_t)max_bands * 16, hipMemcpyDeviceToHost) != hipSuccess)
            return LCS_EIO;
        return LCS_OK;
    }
    if (ctx->band_trace) {
        (void)hipFree(ctx->band_trace);
        ctx->band_trace = nullptr;
        ctx->band_trace_cap = 0;
    }
    if (max_bands == 0) return LCS_OK;
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)max_bands * 16) != hipSuccess) return LCS_ENOMEM;
    ctx->band_trace = (unsigned long long *)p;
    ctx->band_trace_cap = max_bands;
    return LCS_OK;
}

int lcs_permutation_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n,
                                 int32_t *row_to_col, int32_t *col_to_row)
{
    if (m < 0 || n < 0 || (m && !h) || (n && !v)) return LCS_EINVAL;
    const int32_t size = m + n;
    if (size && (!row_to_col || !col_to_row)) return LCS_EINVAL;
    for (int32_t i = 0; i < size; ++i) row_to_col[i] = col_to_row[i] = -1;
    for (int32_t l = 0; l < m; ++l)
        if (h[l] >= 0 && h[l] < size) { row_to_col[h[l]] = n + l; col_to_row[n + l] = h[l]; }
    for (int32_t j = 0; j < n; ++j)
        if (v[j] >= 0 && v[j] < size) { row_to_col[v[j]] = j; col_to_row[j] = v[j]; }
    return LCS_OK;
}

int64_t lcs_permutation_hash(const int32_t *col_to_row, int32_t size)
{
    // result = (R * result + row) % M in 64-bit two's complement, the wrap made
    // explicit with unsigned arithmetic.
    const uint64_t R = 4294967279ull;
    const int64_t M = 4294967291ll;
    int64_t res = 0;
    for (int32_t c = 0; c < size; ++c) {
        const uint64_t t = R * (uint64_t)res + (uint64_t)(int64_t)col_to_row[c];
        res = (int64_t)t % M;
    }
    return res;
}

}  // extern "C"
