// Score-only LCS length (bit-parallel kernel) behind the C ABI; declarations
// in include/lcs_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"
#include "llcs_kernels.h"

using namespace lcs;

namespace {

struct LlcsPair {
    const int32_t *a, *b;   // device
    int32_t m, n;
    int32_t *out;           // device int32
};

// Steps from the first band's start to the last band's end: a band lags the
// one below by about kLag steps (63 of skew, two chunks of hand-off, the
// fetch round trip), and bands beyond the resident ones run in later rounds.
constexpr double kLag = 160.0;

double pipeline_steps(int64_t rows, int64_t cols, int64_t resident)
{
    const int64_t bands = (rows + lcsl::kBandRows - 1) / lcsl::kBandRows;
    const int64_t rounds = (bands + resident - 1) / std::max<int64_t>(resident, 1);
    return (double)rounds * (double)(cols + lcsl::kWave - 1) + (double)bands * kLag;
}

// bands in flight: a band is a workgroup of two waves (comb + fetch), four per CU
int resident_bands(const lcs_ctx *ctx) { return 4 * ctx->num_cus; }

// Enqueues one launch over `pairs` on `s`: the longer-pipelined orientation
// is chosen per pair (LCS is symmetric).  Rings and descriptors live in the
// context's buffers, error words after the descriptors.
static int enqueue_llcs_ws(lcs_ctx *ctx, const std::vector<LlcsPair> &pairs, int32_t sigma, hipStream_t s);

int enqueue_llcs(lcs_ctx *ctx, const std::vector<LlcsPair> &pairs, int32_t sigma, hipStream_t s)
{
    int rc = ws_acquire(ctx, s);
    if (rc) return rc;
    rc = enqueue_llcs_ws(ctx, pairs, sigma, s);
    const int rc2 = ws_release(ctx, s);
    return rc ? rc : rc2;
}

static int enqueue_llcs_ws(lcs_ctx *ctx, const std::vector<LlcsPair> &pairs, int32_t sigma, hipStream_t s)
{
    const int np = (int)pairs.size();
    if (np == 0) return LCS_OK;
    const int64_t resident = resident_bands(ctx);
    std::vector<lcsl::LlcsDesc> descs((size_t)np);
    size_t ring_words = 0;
    int max_ring = 0;
    for (int p = 0; p < np; ++p) {
        const LlcsPair &P = pairs[(size_t)p];
        const bool swap = pipeline_steps(P.n, P.m, resident) < pipeline_steps(P.m, P.n, resident);
        lcsl::LlcsDesc &D = descs[(size_t)p];
        D = lcsl::LlcsDesc{};
        D.rows = swap ? P.b : P.a;
        D.cols = swap ? P.a : P.b;
        D.nrows = swap ? P.n : P.m;
        D.ncols = swap ? P.m : P.n;
        // the band loop counts columns up to ncols + 3 chunks in int
        if (D.ncols > 0x7fffffff - 4096) return LCS_ERANGE;
        D.nbands = lcsl::llcs_nbands(D.nrows);
        D.nchunks = lcsl::llcs_nchunks(D.ncols);
        D.out = P.out;
        ring_words += 2 * (size_t)D.nchunks;
        max_ring = std::max(max_ring, 2 * D.nchunks);
    }
    int rc = ctx->ring.reserve(ctx, ring_words * sizeof(unsigned long long));
    if (rc) return rc;
    std::vector<int2> tasks;
    for (int p = 0; p < np; ++p)
        for (int32_t g = 0; g < descs[(size_t)p].nbands; ++g) tasks.push_back(make_int2(p, g));
    const size_t desc_bytes = sizeof(lcsl::LlcsDesc) * (size_t)np, task_bytes = sizeof(int2) * tasks.size();
    const size_t err_off = (desc_bytes + task_bytes + 255) / 256 * 256;
    if ((rc = ctx->descs.reserve(ctx, err_off + 4 * (size_t)np + 256))) return rc;
    unsigned long long *ring = (unsigned long long *)ctx->ring.ptr;
    int32_t *err = (int32_t *)((unsigned char *)ctx->descs.ptr + err_off);
    for (int p = 0; p < np; ++p) {
        descs[(size_t)p].ring = ring;
        ring += 2 * (size_t)descs[(size_t)p].nchunks;
        descs[(size_t)p].err = err + p;
    }
    std::vector<unsigned char> blob(desc_bytes + task_bytes);
    std::memcpy(blob.data(), descs.data(), desc_bytes);
    if (task_bytes) std::memcpy(blob.data() + desc_bytes, tasks.data(), task_bytes);
    if ((rc = upload(ctx, ctx->descs.ptr, blob.data(), blob.size(), s))) return rc;
    lcsl::LlcsArgs A{};
    A.descs = (const lcsl::LlcsDesc *)ctx->descs.ptr;
    A.tasks = (const int2 *)((unsigned char *)ctx->descs.ptr + desc_bytes);
    A.ntasks = (int32_t)tasks.size();
    A.sigma = sigma;
    A.ticket = ctx->ctrl;
    A.abort_flag = ctx->ctrl + 1;
    A.sticky = ctx->ctrl + kStickyWord;
    A.timeout_ticks = ctx->timeout_ticks;
    A.trace = ctx->band_trace;
    A.trace_cap = ctx->band_trace_cap;
    const int blocks = (int)std::min<int64_t>((int64_t)tasks.size(), resident);
    return lcsl_launch(A, A.descs, np, max_ring, blocks, s) == 0 ? LCS_OK : LCS_EIO;
}

bool valid_sigma(int32_t sigma) { return sigma >= 1 && sigma <= lcsl::kMaxSigma; }

// Runs host pairs: copies the strings in, launches, copies the lengths out.
int run_host(lcs_ctx *ctx, int32_t npairs, const int32_t *a_all, const int64_t *a_off, const int32_t *m,
             const int32_t *b_all, const int64_t *b_off, const int32_t *n, int32_t sigma, int32_t *out)
{
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    int64_t a_len = 0, b_len = 0;
    for (int32_t p = 0; p < npairs; ++p) {
        a_len = std::max<int64_t>(a_len, a_off[p] + m[p]);
        b_len = std::max<int64_t>(b_len, b_off[p] + n[p]);
    }
    int rc = ctx->in_a.reserve(ctx, (size_t)std::max<int64_t>(a_len, 1) * 4);
    if (!rc) rc = ctx->in_b.reserve(ctx, (size_t)std::max<int64_t>(b_len, 1) * 4);
    if (!rc) rc = ctx->out_h.reserve(ctx, (size_t)std::max(npairs, 1) * 4);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    int32_t *da = (int32_t *)ctx->in_a.ptr, *db = (int32_t *)ctx->in_b.ptr, *dout = (int32_t *)ctx->out_h.ptr;
    if (a_len && hipMemcpyAsync(da, a_all, (size_t)a_len * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (b_len && hipMemcpyAsync(db, b_all, (size_t)b_len * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    std::vector<LlcsPair> pairs((size_t)npairs);
    for (int32_t p = 0; p < npairs; ++p)
        pairs[(size_t)p] = LlcsPair{da + a_off[p], db + b_off[p], m[p], n[p], dout + p};
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    if ((rc = enqueue_llcs(ctx, pairs, sigma, s))) return rc;
    if (hipEventRecord(ctx->ev1, s) != hipSuccess) return LCS_EIO;
    ctx->pending_timing = true;
    if (hipMemcpyAsync(out, dout, (size_t)npairs * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return LCS_EIO;
    if ((rc = check_abort(ctx))) return rc;
    for (int32_t p = 0; p < npairs; ++p)
        if (out[p] < 0) return LCS_ERANGE;
    return LCS_OK;
}

}  // namespace

extern "C" int lcs_llcs(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n, int32_t sigma,
                        int32_t *llcs)
{
    if (!ctx || !llcs || m < 0 || n < 0 || (m && !a) || (n && !b) || !valid_sigma(sigma)) return LCS_EINVAL;
    const int64_t zero = 0;
    return run_host(ctx, 1, a, &zero, &m, b, &zero, &n, sigma, llcs);
}

extern "C" int lcs_llcs_device(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                               int32_t sigma, int32_t *d_llcs)
{
    if (!ctx || !d_llcs || m < 0 || n < 0 || (m && !d_a) || (n && !d_b) || !valid_sigma(sigma)) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    hipStream_t s = ctx->stream;
    std::vector<LlcsPair> pairs{LlcsPair{d_a, d_b, m, n, d_llcs}};
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    int rc = enqueue_llcs(ctx, pairs, sigma, s);
    if (rc) return rc;
    if (hipEventRecord(ctx->ev1, s) != hipSuccess) return LCS_EIO;
    ctx->pending_timing = true;
    return LCS_OK;
}

extern "C" int lcs_llcs_batch(lcs_ctx *ctx, int32_t npairs, const int32_t *a_all, const int64_t *a_off,
                              const int32_t *m, const int32_t *b_all, const int64_t *b_off, const int32_t *n,
                              int32_t sigma, int32_t *out)
{
    if (!ctx || npairs < 0 || !valid_sigma(sigma)) return LCS_EINVAL;
    if (npairs == 0) return LCS_OK;
    if (!a_off || !m || !b_off || !n || !out) return LCS_EINVAL;
    for (int32_t p = 0; p < npairs; ++p)
        if (m[p] < 0 || n[p] < 0 || a_off[p] < 0 || b_off[p] < 0 || (m[p] && !a_all) || (n[p] && !b_all))
            return LCS_EINVAL;
    return run_host(ctx, npairs, a_all, a_off, m, b_all, b_off, n, sigma, out);
}
