// lcs_batch_comb: many pairs of uneven lengths in one persistent launch per
// tile height, the launches side by side on their own streams and captured
// in one hipGraph.  No counterpart in the reference, which combs one pair
// per call (src/main.cpp:46-64); each pair's result is the same as
// lcs_comb's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"

namespace lcs {

// A batch is grouped by tile height and variant (one persistent launch per
// class): classes 0..5 register chain, 6..11 LDS groups.
constexpr int kTileClasses[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumTiles = 6;
constexpr int kNumClasses = 2 * kNumTiles;

struct Extras {
    hipStream_t gstream[kNumClasses] = {};
    hipEvent_t fork = nullptr, join[kNumClasses] = {};
    DevBuf gdescs[kNumClasses];
    DevBuf batch_a, batch_b, batch_h, batch_v, batch_ring;
};

static int class_of(int tm, int group_w)
{
    for (int i = 0; i < kNumTiles; ++i)
        if (kTileClasses[i] == tm) return group_w == 1 ? i : group_w == 4 ? kNumTiles + i : -1;
    return -1;
}

Extras *get_extras(lcs_ctx *ctx)
{
    if (ctx->extras) return ctx->extras;
    Extras *X = new (std::nothrow) Extras();
    if (!X) return nullptr;
    bool ok = hipEventCreateWithFlags(&X->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < kNumClasses; ++i)
        ok = hipStreamCreateWithFlags(&X->gstream[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&X->join[i], hipEventDisableTiming) == hipSuccess;
    ctx->extras = X;
    if (!ok) {
        destroy_extras(ctx);
        return nullptr;
    }
    return X;
}

void destroy_extras(lcs_ctx *ctx)
{
    Extras *X = ctx->extras;
    if (!X) return;
    for (int i = 0; i < kNumClasses; ++i) {
        if (X->gstream[i]) {
            (void)hipStreamSynchronize(X->gstream[i]);
            (void)hipStreamDestroy(X->gstream[i]);
        }
        if (X->join[i]) (void)hipEventDestroy(X->join[i]);
        X->gdescs[i].release(ctx);
    }
    if (X->fork) (void)hipEventDestroy(X->fork);
    X->batch_a.release(ctx);
    X->batch_b.release(ctx);
    X->batch_h.release(ctx);
    X->batch_v.release(ctx);
    X->batch_ring.release(ctx);
    delete X;
    ctx->extras = nullptr;
}

// One band covers a short string; long strings take 16 rows per lane.
int default_batch_tile_m(int32_t m)
{
    if (m <= 64) return 1;
    if (m <= 128) return 2;
    if (m <= 256) return 4;
    if (m <= 512) return 8;
    return 16;
}

// Runs the prepared launches side by side (fork/join over the class
// streams), optionally captured in a hipGraph, between ev0 and ev1 on the
// context stream.
int run_launches(lcs_ctx *ctx, Extras *X, const std::vector<int> &classes,
                 const std::vector<PreparedLaunch> &launches, bool use_graph)
{
    hipStream_t s = ctx->stream;
    // timing brackets the kernels only: in the graph path ev0 is recorded
    // after instantiation, right before the graph launch
    if (!use_graph && hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    hipGraph_t graph = nullptr;
    if (use_graph && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return LCS_EIO;
    int rc = LCS_OK;
    if (hipEventRecord(X->fork, s) != hipSuccess) rc = LCS_EIO;
    for (size_t i = 0; !rc && i < classes.size(); ++i) {
        const int c = classes[i];
        if (hipStreamWaitEvent(X->gstream[c], X->fork, 0) != hipSuccess) { rc = LCS_EIO; break; }
        rc = launch_prepared(launches[i], X->gstream[c]);
        if (!rc && hipEventRecord(X->join[c], X->gstream[c]) != hipSuccess) rc = LCS_EIO;
        if (!rc && hipStreamWaitEvent(s, X->join[c], 0) != hipSuccess) rc = LCS_EIO;
    }
    if (use_graph) {
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (!rc && e != hipSuccess) rc = LCS_EIO;
        if (!rc) {
            hipGraphExec_t exec = nullptr;
            if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = LCS_EIO;
            else {
                if (hipEventRecord(ctx->ev0, s) != hipSuccess || hipGraphLaunch(exec, s) != hipSuccess) rc = LCS_EIO;
                if (!rc && hipEventRecord(ctx->ev1, s) != hipSuccess) rc = LCS_EIO;
                if (hipStreamSynchronize(s) != hipSuccess) rc = LCS_EIO;
                (void)hipGraphExecDestroy(exec);
            }
        }
        if (graph) (void)hipGraphDestroy(graph);
    }
    if (!use_graph && !rc && hipEventRecord(ctx->ev1, s) != hipSuccess) rc = LCS_EIO;
    return rc;
}

}  // namespace lcs

using namespace lcs;

namespace {

// Groups the pairs by (tile height, variant), prepares one persistent launch
// per class and runs them (captured in a hipGraph when use_graph).  All
// arrays of symbols and strands are device pointers; offsets, lengths and
// schedules are host arrays.
int batch_core(lcs_ctx *ctx, Extras *X, int32_t npairs, const int32_t *da, const int64_t *a_off, const int32_t *m,
               const int32_t *db, const int64_t *b_off, const int32_t *n, const lcs_schedule *scheds,
               int32_t use_graph, int32_t *dh, int32_t *dv)
{
    hipStream_t s = ctx->stream;
    std::vector<std::vector<PairDesc>> groups(kNumClasses);
    std::vector<int> wps(kNumClasses, 1);
    for (int32_t p = 0; p < npairs; ++p) {
        int tm = (scheds && scheds[p].tile_m) ? scheds[p].tile_m : default_batch_tile_m(m[p]);
        const int c = class_of(tm, group_width(scheds ? scheds[p].variant : LCS_VARIANT_AUTO));
        if (c < 0) return LCS_EINVAL;
        if (scheds && scheds[p].waves_per_simd) {
            if (scheds[p].waves_per_simd < 1 || scheds[p].waves_per_simd > 8) return LCS_EINVAL;
            wps[(size_t)c] = std::max(wps[(size_t)c], (int)scheds[p].waves_per_simd);
        }
        PairDesc P{};
        P.a = da + a_off[p];
        P.b = db + b_off[p];
        P.h_out = dh + a_off[p];
        P.v_out = dv + b_off[p];
        P.m = m[p];
        P.n = n[p];
        groups[(size_t)c].push_back(P);
    }
    size_t ring_words = 0;
    for (auto &g : groups) ring_words += pairs_ring_words(g);
    int rc;
    if ((rc = X->batch_ring.reserve(ctx, std::max<size_t>(ring_words, 1) * 8))) return rc;
    unsigned long long *ring = (unsigned long long *)X->batch_ring.ptr;
    std::vector<int> classes;
    std::vector<PreparedLaunch> launches;
    for (int c = 0; c < kNumClasses; ++c) {
        auto &g = groups[(size_t)c];
        if (g.empty()) continue;
        PreparedLaunch L;
        rc = prepare_pairs(ctx, g, kTileClasses[c % kNumTiles], wps[(size_t)c], ring, X->gdescs[c],
                           ctx->ctrl + 2 + 2 * c, L, s, c < kNumTiles ? 1 : 4);
        if (rc) return rc;
        ring += pairs_ring_words(g);
        classes.push_back(c);
        launches.push_back(L);
    }
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    if ((rc = run_launches(ctx, X, classes, launches, use_graph != 0))) return rc;
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_ms = ms;
    unsigned flags[2 * kNumClasses + 2] = {};
    if (hipMemcpy(flags, ctx->ctrl, sizeof(flags), hipMemcpyDeviceToHost) != hipSuccess) return LCS_EIO;
    for (int c : classes)
        if (flags[2 + 2 * c + 1]) return LCS_ETIMEDOUT;
    return LCS_OK;
}

// Schedules of a batch's pairs: those left to the library come from the
// context's batch-mode selection (table, model, autotune on a miss,
// default).  Host strings give the model each pair's alphabet.
int batch_schedules(lcs_ctx *ctx, int32_t npairs, const int32_t *m, const int32_t *n, const int32_t *a_host,
                    const int64_t *a_off, const lcs_schedule *scheds, std::vector<lcs_schedule> &out)
{
    const Selector &S = ctx->select[LCS_TUNE_BATCH];
    if (S.table.empty() && !S.model && S.miss_cands.empty()) {   // nothing to choose with: keep the caller's
        out.assign((size_t)npairs, lcs_schedule{});
        if (scheds) std::copy(scheds, scheds + npairs, out.begin());
        return LCS_OK;
    }
    std::vector<int32_t> alphabet((size_t)npairs, 4);
    if (a_host && (S.model || !S.miss_cands.empty()))
        for (int32_t p = 0; p < npairs; ++p) alphabet[(size_t)p] = alphabet_of(a_host + a_off[p], m[p]);
    return choose_schedules(ctx, LCS_TUNE_BATCH, npairs, m, n, alphabet.data(), scheds, out);
}

int check_batch_args(int32_t npairs, const int64_t *a_off, const int32_t *m, const int64_t *b_off,
                     const int32_t *n, int64_t &total_a, int64_t &total_b)
{
    if (!a_off || !m || !b_off || !n) return LCS_EINVAL;
    total_a = total_b = 0;
    for (int32_t p = 0; p < npairs; ++p) {
        if (m[p] < 0 || n[p] < 0 || a_off[p] < 0 || b_off[p] < 0) return LCS_EINVAL;
        if ((long long)m[p] + n[p] > 0x7fffffffLL) return LCS_ERANGE;
        total_a = std::max<int64_t>(total_a, a_off[p] + m[p]);
        total_b = std::max<int64_t>(total_b, b_off[p] + n[p]);
    }
    return LCS_OK;
}

}  // namespace

extern "C" int lcs_batch_comb(lcs_ctx *ctx, int32_t npairs, const int32_t *a_all, const int64_t *a_off,
                              const int32_t *m, const int32_t *b_all, const int64_t *b_off, const int32_t *n,
                              const lcs_schedule *scheds, int32_t use_graph, int32_t *h_out, int32_t *v_out)
{
    if (!ctx || npairs < 0) return LCS_EINVAL;
    if (npairs == 0) return LCS_OK;
    int64_t total_a = 0, total_b = 0;
    int rc = check_batch_args(npairs, a_off, m, b_off, n, total_a, total_b);
    if (rc) return rc;
    if ((total_a && (!a_all || !h_out)) || (total_b && (!b_all || !v_out))) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    // before any batch buffer is touched: an autotune on a miss runs batches itself
    std::vector<lcs_schedule> resolved;
    if ((rc = batch_schedules(ctx, npairs, m, n, a_all, a_off, scheds, resolved))) return rc;
    Extras *X = get_extras(ctx);
    if (!X) return LCS_ENOMEM;
    if ((rc = X->batch_a.reserve(ctx, (size_t)std::max<int64_t>(total_a, 1) * 4))) return rc;
    if ((rc = X->batch_b.reserve(ctx, (size_t)std::max<int64_t>(total_b, 1) * 4))) return rc;
    if ((rc = X->batch_h.reserve(ctx, (size_t)std::max<int64_t>(total_a, 1) * 4))) return rc;
    if ((rc = X->batch_v.reserve(ctx, (size_t)std::max<int64_t>(total_b, 1) * 4))) return rc;
    int32_t *da = (int32_t *)X->batch_a.ptr, *db = (int32_t *)X->batch_b.ptr;
    int32_t *dh = (int32_t *)X->batch_h.ptr, *dv = (int32_t *)X->batch_v.ptr;
    hipStream_t s = ctx->stream;
    if (total_a && hipMemcpyAsync(da, a_all, (size_t)total_a * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (total_b && hipMemcpyAsync(db, b_all, (size_t)total_b * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if ((rc = batch_core(ctx, X, npairs, da, a_off, m, db, b_off, n, resolved.data(), use_graph, dh, dv))) return rc;
    if (total_a && hipMemcpyAsync(h_out, dh, (size_t)total_a * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (total_b && hipMemcpyAsync(v_out, dv, (size_t)total_b * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    return LCS_OK;
}

extern "C" int lcs_batch_comb_device(lcs_ctx *ctx, int32_t npairs, const int32_t *d_a_all, const int64_t *a_off,
                                     const int32_t *m, const int32_t *d_b_all, const int64_t *b_off,
                                     const int32_t *n, const lcs_schedule *scheds, int32_t use_graph,
                                     int32_t *d_h_out, int32_t *d_v_out)
{
    if (!ctx || npairs < 0) return LCS_EINVAL;
    if (npairs == 0) return LCS_OK;
    int64_t total_a = 0, total_b = 0;
    int rc = check_batch_args(npairs, a_off, m, b_off, n, total_a, total_b);
    if (rc) return rc;
    if ((total_a && (!d_a_all || !d_h_out)) || (total_b && (!d_b_all || !d_v_out))) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    std::vector<lcs_schedule> resolved;
    if ((rc = batch_schedules(ctx, npairs, m, n, nullptr, a_off, scheds, resolved))) return rc;
    Extras *X = get_extras(ctx);
    if (!X) return LCS_ENOMEM;
    return batch_core(ctx, X, npairs, d_a_all, a_off, m, d_b_all, b_off, n, resolved.data(), use_graph, d_h_out,
                      d_v_out);
}
