// Internal state of a context (shared by the C ABI translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "comb_kernels.h"
#include "lcs_hip.h"

namespace lcs {

// A grow-only device buffer whose bytes count against the context's budget.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(lcs_ctx *ctx, size_t want);
    void release(lcs_ctx *ctx);
};

struct Extras;  // batch / server resources, created on first use

// A chain launch whose descriptors are already on the device.
struct PreparedLaunch {
    ChainArgs args{};
    int blocks = 0, threads = 0, npairs = 0, max_n = 0, tile_m = 0;
};

}  // namespace lcs

namespace lcs {
// ctrl word that collects the abort flags of finished launches (folded in by
// each launch's prep kernel); lcs_ctx_status reads and clears it
constexpr int kStickyWord = 63;
}  // namespace lcs

struct lcs_ctx {
    int device = 0;
    int num_cus = 0;
    std::string name, arch;
    size_t budget = 0;
    size_t held = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned *ctrl = nullptr;            // [0] ticket, [1] abort flag, rest spare
    unsigned long long timeout_ticks = 500000000ull;  // 5 s at 100 MHz
    double last_kernel_ms = 0.0;
    bool pending_timing = false;
    lcs::DevBuf ring, in_a, in_b, out_h, out_v, descs, strip_ws, perm;
    hipEvent_t upload_done = nullptr;    // last copy out of `pinned`
    // The shared workspace (ring, descriptors, ctrl words) is used by one
    // launch at a time: a call on another stream first waits for ws_done,
    // recorded after the last call's kernels on ws_stream.
    hipEvent_t ws_done = nullptr;
    hipStream_t ws_stream = nullptr;
    void *pinned = nullptr;              // pinned staging for descriptor uploads
    size_t pinned_bytes = 0;
    lcs::Extras *extras = nullptr;
    struct lcs_mlp *cost_model = nullptr;   // not owned
    unsigned long long *band_trace = nullptr;  // device buffer set by lcs_debug_band_trace
    int32_t band_trace_cap = 0;
};

namespace lcs {
int default_tile_m(const lcs_ctx *ctx, int32_t m);
int enqueue_chain(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                  const lcs_schedule *sched, int32_t *d_h, int32_t *d_v, hipStream_t stream,
                  const int32_t *d_h_in = nullptr, const int32_t *d_v_in = nullptr);
int check_abort(lcs_ctx *ctx);
// order a call on `s` after the previous user of the shared workspace, and
// mark `s` as its latest user once the call's kernels are enqueued
int ws_acquire(lcs_ctx *ctx, hipStream_t s);
int ws_release(lcs_ctx *ctx, hipStream_t s);
int upload(lcs_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t stream);
size_t pairs_ring_words(const std::vector<PairDesc> &pairs);
// group_w: 1 = register chain (one wave per band), 4 = LDS groups of 4 bands
int prepare_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, unsigned long long *ring,
                  DevBuf &descs, unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w = 1);
// group width of a schedule variant (LCS_VARIANT_*), 0 if unknown
int group_width(int variant);
int launch_prepared(const PreparedLaunch &L, hipStream_t stream);
int prepare_pairs_own_rings(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, DevBuf &descs,
                            unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w = 1);
int enqueue_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, hipStream_t stream, int group_w = 1);
int default_batch_tile_m(int32_t m);
bool valid_tile_m(int tm);
int32_t bands_of(int32_t m, int32_t n, int tm);
void destroy_extras(lcs_ctx *ctx);
}  // namespace lcs
