// Internal state of a context (shared by the C ABI translation units).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <tuple>
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

// Pinned host staging for descriptor uploads.  `done` is recorded after the
// last copy out of `pinned`; the next upload through the same staging waits
// for it before rewriting the buffer.
struct Staging {
    void *pinned = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    int init();
    void destroy();
};

// A chain launch whose descriptors are already on the device.  Between
// build_launch and rebase_launch, args.pairs and args.tasks hold byte
// offsets into the host descriptor blob instead of device addresses.
struct PreparedLaunch {
    ChainArgs args{};
    int blocks = 0, threads = 0, npairs = 0, max_n = 0, tile_m = 0;
    bool packed = false;   // args.couples (the 16-bit packed chain), npairs couples
};

// Schedule selection state of one mode (LCS_TUNE_SINGLE / LCS_TUNE_BATCH):
// the autotune table, the attached cost model and the autotune-on-miss
// candidates.  Model and on-miss choices are memoised per shape.
struct TunedShape {
    int32_t m, n;
    lcs_schedule s;
};
struct Selector {
    std::vector<TunedShape> table;
    lcs_mlp *model = nullptr;                 // not owned; same context
    std::vector<lcs_schedule> model_cands;
    std::vector<lcs_schedule> miss_cands;     // empty: no autotune on a miss
    int32_t miss_reps = 1;
    std::map<std::tuple<int32_t, int32_t, int32_t>, std::pair<lcs_schedule, int>> memo;
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
    // around the band-chain launch (seed, chain, extract) of the last
    // single-pair call: its time alone, without merges or conversions
    hipEvent_t evc0 = nullptr, evc1 = nullptr;
    bool chain_timed = false;
    unsigned *ctrl = nullptr;            // [0] ticket, [1] abort flag, rest spare
    unsigned long long timeout_ticks = 500000000ull;  // 5 s at 100 MHz
    double last_kernel_ms = 0.0;
    bool pending_timing = false;
    lcs::DevBuf ring, in_a, in_b, out_h, out_v, descs, strip_ws, perm;
    lcs::DevBuf sym8;                    // byte copies of b and their range flag (single-pair chains)
    // The shared workspace (ring, descriptors, ctrl words) is used by one
    // launch at a time: a call on another stream first waits for ws_done,
    // recorded after the last call's kernels on ws_stream.
    hipEvent_t ws_done = nullptr;
    hipStream_t ws_stream = nullptr;
    lcs::Staging staging;                // descriptor uploads of the context's own calls
    lcs::Extras *extras = nullptr;
    lcs::Selector select[2];             // schedule selection per mode (single pair, batch)
    unsigned long long *band_trace = nullptr;  // device buffer set by lcs_debug_band_trace
    int32_t band_trace_cap = 0;
};

namespace lcs {
int default_tile_m(const lcs_ctx *ctx, int32_t m);
int small_pair_tile_m(int variant, int32_t m, int32_t n);
int enqueue_chain(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                  const lcs_schedule *sched, int32_t *d_h, int32_t *d_v, hipStream_t stream,
                  const int32_t *d_h_in = nullptr, const int32_t *d_v_in = nullptr);
int check_abort(lcs_ctx *ctx);
// order a call on `s` after the previous user of the shared workspace, and
// mark `s` as its latest user once the call's kernels are enqueued
int ws_acquire(lcs_ctx *ctx, hipStream_t s);
int ws_release(lcs_ctx *ctx, hipStream_t s);
int upload(lcs_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t stream);
int upload_via(Staging &st, void *dst, const void *src, size_t bytes, hipStream_t stream);
size_t pairs_ring_words(const std::vector<PairDesc> &pairs);
// Appends one chain launch over `pairs` (descriptors and task list) to
// `blob` and fills `L` with offsets into it; several launches share one
// blob, so they go to the device in one upload.  rebase_launch turns the
// offsets into addresses once the blob lies at `device_blob`.
int build_launch(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, unsigned long long *ring,
                 unsigned *ctrl, int group_w, std::vector<unsigned char> &blob, PreparedLaunch &L);
void rebase_launch(PreparedLaunch &L, const void *device_blob);
// The same for couples of pairs combed by the 16-bit packed chain.
int build_launch16(lcs_ctx *ctx, std::vector<CoupleDesc> &couples, int tm, int wps, unsigned long long *ring,
                   unsigned *ctrl, std::vector<unsigned char> &blob, PreparedLaunch &L);
size_t couples_ring_words(const std::vector<CoupleDesc> &couples);
// group_w: 1 = register chain (one wave per band), 4 = LDS groups of 4 bands
int prepare_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, unsigned long long *ring,
                  DevBuf &descs, unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w = 1);
// group width of a schedule variant (LCS_VARIANT_*), 0 if unknown
int group_width(int variant);
int launch_prepared(const PreparedLaunch &L, hipStream_t stream);
int prepare_pairs_own_rings(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, DevBuf &descs,
                            unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w = 1);
int enqueue_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, hipStream_t stream, int group_w = 1);
// whether a single pair's register chain stages its b as byte symbols
// (lcsk_launch_sym8): pairs of 2^22 cells or more, or LCS_VARIANT_CHAIN_BYTES
bool use_sym8(int variant, int32_t m, int32_t n);
int default_batch_tile_m(int32_t m);
bool valid_tile_m(int tm);
int32_t bands_of(int32_t m, int32_t n, int tm);
void destroy_extras(lcs_ctx *ctx);
// layer sizes of a cost model
int mlp_shape(const lcs_mlp *M, int *n_in, int *hidden, int *n_out);
const lcs_ctx *mlp_ctx(const lcs_mlp *M);
// Schedule of a pair whose schedule leaves tile_m at 0: the context's
// autotune table, then its cost model, then the autotune on a miss, then
// the built-in default (all fields 0).  Non-zero fields of `given` are
// kept.  `source` (nullable) receives LCS_SCHED_FROM_*.
int choose_schedule(lcs_ctx *ctx, int mode, int32_t m, int32_t n, int32_t alphabet, float rate,
                    const lcs_schedule *given, lcs_schedule *out, int32_t *source);
// The same for npairs pairs at once (one model evaluation for all misses).
int choose_schedules(lcs_ctx *ctx, int mode, int32_t npairs, const int32_t *m, const int32_t *n,
                     const int32_t *alphabet, const lcs_schedule *given, std::vector<lcs_schedule> &out);
// 1 + the largest symbol of a host string, at least 2 (the cost model's
// alphabet input); 4 when unknown
int32_t alphabet_of(const int32_t *x, int32_t len);
void detach_model(lcs_ctx *ctx, const lcs_mlp *M);
}  // namespace lcs
