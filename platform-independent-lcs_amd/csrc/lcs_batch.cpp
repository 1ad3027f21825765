// lcs_batch_comb: many pairs of uneven lengths in one persistent launch per
// tile height, the launches side by side on their own streams and captured
// in one hipGraph.  No counterpart in the reference, which combs one pair
// per call (src/main.cpp:46-64); each pair's result is the same as
// lcs_comb's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <vector>

#include "braid_kernels.h"
#include "lcs_hip.h"
#include "lcs_internal.h"

// LCS_ONEHOT=0 keeps every couple on the xor cells (A/B timing of the one-hot cells).
#ifndef LCS_ONEHOT
#define LCS_ONEHOT 1
#endif

namespace lcs {

// A batch is grouped by tile height and variant (one persistent launch per
// class): classes 0..5 register chain, 6..11 LDS groups, 12..17 the 16-bit
// packed chain (two pairs per wave).
constexpr int kTileClasses[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumTiles = 6;
constexpr int kNumClasses = 3 * kNumTiles;

struct Extras {
    hipStream_t gstream[kNumClasses] = {};
    hipEvent_t fork = nullptr, join[kNumClasses] = {};
    DevBuf gdescs[kNumClasses];
    DevBuf batch_a, batch_b, batch_h, batch_v, batch_ring;
    DevBuf check16;                  // symbol-range flags of the packed candidates
    DevBuf merge_ws, merge_descs;    // striped pairs: strip strands, permutations, steady-ant workspace
    // the instantiated graph of the last batch: the next batch's capture
    // updates it in place when its topology (the set of classes) matches
    hipGraphExec_t exec = nullptr;
    std::vector<int> exec_classes;
};

static int class_of(int tm, int group_w)
{
    for (int i = 0; i < kNumTiles; ++i)
        if (kTileClasses[i] == tm) return group_w == 1 ? i : group_w == 4 ? kNumTiles + i : -1;
    return -1;
}

// A pair the packed chain can take: 16-bit strand ids below 0xffff (the
// padding value) and both strings non-empty; its symbols are checked on the
// device.
static bool packable(int32_t m, int32_t n) { return m >= 1 && n >= 1 && (long long)m + n <= 65535; }

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
    if (X->exec) (void)hipGraphExecDestroy(X->exec);
    if (X->fork) (void)hipEventDestroy(X->fork);
    X->batch_a.release(ctx);
    X->batch_b.release(ctx);
    X->batch_h.release(ctx);
    X->batch_v.release(ctx);
    X->batch_ring.release(ctx);
    X->check16.release(ctx);
    X->merge_ws.release(ctx);
    X->merge_descs.release(ctx);
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

// The packed chain's default for a pair left to the library: one band for a
// short string, else 8 rows per lane below m = 8192 and 16 from there.  On
// the 10,000-pair batch (m, n log-uniform in [256, 16384], alphabet 4) this
// took 12.2-12.6 ms against 11.6-11.8 for the learned cost model's choice
// and 27.5 for the register chain at default_batch_tile_m.
static int default_batch_packed_tile_m(int32_t m)
{
    if (m <= 512) return default_batch_tile_m(m);
    return m < 8192 ? 8 : 16;
}

// Runs the prepared launches side by side (fork/join over the class
// streams), optionally captured in a hipGraph, between ev0 and ev1 on the
// context stream.
int run_launches(lcs_ctx *ctx, Extras *X, const std::vector<int> &classes,
                 const std::vector<PreparedLaunch> &launches, bool use_graph,
                 const std::function<int(hipStream_t)> &post)
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
    // work after every chain (the steady-ant merges of striped pairs)
    if (!rc && post) rc = post(s);
    if (use_graph) {
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (!rc && e != hipSuccess) rc = LCS_EIO;
        if (!rc) {
            // reuse the last instantiation when the captured topology matches
            bool ready = false;
            if (X->exec && X->exec_classes == classes) {
                hipGraphNode_t err_node = nullptr;
                hipGraphExecUpdateResult res;
                ready = hipGraphExecUpdate(X->exec, graph, &err_node, &res) == hipSuccess &&
                        res == hipGraphExecUpdateSuccess;
                if (!ready) (void)hipGetLastError();
            }
            if (!ready) {
                if (X->exec) (void)hipGraphExecDestroy(X->exec);
                X->exec = nullptr;
                X->exec_classes.clear();
                if (hipGraphInstantiate(&X->exec, graph, nullptr, nullptr, 0) != hipSuccess) {
                    X->exec = nullptr;
                    rc = LCS_EIO;
                } else {
                    X->exec_classes = classes;
                }
            }
            if (!rc) {
                if (hipEventRecord(ctx->ev0, s) != hipSuccess || hipGraphLaunch(X->exec, s) != hipSuccess) rc = LCS_EIO;
                if (!rc && hipEventRecord(ctx->ev1, s) != hipSuccess) rc = LCS_EIO;
                if (hipStreamSynchronize(s) != hipSuccess) rc = LCS_EIO;
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
    std::vector<int32_t> cand;                  // packed-chain candidates
    std::vector<int> cand_tm;
    std::vector<int32_t> striped;               // pairs combed as column strips
    std::vector<int> striped_class;
    auto desc = [&](int32_t p) {
        PairDesc P{};
        P.a = da + a_off[p];
        P.b = db + b_off[p];
        P.h_out = dh + a_off[p];
        P.v_out = dv + b_off[p];
        P.m = m[p];
        P.n = n[p];
        return P;
    };
    for (int32_t p = 0; p < npairs; ++p) {
        int tm = (scheds && scheds[p].tile_m) ? scheds[p].tile_m : default_batch_tile_m(m[p]);
        int variant = scheds ? scheds[p].variant : LCS_VARIANT_AUTO;
        if (variant == LCS_VARIANT_AUTO && !(scheds && (scheds[p].tile_m || scheds[p].strips > 1)) &&
            packable(m[p], n[p])) {
            // a pair left to the library that the packed chain can take
            variant = LCS_VARIANT_PACKED16;
            tm = default_batch_packed_tile_m(m[p]);
        }
        const int c = class_of(tm, group_width(variant));
        if (c < 0) return LCS_EINVAL;
        if (scheds && scheds[p].waves_per_simd) {
            if (scheds[p].waves_per_simd < 1 || scheds[p].waves_per_simd > 8) return LCS_EINVAL;
            wps[(size_t)c] = std::max(wps[(size_t)c], (int)scheds[p].waves_per_simd);
            if (variant == LCS_VARIANT_PACKED16)
                wps[(size_t)(2 * kNumTiles + c)] = std::max(wps[(size_t)(2 * kNumTiles + c)],
                                                            (int)scheds[p].waves_per_simd);
        }
        const int strips = scheds ? scheds[p].strips : 0;
        if (strips < 0) return LCS_EINVAL;
        if (strips > 1 && m[p] > 0 && n[p] > 1) {   // combed as strips, merged after the chains
            if (m[p] > kMaxCoreK) return LCS_ERANGE;
            striped.push_back(p);
            striped_class.push_back(c);
            continue;
        }
        if (variant == LCS_VARIANT_PACKED16 && packable(m[p], n[p])) {
            cand.push_back(p);
            cand_tm.push_back(c);   // the register-chain class of the same tile height
            continue;
        }
        groups[(size_t)c].push_back(desc(p));
    }
    // Striped pairs: strip s of pair p is a[p] against b[p][s W, (s+1) W),
    // combed in the chain launch of p's class with its strands in the merge
    // workspace (h) and in p's own v_out (scratch until the end); after the
    // chains, the strips' permutations are merged level by level with the
    // steady ant (k = m, all pairs' products of a level in one set of
    // launches) and the product's strands written to h_out, v_out.
    struct Striped {
        int32_t p, S, W;
        size_t strip_h, permA, permB, core_ws, c2r, ws_words, perm_off;   // word offsets in merge_ws
    };
    std::vector<Striped> sp;
    std::vector<PairDesc> strip_descs;
    std::vector<int64_t> strip_perm_off;
    size_t merge_words = 0;
    int max_strip_len = 0;
    for (size_t i = 0; i < striped.size(); ++i) {
        const int32_t p = striped[i];
        Striped t{};
        t.p = p;
        const int req = std::min<int32_t>(scheds[p].strips, n[p]);
        t.W = (n[p] + req - 1) / req;
        t.S = (n[p] + t.W - 1) / t.W;
        const size_t mm = (size_t)m[p], perm_words = (size_t)t.S * mm + (size_t)n[p];
        t.ws_words = lcsb_multiply_ws_words(m[p] + n[p], m[p] + n[p], m[p]);
        auto take = [&](size_t words) { const size_t o = merge_words; merge_words += (words + 63) & ~(size_t)63; return o; };
        t.strip_h = take((size_t)t.S * mm);
        t.permA = take(perm_words);
        t.permB = take(perm_words);
        t.core_ws = take((size_t)std::max(1, t.S / 2) * t.ws_words);
        t.c2r = take(mm + (size_t)n[p]);
        sp.push_back(t);
    }
    int rc;
    if (!sp.empty() && (rc = X->merge_ws.reserve(ctx, std::max<size_t>(merge_words, 1) * 4))) return rc;
    size_t total_strips = 0;
    for (const auto &t : sp) total_strips += (size_t)t.S;
    if (total_strips > 65535) return LCS_ERANGE;   // one launch row per strip
    int32_t *mws = (int32_t *)X->merge_ws.ptr;
    for (size_t i = 0; i < sp.size(); ++i) {
        const Striped &t = sp[i];
        const int32_t p = t.p;
        int64_t acc = (int64_t)t.permA;
        for (int st = 0; st < t.S; ++st) {
            PairDesc P{};
            P.a = da + a_off[p];
            P.b = db + b_off[p] + (size_t)st * t.W;
            P.m = m[p];
            P.n = std::min<int32_t>(t.W, n[p] - st * t.W);
            P.h_out = mws + t.strip_h + (size_t)st * m[p];
            P.v_out = dv + b_off[p] + (size_t)st * t.W;
            groups[(size_t)striped_class[i]].push_back(P);
            strip_descs.push_back(P);
            strip_perm_off.push_back(acc);
            acc += (int64_t)P.m + P.n;
            max_strip_len = std::max(max_strip_len, P.m + P.n);
        }
    }
    // couples: candidates whose symbols all lie in [0, 0xfffd], paired within
    // a tile height by (bands, columns) so the padding stays small; an odd one
    // out and every failed candidate take the register chain
    std::vector<std::vector<CoupleDesc>> couples(kNumClasses);
    if (!cand.empty()) {
        std::vector<int32_t> bad(cand.size(), 0);
        const size_t flag_bytes = (cand.size() * 4 + 255) & ~(size_t)255;   // ranges 256-B aligned after the flags
        if ((rc = X->check16.reserve(ctx, flag_bytes + sizeof(SymbolRange) * 2 * cand.size()))) return rc;
        std::vector<SymbolRange> ranges;
        for (size_t i = 0; i < cand.size(); ++i) {
            const int32_t p = cand[i];
            ranges.push_back(SymbolRange{da + a_off[p], m[p], (int32_t)i});
            ranges.push_back(SymbolRange{db + b_off[p], n[p], (int32_t)i});
        }
        int32_t *d_bad = (int32_t *)X->check16.ptr;
        SymbolRange *d_ranges = (SymbolRange *)((char *)X->check16.ptr + flag_bytes);
        if ((rc = upload(ctx, d_ranges, ranges.data(), ranges.size() * sizeof(SymbolRange), s))) return rc;
        if (lcsk_launch_check16(d_ranges, (int32_t)ranges.size(), d_bad, (int32_t)cand.size(), s)) return LCS_EIO;
        if (hipMemcpyAsync(bad.data(), d_bad, bad.size() * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return LCS_EIO;
        std::vector<char> wide_of((size_t)npairs, 0);
        for (int c = 0; c < kNumTiles; ++c) {
            // candidates of this tile height ordered by (bands, columns, pair):
            // one 64-bit key each, so the sort compares integers
            const int tm = kTileClasses[c];
            std::vector<unsigned long long> keyed;
            for (size_t i = 0; i < cand.size(); ++i) {
                if (cand_tm[i] != c) continue;
                const int32_t p = cand[i];
                if (bad[i] & 1) {
                    groups[(size_t)c].push_back(desc(p));
                    continue;
                }
                // bands < 2^10 and n <= 65535 (packable), p < 2^31; pairs with
                // a symbol above 15 sort after the rest, so one-hot pairs couple
                // with each other
                const unsigned long long bands = (unsigned long long)bands_of(m[p], n[p], tm);
                const unsigned long long wide = (bad[i] & 2) ? 1ull : 0ull;
                keyed.push_back(wide << 63 | bands << 48 | (unsigned long long)n[p] << 32 | (uint32_t)p);
                wide_of[(size_t)p] = (char)wide;
            }
            std::sort(keyed.begin(), keyed.end());
            std::vector<int32_t> mine(keyed.size());
            for (size_t i = 0; i < keyed.size(); ++i) mine[i] = (int32_t)(uint32_t)keyed[i];
            size_t i = 0;
            for (; i + 1 < mine.size(); i += 2) {
                CoupleDesc D{};
                for (int q = 0; q < 2; ++q) {
                    const int32_t p = mine[i + (size_t)q];
                    D.a[q] = da + a_off[p];
                    D.b[q] = db + b_off[p];
                    D.h_out[q] = dh + a_off[p];
                    D.v_out[q] = dv + b_off[p];
                    D.m_[q] = m[p];
                    D.n_[q] = n[p];
                }
                D.m = std::max(D.m_[0], D.m_[1]);
                D.n = std::max(D.n_[0], D.n_[1]);
                D.onehot = (LCS_ONEHOT && !wide_of[(size_t)mine[i]] && !wide_of[(size_t)mine[i + 1]]) ? 1 : 0;
                couples[(size_t)(2 * kNumTiles + c)].push_back(D);
            }
            if (i < mine.size()) groups[(size_t)c].push_back(desc(mine[i]));
        }
    }
    size_t ring_words = 0;
    for (auto &g : groups) ring_words += pairs_ring_words(g);
    for (auto &g : couples) ring_words += couples_ring_words(g);
    if ((rc = X->batch_ring.reserve(ctx, std::max<size_t>(ring_words, 1) * 8))) return rc;
    unsigned long long *ring = (unsigned long long *)X->batch_ring.ptr;
    unsigned long long *const ring_end = ring + ring_words;
    std::vector<int> classes;
    std::vector<PreparedLaunch> launches;
    for (int c = 0; c < kNumClasses; ++c) {
        PreparedLaunch L;
        if (c < 2 * kNumTiles) {
            auto &g = groups[(size_t)c];
            if (g.empty()) continue;
            rc = prepare_pairs(ctx, g, kTileClasses[c % kNumTiles], wps[(size_t)c], ring, X->gdescs[c],
                               ctx->ctrl + 2 + 2 * c, L, s, c < kNumTiles ? 1 : 4);
            if (rc) return rc;
            ring += pairs_ring_words(g);
        } else {
            auto &g = couples[(size_t)c];
            if (g.empty()) continue;
            std::vector<unsigned char> blob;
            if ((rc = build_launch16(ctx, g, kTileClasses[c % kNumTiles], wps[(size_t)c], ring,
                                     ctx->ctrl + 2 + 2 * c, blob, L)))
                return rc;
            if ((rc = X->gdescs[c].reserve(ctx, blob.size() + 256))) return rc;
            if ((rc = upload(ctx, X->gdescs[c].ptr, blob.data(), blob.size(), s))) return rc;
            rebase_launch(L, X->gdescs[c].ptr);
            ring += couples_ring_words(g);
        }
        classes.push_back(c);
        launches.push_back(L);
    }
    if (ring > ring_end) return LCS_EIO;   // every ring lies inside the reserved buffer
    // classes side by side share the SIMDs: their waiting waves sleep between polls
    if (launches.size() > 1)
        for (auto &L : launches) L.args.poll_sleep = 1;
    // the striped pairs' strip descriptors and permutation offsets, uploaded
    // before any capture
    const PairDesc *d_strips = nullptr;
    const int64_t *d_strip_off = nullptr;
    if (!sp.empty()) {
        const size_t db_bytes = sizeof(PairDesc) * strip_descs.size();
        const size_t off_at = (db_bytes + 15) & ~(size_t)15;
        std::vector<unsigned char> blob(off_at + sizeof(int64_t) * strip_perm_off.size());
        std::memcpy(blob.data(), strip_descs.data(), db_bytes);
        std::memcpy(blob.data() + off_at, strip_perm_off.data(), sizeof(int64_t) * strip_perm_off.size());
        if ((rc = X->merge_descs.reserve(ctx, blob.size() + 256))) return rc;
        if ((rc = upload(ctx, X->merge_descs.ptr, blob.data(), blob.size(), s))) return rc;
        d_strips = (const PairDesc *)X->merge_descs.ptr;
        d_strip_off = (const int64_t *)((unsigned char *)X->merge_descs.ptr + off_at);
        if (lcsb_init()) return LCS_EIO;
    }
    auto merges = [&](hipStream_t st) -> int {
        if (sp.empty()) return LCS_OK;
        if (lcsk_launch_from_strands(d_strips, (int)strip_descs.size(), max_strip_len, mws, d_strip_off, st))
            return LCS_EIO;
        // merge levels over every striped pair at once
        struct Cur {
            std::vector<int32_t *> perm;
            std::vector<int64_t> size;
            bool in_a = true;
        };
        std::vector<Cur> cur(sp.size());
        size_t k = 0;
        for (size_t i = 0; i < sp.size(); ++i) {
            for (int t = 0; t < sp[i].S; ++t, ++k) {
                cur[i].perm.push_back(mws + strip_perm_off[k]);
                cur[i].size.push_back((int64_t)strip_descs[k].m + strip_descs[k].n);
            }
        }
        for (;;) {
            std::vector<MulSpec> specs;
            std::vector<Cur> nxt(sp.size());
            bool more = false;
            for (size_t i = 0; i < sp.size(); ++i) {
                const Striped &t = sp[i];
                const int32_t mm = m[t.p];
                Cur &c = cur[i];
                nxt[i].in_a = !c.in_a;
                if (c.perm.size() == 1) {
                    nxt[i].perm = c.perm;
                    nxt[i].size = c.size;
                    nxt[i].in_a = c.in_a;
                    continue;
                }
                more = true;
                int32_t *dst = mws + (c.in_a ? t.permB : t.permA);
                int slot = 0;
                for (size_t j = 0; j + 1 < c.perm.size(); j += 2, ++slot) {
                    const int64_t out_size = c.size[j] + c.size[j + 1] - mm;
                    specs.push_back(MulSpec{c.perm[j], (int32_t)c.size[j], c.perm[j + 1], (int32_t)c.size[j + 1], mm, 0,
                                            dst, mws + t.core_ws + (size_t)slot * t.ws_words});
                    nxt[i].perm.push_back(dst);
                    nxt[i].size.push_back(out_size);
                    dst += out_size;
                }
                if (c.perm.size() % 2) {   // the odd permutation moves on unchanged
                    if (lcsk_launch_copy(dst, c.perm.back(), c.size.back(), st)) return LCS_EIO;
                    nxt[i].perm.push_back(dst);
                    nxt[i].size.push_back(c.size.back());
                }
            }
            if (!more) break;
            if (lcsb_multiply_many_device(specs.data(), (int)specs.size(), st)) return LCS_EIO;
            cur.swap(nxt);
        }
        for (size_t i = 0; i < sp.size(); ++i) {
            const int32_t p = sp[i].p;
            if (lcsk_launch_strands_from_perm(cur[i].perm[0], m[p], n[p], mws + sp[i].c2r, dh + a_off[p],
                                              dv + b_off[p], st))
                return LCS_EIO;
        }
        return LCS_OK;
    };
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    if ((rc = run_launches(ctx, X, classes, launches, use_graph != 0, merges))) return rc;
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
