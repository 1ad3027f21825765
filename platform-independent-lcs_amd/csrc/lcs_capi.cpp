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
#include <vector>

#include "lcs_hip.h"
#include "braid_kernels.h"
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

bool valid_tile_m(int tm) { return tm == 1 || tm == 2 || tm == 4 || tm == 8 || tm == 16 || tm == 32; }

// One strip whose bands cannot fill the SIMDs at one row per lane (m below
// 64 rows per SIMD): the launch is the pipeline fill, (bands - 1) lags, plus
// one band's sweep of n columns, so taller bands pay off unless the pair is
// wide.  Measured on MI355X (one strip, byte symbols, m = 1024 .. 32768):
// 4 rows per lane is fastest for n < 1.4 m (m = n = 4096: 0.301 ms against
// 0.343 at one row, 0.321 at two), two rows between 1.4 m and 1.7 m by the
// fitted per-height costs, and one row from n = 1.7 m (n = 4m: one row
// 0.511 ms, four 0.722).  Taller bands are taken only where no partial band
// would need the masked 32-bit loop (byte symbols pad their partial bands).
int small_pair_tile_m(int variant, int32_t m, int32_t n)
{
    if (m < 1024) return 1;
    const bool bytes = (variant == LCS_VARIANT_AUTO || variant == LCS_VARIANT_CHAIN ||
                        variant == LCS_VARIANT_CHAIN_BYTES) && use_sym8(variant, m, n);
    const long long n10 = 10LL * n;
    if (n10 < 14LL * m && (bytes || m % 256 == 0)) return 4;
    if (n10 < 17LL * m && (bytes || m % 128 == 0)) return 2;
    return 1;
}

int32_t bands_of(int32_t m, int32_t n, int tm)
{
    if (m <= 0 || n <= 0) return 0;
    return (int32_t)(((long long)m + 64LL * tm - 1) / (64LL * tm));
}

int Staging::init()
{
    return hipEventCreateWithFlags(&done, hipEventDisableTiming) == hipSuccess ? LCS_OK : LCS_EIO;
}

void Staging::destroy()
{
    if (done) (void)hipEventSynchronize(done);
    if (pinned) (void)hipHostFree(pinned);
    if (done) (void)hipEventDestroy(done);
    pinned = nullptr;
    bytes = 0;
    done = nullptr;
}

// Stages host bytes into a pinned buffer and copies them to `dst` on
// `stream`.  The previous upload through the same staging is waited for
// first, so the buffer is never rewritten while a copy from it is pending.
int upload_via(Staging &st, void *dst, const void *src, size_t bytes, hipStream_t stream)
{
    if (bytes == 0) return LCS_OK;
    if (hipEventSynchronize(st.done) != hipSuccess) return LCS_EIO;
    if (bytes > st.bytes) {
        if (st.pinned) (void)hipHostFree(st.pinned);
        st.pinned = nullptr;
        st.bytes = 0;
        void *p = nullptr;
        if (hipHostMalloc(&p, bytes * 2, hipHostMallocDefault) != hipSuccess) return LCS_ENOMEM;
        st.pinned = p;
        st.bytes = bytes * 2;
    }
    std::memcpy(st.pinned, src, bytes);
    if (hipMemcpyAsync(dst, st.pinned, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) return LCS_EIO;
    if (hipEventRecord(st.done, stream) != hipSuccess) return LCS_EIO;
    return LCS_OK;
}

int upload(lcs_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t stream)
{
    return upload_via(ctx->staging, dst, src, bytes, stream);
}

// Ring words of a launch over `pairs` (one ring of lcsk_ring_words(n) each).
size_t pairs_ring_words(const std::vector<PairDesc> &pairs)
{
    size_t w = 0;
    for (const auto &P : pairs) w += lcsk_ring_words(P.n);
    return w;
}

int group_width(int variant)
{
    switch (variant) {
    case LCS_VARIANT_AUTO:
    case LCS_VARIANT_CHAIN:
    case LCS_VARIANT_CHAIN_BYTES: return 1;
    case LCS_VARIANT_CHAIN_LDS: return 4;
    case LCS_VARIANT_PACKED16: return 1;   // one pair per wave outside batches
    default: return 0;
    }
}

constexpr int32_t kMaxChainN = 0x7fffffff - 4096;   // widest b a band chain combs

// One chain launch over `pairs` (one tile height): assigns each pair its
// ring inside `ring` (which must hold pairs_ring_words(pairs) words; NULL
// when every pair carries its own) and lists the tasks band-major (all pairs
// start at once, and every band comes after the band above it).
int build_launch(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, unsigned long long *ring,
                 unsigned *ctrl, int group_w, std::vector<unsigned char> &blob, PreparedLaunch &L)
{
    if (!valid_tile_m(tm) || wps < 1 || wps > 8 || (group_w != 1 && group_w != 4)) return LCS_EINVAL;
    const int np = (int)pairs.size();
    L = PreparedLaunch{};
    if (np == 0) return LCS_OK;
    int32_t max_bands = 0, max_n = 0;
    unsigned long long *E = ring;
    for (auto &P : pairs) {
        // the band loop counts steps up to n + 64 + 2 chunks in int
        if (P.n > kMaxChainN) return LCS_ERANGE;
        P.tile_m = tm;
        P.nbands = bands_of(P.m, P.n, tm);
        if (ring) {  // else every pair already carries its own ring
            P.E = E;
            E += lcsk_ring_words(P.n);
        }
        max_bands = std::max(max_bands, P.nbands);
        max_n = std::max(max_n, std::max(P.n, P.m));
    }
    // tasks: bands (or groups of group_w bands), band-major over the pairs
    std::vector<int2> tasks;
    const int32_t max_units = (max_bands + group_w - 1) / group_w;
    for (int32_t g = 0; g < max_units; ++g)
        for (int p = 0; p < np; ++p)
            if (g < (pairs[(size_t)p].nbands + group_w - 1) / group_w) tasks.push_back(make_int2(p, g));
    const size_t pair_bytes = sizeof(PairDesc) * (size_t)np, task_bytes = sizeof(int2) * tasks.size();
    const size_t base = (blob.size() + 15) & ~(size_t)15;
    blob.resize(base + pair_bytes + task_bytes);
    std::memcpy(blob.data() + base, pairs.data(), pair_bytes);
    if (task_bytes) std::memcpy(blob.data() + base + pair_bytes, tasks.data(), task_bytes);
    ChainArgs &A = L.args;
    A.pairs = (const PairDesc *)(uintptr_t)base;
    A.tasks = (const int2 *)(uintptr_t)(base + pair_bytes);
    A.ntasks = (int32_t)tasks.size();
    A.ticket = ctrl;
    A.abort_flag = ctrl + 1;
    A.sticky = ctx->ctrl + kStickyWord;
    A.timeout_ticks = ctx->timeout_ticks;
    A.trace = ctx->band_trace;
    A.trace_cap = ctx->band_trace_cap;
    A.group_w = group_w;
    const int waves_per_block = 4;
    if (group_w == 1) {
        const long long waves = std::min<long long>((long long)tasks.size(), 4LL * ctx->num_cus * wps);
        L.blocks = (int)std::max<long long>(1, (waves + waves_per_block - 1) / waves_per_block);
    } else {
        // one group per workgroup, wps groups per CU
        L.blocks = (int)std::max<long long>(1, std::min<long long>((long long)tasks.size(), (long long)ctx->num_cus * wps));
    }
    L.threads = 64 * waves_per_block;
    L.npairs = np;
    L.max_n = max_n;
    L.tile_m = tm;
    return LCS_OK;
}

void rebase_launch(PreparedLaunch &L, const void *device_blob)
{
    const uintptr_t b = (uintptr_t)device_blob;
    if (L.packed) L.args.couples = (const CoupleDesc *)(b + (uintptr_t)L.args.couples);
    else L.args.pairs = (const PairDesc *)(b + (uintptr_t)L.args.pairs);
    L.args.tasks = (const int2 *)(b + (uintptr_t)L.args.tasks);
}

// Ring words of a packed launch: one ring of lcsk_ring_words(max n) per
// couple, from the pairs' own extents (D.n is only set by build_launch16).
size_t couples_ring_words(const std::vector<CoupleDesc> &couples)
{
    size_t w = 0;
    for (const auto &D : couples) w += lcsk_ring_words(std::max(D.n_[0], D.n_[1]));
    return w;
}

// One packed-chain launch over `couples` (one tile height): rings carved
// from `ring`, tasks band-major over the couples.
int build_launch16(lcs_ctx *ctx, std::vector<CoupleDesc> &couples, int tm, int wps, unsigned long long *ring,
                   unsigned *ctrl, std::vector<unsigned char> &blob, PreparedLaunch &L)
{
    if (!valid_tile_m(tm) || wps < 1 || wps > 8) return LCS_EINVAL;
    const int nc = (int)couples.size();
    L = PreparedLaunch{};
    L.packed = true;
    if (nc == 0) return LCS_OK;
    int32_t max_bands = 0, max_n = 0;
    unsigned long long *E = ring;
    for (auto &D : couples) {
        D.tile_m = tm;
        D.m = std::max(D.m_[0], D.m_[1]);
        D.n = std::max(D.n_[0], D.n_[1]);
        D.nbands = bands_of(D.m, D.n, tm);
        D.E = E;
        E += lcsk_ring_words(D.n);
        max_bands = std::max(max_bands, D.nbands);
        max_n = std::max(max_n, std::max(D.n, D.m));
    }
    std::vector<int2> tasks;
    for (int32_t g = 0; g < max_bands; ++g)
        for (int c = 0; c < nc; ++c)
            if (g < couples[(size_t)c].nbands) tasks.push_back(make_int2(c, g));
    const size_t cbytes = sizeof(CoupleDesc) * (size_t)nc, task_bytes = sizeof(int2) * tasks.size();
    const size_t base = (blob.size() + 15) & ~(size_t)15;
    blob.resize(base + cbytes + task_bytes);
    std::memcpy(blob.data() + base, couples.data(), cbytes);
    if (task_bytes) std::memcpy(blob.data() + base + cbytes, tasks.data(), task_bytes);
    ChainArgs &A = L.args;
    A.couples = (const CoupleDesc *)(uintptr_t)base;
    A.tasks = (const int2 *)(uintptr_t)(base + cbytes);
    A.ntasks = (int32_t)tasks.size();
    A.ticket = ctrl;
    A.abort_flag = ctrl + 1;
    A.sticky = ctx->ctrl + kStickyWord;
    A.timeout_ticks = ctx->timeout_ticks;
    A.trace = ctx->band_trace;
    A.trace_cap = ctx->band_trace_cap;
    A.group_w = 1;
    const long long waves = std::min<long long>((long long)tasks.size(), 4LL * ctx->num_cus * wps);
    L.blocks = (int)std::max<long long>(1, (waves + 3) / 4);
    L.threads = 256;
    L.npairs = nc;
    L.max_n = max_n;
    L.tile_m = tm;
    return LCS_OK;
}

// Prepares one chain launch over `pairs` (one tile height): assigns each
// pair its ring inside `ring`, lists the tasks and uploads descriptors and
// tasks to `descs` through the context's staging.
int prepare_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, unsigned long long *ring,
                  DevBuf &descs, unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w)
{
    std::vector<unsigned char> blob;
    int rc = build_launch(ctx, pairs, tm, wps, ring, ctrl, group_w, blob, L);
    if (rc || pairs.empty()) return rc;
    if ((rc = descs.reserve(ctx, blob.size() + 256))) return rc;
    if ((rc = upload(ctx, descs.ptr, blob.data(), blob.size(), stream))) return rc;
    rebase_launch(L, descs.ptr);
    return LCS_OK;
}

int prepare_pairs_own_rings(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, DevBuf &descs,
                            unsigned *ctrl, PreparedLaunch &L, hipStream_t stream, int group_w)
{
    return prepare_pairs(ctx, pairs, tm, wps, nullptr, descs, ctrl, L, stream, group_w);
}

// Prep, chain and extract kernels of a prepared launch (no copies or
// memsets: capturable in a hipGraph).
int launch_prepared(const PreparedLaunch &L, hipStream_t stream)
{
    if (L.npairs == 0) return LCS_OK;
    if (L.packed)
        return lcsk_launch_chain16(L.args, L.args.couples, L.npairs, L.max_n, L.tile_m, L.blocks, L.threads, stream);
    return lcsk_launch_chain(L.args, L.args.pairs, L.npairs, L.max_n, L.tile_m, L.blocks, L.threads, stream);
}

// Launches the band chain over `pairs` on `stream` with the context's ring,
// descriptor buffer and ticket words.
int enqueue_pairs(lcs_ctx *ctx, std::vector<PairDesc> &pairs, int tm, int wps, hipStream_t stream, int group_w)
{
    int rc = ctx->ring.reserve(ctx, pairs_ring_words(pairs) * sizeof(unsigned long long));
    if (rc) return rc;
    PreparedLaunch L;
    if ((rc = prepare_pairs(ctx, pairs, tm, wps, (unsigned long long *)ctx->ring.ptr, ctx->descs, ctx->ctrl, L,
                            stream, group_w)))
        return rc;
    if (hipEventRecord(ctx->evc0, stream) != hipSuccess) return LCS_EIO;
    if ((rc = launch_prepared(L, stream))) return rc;
    if (hipEventRecord(ctx->evc1, stream) != hipSuccess) return LCS_EIO;
    ctx->chain_timed = true;
    return LCS_OK;
}

// Strips a single pair is cut into.  A thin pair (m <= kDefaultStripMaxM)
// whose bands alone would leave SIMDs idle is cut until the strips' bands
// fill them.  A tall pair whose bands already fill every SIMD is cut when it
// is wide: the persistent waves then comb each strip's lower bands as later
// rounds, so the chain's fill (one band-to-band lag per band before the
// last band starts) is paid for a fraction of the bands, at the price of
// whole-GPU steady-ant merges with k = m (S - 1 of them).  At m = n = 2^20
// (the same kernels, same box): one strip 145.0 ms, two 138.0 (chain 136.0
// + merge 2.0), four 136.7 (131.6 + 5.1), eight 140.3 (129.3 + 11.0).
constexpr int32_t kDefaultStripMaxM = 32768;
constexpr double kLargeThinCells = 34359738368.0;   // 2^35
constexpr int32_t kTallStripMinN = 1 << 18;    // two strips from here
constexpr int32_t kTallStrip4MinN = 1 << 20;   // four strips from here
int default_strips(const lcs_ctx *ctx, int32_t m, int32_t n, int tm)
{
    const long long simds = 4LL * ctx->num_cus;
    const long long bands = bands_of(m, n, tm);
    if (m > kDefaultStripMaxM) {
        if (m > kMaxCoreK || bands < simds || n < kTallStripMinN) return 1;
        return n >= kTallStrip4MinN ? 4 : 2;
    }
    if (n < 2 * 8192) return 1;
    int s = 1;
    while ((long long)s * bands < simds && s < 64 && (long long)n / (2 * s) >= 8192) s *= 2;
    return s;
}

// Byte symbols for the register chain (lcsk_launch_sym8): the pairs, all
// reading columns of the one string d_b (pair p from column col0[p]), get
// the byte copies of d_b; a chain band then takes each lane's symbols from
// them when every symbol of d_b lies in [0, 255] (which the conversion
// kernel finds out on the device) and from d_b otherwise.  A b longer than
// kSym8MaxN, or a context whose budget cannot hold the 4(n + 256) bytes,
// keeps the 32-bit symbols.  The staging is two small launches (about 8 us
// of device time at n = 16384); a pair below kSym8MinCells cells, whose
// comb that would not repay it, skips it unless its schedule asks for
// LCS_VARIANT_CHAIN_BYTES.  Besides the per-step saving, bytes let a
// partial last band (m not a multiple of 64 * tile_m) pad its rows and run
// the unmasked loop, where with 32-bit symbols it masks every cell and, as
// the chain's last band, sets the pace of the whole launch (TM=1,
// n = 103691: m = 1387 took 2.75 ms against 2.14 for the full-band
// m = 1408; with bytes 1.84).
#ifndef LCS_SYM8
#define LCS_SYM8 1
#endif
constexpr int32_t kSym8MaxN = 1 << 30;
constexpr double kSym8MinCells = 4194304.0;   // 2^22
bool use_sym8(int variant, int32_t m, int32_t n)
{
    if (!LCS_SYM8 || n <= 0 || n > kSym8MaxN) return false;
    return variant == LCS_VARIANT_CHAIN_BYTES || (double)m * (double)n >= kSym8MinCells;
}
int stage_sym8(lcs_ctx *ctx, const int32_t *d_b, int32_t n, std::vector<PairDesc> &pairs,
               const std::vector<int32_t> &col0, hipStream_t stream)
{
    const int32_t stride = lcsk_sym8_stride(n);
    const size_t bytes = (size_t)4 * 4 * (size_t)stride + 64;
    if (ctx->sym8.reserve(ctx, bytes) != LCS_OK) return LCS_OK;
    uint32_t *copies = (uint32_t *)ctx->sym8.ptr;
    int32_t *bad = (int32_t *)(copies + (size_t)4 * stride);
    if (lcsk_launch_sym8(d_b, n, copies, stride, bad, stream)) return LCS_EIO;
    for (size_t p = 0; p < pairs.size(); ++p) {
        pairs[p].sym8 = copies;
        pairs[p].sym8_bad = bad;
        pairs[p].sym8_col0 = col0[p];
        pairs[p].sym8_stride = stride;
    }
    return LCS_OK;
}

// One pair cut into `strips` column strips: comb the strips in one chain
// launch, turn each into its permutation, merge them pairwise with the
// steady ant (b-split, k = m), and read the strands off the product.
int enqueue_strips(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n, int strips, int tm,
                   int wps, int group_w, bool sym8, int32_t *d_h, int32_t *d_v, hipStream_t stream)
{
    const int32_t W = (n + strips - 1) / strips;
    const int S = (n + W - 1) / W;
    size_t perm_words = (size_t)S * m + n;
    // one workspace per product of a merge-tree level (at most S/2 of them)
    const size_t ws_words = lcsb_multiply_ws_words(m + n, m + n, m);
    const size_t nws = (size_t)(S / 2 > 0 ? S / 2 : 1);
    const size_t words = (size_t)S * m + n + 2 * perm_words + nws * ws_words + (size_t)(m + n) + 2 * (size_t)(S + 1);
    int rc = ctx->strip_ws.reserve(ctx, words * 4);
    if (rc) return rc;
    int32_t *hs = (int32_t *)ctx->strip_ws.ptr, *vs = hs + (size_t)S * m;
    int32_t *permA = vs + n, *permB = permA + perm_words, *mws = permB + perm_words;
    int32_t *c2r = mws + nws * ws_words;
    int64_t *d_off = (int64_t *)(((uintptr_t)(c2r + m + n) + 7) & ~(uintptr_t)7);
    std::vector<PairDesc> pairs((size_t)S);
    std::vector<int64_t> off((size_t)S), size((size_t)S);
    int64_t acc = 0;
    for (int s = 0; s < S; ++s) {
        PairDesc &P = pairs[(size_t)s];
        P = PairDesc{};
        P.a = d_a;
        P.b = d_b + (size_t)s * W;
        P.m = m;
        P.n = std::min<int32_t>(W, n - s * W);
        P.h_out = hs + (size_t)s * m;
        P.v_out = vs + (size_t)s * W;
        off[(size_t)s] = acc;
        size[(size_t)s] = (int64_t)m + P.n;
        acc += size[(size_t)s];
    }
    if (group_w == 1 && sym8) {
        std::vector<int32_t> col0((size_t)S);
        for (int s = 0; s < S; ++s) col0[(size_t)s] = s * W;
        if ((rc = stage_sym8(ctx, d_b, n, pairs, col0, stream))) return rc;
    }
    if ((rc = enqueue_pairs(ctx, pairs, tm, wps, stream, group_w))) return rc;
    // enqueue_pairs staged the descriptors into ctx->descs; reuse them for the permutations
    if ((rc = upload(ctx, d_off, off.data(), off.size() * sizeof(int64_t), stream))) return rc;
    if ((rc = lcsk_launch_from_strands((const PairDesc *)ctx->descs.ptr, S, m + W, permA, d_off, stream))) return rc;
    // merge tree: the products of a level share one core launch
    std::vector<int32_t *> cur((size_t)S);
    std::vector<int64_t> cur_size = size;
    for (int s = 0; s < S; ++s) cur[(size_t)s] = permA + off[(size_t)s];
    int32_t *dst_base = permB;
    while (cur.size() > 1) {
        std::vector<int32_t *> nxt;
        std::vector<int64_t> nxt_size;
        std::vector<MulSpec> specs;
        int32_t *dst = dst_base;
        for (size_t i = 0; i + 1 < cur.size(); i += 2) {
            const int64_t out_size = cur_size[i] + cur_size[i + 1] - m;
            specs.push_back(MulSpec{cur[i], (int32_t)cur_size[i], cur[i + 1], (int32_t)cur_size[i + 1], m, 0, dst,
                                    mws + specs.size() * ws_words});
            nxt.push_back(dst);
            nxt_size.push_back(out_size);
            dst += out_size;
        }
        if ((rc = lcsb_multiply_many_device(specs.data(), (int)specs.size(), stream))) return rc;
        if (cur.size() % 2) {
            // carry the odd permutation over to the next buffer
            if (hipMemcpyAsync(dst, cur.back(), (size_t)cur_size.back() * 4, hipMemcpyDeviceToDevice, stream) !=
                hipSuccess)
                return LCS_EIO;
            nxt.push_back(dst);
            nxt_size.push_back(cur_size.back());
        }
        cur.swap(nxt);
        cur_size.swap(nxt_size);
        dst_base = (dst_base == permB) ? permA : permB;
    }
    return lcsk_launch_strands_from_perm(cur[0], m, n, c2r, d_h, d_v, stream);
}

int ws_acquire(lcs_ctx *ctx, hipStream_t s)
{
    if (ctx->ws_stream && ctx->ws_stream != s && hipStreamWaitEvent(s, ctx->ws_done, 0) != hipSuccess) return LCS_EIO;
    return LCS_OK;
}

int ws_release(lcs_ctx *ctx, hipStream_t s)
{
    if (hipEventRecord(ctx->ws_done, s) != hipSuccess) return LCS_EIO;
    ctx->ws_stream = s;
    return LCS_OK;
}

static int enqueue_chain_ws(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                            const lcs_schedule *sched, int32_t *d_h, int32_t *d_v, hipStream_t stream,
                            const int32_t *d_h_in, const int32_t *d_v_in);

// Enqueue one pair on `stream` with device pointers, following `sched`,
// after the previous user of the context's workspace.
int enqueue_chain(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                  const lcs_schedule *sched, int32_t *d_h, int32_t *d_v, hipStream_t stream,
                  const int32_t *d_h_in, const int32_t *d_v_in)
{
    int rc = ws_acquire(ctx, stream);
    if (rc) return rc;
    rc = enqueue_chain_ws(ctx, d_a, m, d_b, n, sched, d_h, d_v, stream, d_h_in, d_v_in);
    const int rc2 = ws_release(ctx, stream);
    return rc ? rc : rc2;
}

static int enqueue_chain_ws(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
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
    if (n > kMaxChainN) return LCS_ERANGE;   // before any buffer is reserved
    const int strips_req = sched ? sched->strips : 0;
    int tm = (sched && sched->tile_m) ? sched->tile_m : 0;
    if (tm && !valid_tile_m(tm)) return LCS_EINVAL;
    const int wps = (sched && sched->waves_per_simd) ? sched->waves_per_simd : 1;
    if (wps < 1 || wps > 8) return LCS_EINVAL;
    if (strips_req < 0) return LCS_EINVAL;
    const int group_w = group_width(sched ? sched->variant : LCS_VARIANT_AUTO);
    if (!group_w) return LCS_EINVAL;
    if (!tm) {
        tm = default_tile_m(ctx, m);
        if (strips_req != 1 && !d_h_in && !d_v_in && tm < 4 && default_strips(ctx, m, n, 4) > 1) {
            // a thin pair cut into strips: 16 rows per lane for a large pair
            // (2^35 cells or more) when at most 64 strips of such bands fill
            // every SIMD -- on the Zipf pair (m = 2^14, n = 2^22) 64 strips of
            // 16 rows took 10.64 ms against 10.91 for 32 strips of 8 rows: the
            // chain's 0.9 ms outweighs a sixth merge level; else 8 rows when
            // at most 32 strips of them fill every SIMD (fewer merge levels
            // than with 4 rows per lane would cost), else 4
            const long long simds = 4LL * ctx->num_cus;
            const int s16 = default_strips(ctx, m, n, 16), s8 = default_strips(ctx, m, n, 8);
            if ((double)m * n >= kLargeThinCells && s16 <= 64 && (long long)s16 * bands_of(m, n, 16) >= simds)
                tm = 16;
            else
                tm = (s8 <= 32 && (long long)s8 * bands_of(m, n, 8) >= simds) ? 8 : 4;
        } else if (tm == 1) {
            tm = small_pair_tile_m(sched ? sched->variant : LCS_VARIANT_AUTO, m, n);
        }
    }
    int strips = strips_req ? strips_req : ((d_h_in || d_v_in) ? 1 : default_strips(ctx, m, n, tm));
    if (strips > 1 && (d_h_in || d_v_in)) return LCS_EINVAL;  // fixups use explicit strands: one strip
    if (strips > 1 && m > kMaxCoreK) return LCS_ERANGE;
    if (strips > n) strips = n;
    const bool sym8 = use_sym8(sched ? sched->variant : LCS_VARIANT_AUTO, m, n);
    if (strips > 1) return enqueue_strips(ctx, d_a, m, d_b, n, strips, tm, wps, group_w, sym8, d_h, d_v, stream);
    std::vector<PairDesc> pairs(1);
    PairDesc &P = pairs[0];
    P = PairDesc{};
    P.a = d_a;
    P.b = d_b;
    P.h_in = d_h_in;
    P.v_in = d_v_in;
    P.h_out = d_h;
    P.v_out = d_v;
    P.m = m;
    P.n = n;
    if (group_w == 1 && sym8) {
        const int rc = stage_sym8(ctx, d_b, n, pairs, std::vector<int32_t>(1, 0), stream);
        if (rc) return rc;
    }
    return enqueue_pairs(ctx, pairs, tm, wps, stream, group_w);
}

// The abort flag of the last launch on the context's own ctrl words (the
// caller has synchronised).  A raised flag is cleared once reported, so the
// next launch's prep kernel does not fold it into the status word again.
int check_abort(lcs_ctx *ctx)
{
    unsigned flag = 0;
    if (hipMemcpy(&flag, ctx->ctrl + 1, sizeof(flag), hipMemcpyDeviceToHost) != hipSuccess) return LCS_EIO;
    if (!flag) return LCS_OK;
    const unsigned zero = 0;
    if (hipMemcpy(ctx->ctrl + 1, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return LCS_EIO;
    return LCS_ETIMEDOUT;
}

// The schedule a single-pair call runs with (host strings give the cost
// model the alphabet; device strings are taken as alphabet 4).
static int single_schedule(lcs_ctx *ctx, const int32_t *a_host, int32_t m, const int32_t *b_host, int32_t n,
                           const lcs_schedule *sched, lcs_schedule &out)
{
    if (sched && sched->tile_m) {
        out = *sched;
        return LCS_OK;
    }
    const Selector &S = ctx->select[LCS_TUNE_SINGLE];
    int32_t alphabet = 4;
    if ((S.model || !S.miss_cands.empty()) && (a_host || b_host))
        alphabet = std::max(alphabet_of(a_host, a_host ? m : 0), alphabet_of(b_host, b_host ? n : 0));
    return choose_schedule(ctx, LCS_TUNE_SINGLE, m, n, alphabet, 0.1f, sched, &out, nullptr);
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
    if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->evc0) != hipSuccess || hipEventCreate(&ctx->evc1) != hipSuccess ||
        ctx->staging.init() != LCS_OK ||
        hipEventCreateWithFlags(&ctx->ws_done, hipEventDisableTiming) != hipSuccess) {
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
    ctx->descs.release(ctx);
    ctx->strip_ws.release(ctx);
    ctx->sym8.release(ctx);
    ctx->staging.destroy();
    if (ctx->ws_done) (void)hipEventDestroy(ctx->ws_done);
    ctx->in_a.release(ctx);
    ctx->in_b.release(ctx);
    ctx->out_h.release(ctx);
    ctx->out_v.release(ctx);
    ctx->perm.release(ctx);
    if (ctx->ctrl) (void)hipFree(ctx->ctrl);
    if (ctx->band_trace) (void)hipFree(ctx->band_trace);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->evc0) (void)hipEventDestroy(ctx->evc0);
    if (ctx->evc1) (void)hipEventDestroy(ctx->evc1);
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

int lcs_ctx_status(lcs_ctx *ctx)
{
    if (!ctx) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return LCS_EIO;
    if (ctx->ws_stream && hipEventSynchronize(ctx->ws_done) != hipSuccess) return LCS_EIO;
    unsigned words[64];
    if (hipMemcpy(words, ctx->ctrl, sizeof(words), hipMemcpyDeviceToHost) != hipSuccess) return LCS_EIO;
    const bool aborted = words[1] || words[kStickyWord];
    if (aborted) {
        const unsigned zero = 0;
        if (hipMemcpy(ctx->ctrl + 1, &zero, 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(ctx->ctrl + kStickyWord, &zero, 4, hipMemcpyHostToDevice) != hipSuccess)
            return LCS_EIO;
    }
    return aborted ? LCS_ETIMEDOUT : LCS_OK;
}

int lcs_ctx_memory(lcs_ctx *ctx, size_t *held, size_t *budget)
{
    if (!ctx) return LCS_EINVAL;
    if (held) *held = ctx->held;
    if (budget) *budget = ctx->budget;
    return LCS_OK;
}

int lcs_ctx_set_timeout(lcs_ctx *ctx, uint64_t ticks)
{
    if (!ctx) return LCS_EINVAL;
    ctx->timeout_ticks = ticks;
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

int lcs_ctx_chain_ms(lcs_ctx *ctx, double *chain_ms)
{
    if (!ctx || !chain_ms) return LCS_EINVAL;
    if (!ctx->chain_timed) return LCS_ENOENT;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    if (hipEventSynchronize(ctx->evc1) != hipSuccess) return LCS_EIO;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->evc0, ctx->evc1) != hipSuccess) return LCS_EIO;
    *chain_ms = ms;
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
    lcs_schedule chosen{};
    if ((rc = single_schedule(ctx, a, m, b, n, sched, chosen))) return rc;
    sched = &chosen;
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

int lcs_comb_permutation(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                         const lcs_schedule *sched, int32_t *row_to_col, int32_t *col_to_row, int32_t *llcs_out)
{
    if (!ctx || m < 0 || n < 0) return LCS_EINVAL;
    if ((m > 0 && !a) || (n > 0 && !b) || (m + n > 0 && (!row_to_col || !col_to_row))) return LCS_EINVAL;
    if ((long long)m + n > 0x7fffffffLL) return LCS_ERANGE;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    int rc;
    lcs_schedule chosen{};
    if ((rc = single_schedule(ctx, a, m, b, n, sched, chosen))) return rc;
    sched = &chosen;
    if ((rc = ctx->in_a.reserve(ctx, (size_t)std::max(m, 1) * 4))) return rc;
    if ((rc = ctx->in_b.reserve(ctx, (size_t)std::max(n, 1) * 4))) return rc;
    if ((rc = ctx->out_h.reserve(ctx, (size_t)std::max(m, 1) * 4))) return rc;
    if ((rc = ctx->out_v.reserve(ctx, (size_t)std::max(n, 1) * 4))) return rc;
    if ((rc = ctx->perm.reserve(ctx, (size_t)std::max(m + n, 1) * 4))) return rc;
    int32_t *da = (int32_t *)ctx->in_a.ptr, *db = (int32_t *)ctx->in_b.ptr;
    int32_t *dh = (int32_t *)ctx->out_h.ptr, *dv = (int32_t *)ctx->out_v.ptr, *dr2c = (int32_t *)ctx->perm.ptr;
    hipStream_t s = ctx->stream;
    if (m && hipMemcpyAsync(da, a, (size_t)m * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (n && hipMemcpyAsync(db, b, (size_t)n * 4, hipMemcpyHostToDevice, s) != hipSuccess) return LCS_EIO;
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    if ((rc = enqueue_chain(ctx, da, m, db, n, sched, dh, dv, s))) return rc;
    if (hipEventRecord(ctx->ev1, s) != hipSuccess) return LCS_EIO;
    if (lcsk_launch_r2c_from_strands(dh, m, dv, n, dr2c, s)) return LCS_EIO;
    // col_to_row is v then h (FromStrands: col j <- v[j], col n+l <- h[l])
    if (n && hipMemcpyAsync(col_to_row, dv, (size_t)n * 4, hipMemcpyDeviceToHost, s) != hipSuccess) return LCS_EIO;
    if (m && hipMemcpyAsync(col_to_row + n, dh, (size_t)m * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return LCS_EIO;
    if (m + n && hipMemcpyAsync(row_to_col, dr2c, (size_t)(m + n) * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return LCS_EIO;
    if (hipStreamSynchronize(s) != hipSuccess) return LCS_EIO;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) == hipSuccess) ctx->last_kernel_ms = ms;
    if ((rc = check_abort(ctx))) return rc;
    if (llcs_out) {
        int32_t c = 0;
        for (int32_t j = 0; j < n; ++j) c += col_to_row[j] < m;
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
    lcs_schedule chosen{};
    int rc = single_schedule(ctx, nullptr, m, nullptr, n, sched, chosen);
    if (rc) return rc;
    if (hipEventRecord(ctx->ev0, s) != hipSuccess) return LCS_EIO;
    rc = enqueue_chain(ctx, d_a, m, d_b, n, &chosen, d_h_out, d_v_out, s);
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
