// Steady-ant braid multiplication on the GPU (gfx950).
//
// Computes staggered_multiply<RowGlue>(p, q, k) of the reference
// (src/braid_multiplication.hpp:273-406): the sticky-braid (distance) product
// of two seaweed permutations glued on k strands.  The k x k core product is
// the steady ant of src/braid_multiplication.hpp:204-270, organised for a
// workgroup instead of a recursion:
//
//  * the core is padded to K = 2^L with fixed points (strands that never meet
//    another strand do not change the product), so every node of the
//    recursion at depth d is an aligned segment of K >> d positions;
//  * top-down, one stable partition per depth orders the points of P by
//    (node, row) and those of Q by (node, column) -- the reference's
//    get_vert_slice / get_horz_slice (:5-36) for all nodes of a depth at
//    once -- and remembers where each point came from;
//  * bottom-up, each depth lifts the children's products into the node's
//    coordinates (the reference's inverse_mapping, :191-202) and runs the ant
//    passage (:123-189) for all nodes at once.  The passage is cut into
//    chunks of rows: the dominance-sum difference
//        delta(r, c) = #{hi points: row < r, col < c} - #{lo points: row >= r, col >= c}
//    is non-decreasing in r and c, so the ant leaves row r at the first
//    column T(r) where delta >= 0.  Waves find T and delta at each chunk
//    boundary with sweeps over the node's columns (and, in nodes of 256
//    rows or more, where the ant crosses every rc-th column with sweeps
//    over its rows); then one thread per piece walks it with the
//    reference's incremental updates and places the good points.
// The union of the children's points, the per-column view and the parents
// of every depth's partition live in LDS (16-bit entries, the lo/hi flag in
// bit 15), so the walks and the lifts run at LDS latency.  Cores up to
// K = 2048 take this path (the parents alone are 4 K log2(K) bytes).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "braid_kernels.h"

namespace lcsb {

constexpr int kThreads = 1024;
// nodes of at least this many rows also cut their walks at column lines
#ifndef LCS_COL_LINE_MIN_S
#define LCS_COL_LINE_MIN_S 256
#endif
constexpr int kColLineMinS = LCS_COL_LINE_MIN_S;
// rows per walked chunk of a node (nodes above 128 chunks take s / 128)
#ifndef LCS_CORE_CHUNK_ROWS
#define LCS_CORE_CHUNK_ROWS 8
#endif
constexpr int kChunkRows = LCS_CORE_CHUNK_ROWS;
static_assert(kChunkRows >= 4 && kChunkRows <= 64 && (kChunkRows & (kChunkRows - 1)) == 0, "chunk rows");
constexpr uint16_t kLoFlag = 0x8000u;
constexpr uint16_t kIdx = 0x7fffu;
// cores up to this size keep keys, products and parents in LDS
// (braid_core_lds_kernel and the leaves of larger cores)
constexpr int kLdsCoreL = 11;
constexpr int kLdsCoreK = 1 << kLdsCoreL;
// row chunks per node are at most 128 (chunk height max(kChunkRows, s/128)); each wave
// keeps 2 x kMaxChunks bucket counters in LDS for the boundary sweeps
constexpr int kMaxChunks = 128;
// dynamic LDS of a core of 2^L points: keys / U, CI, R, Rn (4K uint16), the
// K/64+1 group counts of P and of Q, then the parents of P and of Q at each
// depth (2 L K uint16)
constexpr size_t lds_core_bytes(int L)
{
    return (size_t)4 * (1 << L) * sizeof(uint16_t) + 2 * ((size_t)(1 << L) / 64 + 1) * 4 +
           (size_t)2 * L * (1 << L) * sizeof(uint16_t);
}

// Inclusive prefix sum over the 64 lanes of a wave with DPP: four row_shr
// steps scan each row of 16 lanes, then row_bcast:15 and row_bcast:31 carry
// the row totals forward (lanes without a source add the old value, 0).
__device__ __forceinline__ int wave_incl_scan(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2, 3
    return x;
}

// One bottom-up depth (node size s) of the core: lift the children's
// products into node coordinates (the reference's inverse_mapping, :191-202)
// and run the chunked ant passage (:123-189) for every node of the depth.
// U, CI are in LDS; R (children's products) and Rn (this depth's) are
// either global int32 or LDS uint16 arrays; pp, pq (LDS) map depth d+1
// positions to depth d positions; bT, bD hold K/kChunkRows chunk entries.
template <typename UT, typename RT, typename RNT>
__device__ __forceinline__ void bottom_up_depth(UT *U, UT *CI, const RT *R, RNT *Rn, const uint16_t *pp,
                                                const uint16_t *pq, int K, int s, int32_t *bT, int32_t *bD,
                                                int32_t *cT, int *hist)
{
    const int h = s >> 1;
    // lift the children's products into node coordinates
#pragma unroll 4
    for (int x = threadIdx.x; x < K; x += kThreads) {
        const int p_pos = pp[x];
        const int cseg = x & ~(h - 1);
        const int q_pos = pq[cseg + (int)R[x]];
        const int v0 = p_pos & ~(s - 1);
        const uint16_t lo = (x & h) ? 0 : kLoFlag;
        const int lr = p_pos - v0, lc = q_pos - v0;
        U[p_pos] = (uint16_t)lc | lo;
        CI[v0 + lc] = (uint16_t)lr | lo;
        Rn[p_pos] = lc;
    }
    __syncthreads();
    // chunk height: kChunkRows, or s/128 for big nodes, balancing the
    // boundary scans (O(s) each) against the walks (O(rows) each)
    // (nodes under 64 rows are one chunk: the boundary sweeps read 64 columns per round)
    const int rc = s < 64 ? s : (s / 128 > kChunkRows ? s / 128 : kChunkRows);
    const int rc_log = __builtin_ctz(rc);   // rc and s are powers of two
    const int nch = s / rc;            // chunks per node
    // chunk entry points: chunk j of a node covers rows [s-(j+1)rc, s-j rc).
    // A node's first chunk (j = 0) enters at T = 0; its entry is computed in
    // the walk, so only chunks j > 0 (nodes of 64 rows or more, at most
    // K/kChunkRows chunks) store entries in bT, bD.  For chunk j > 0 entering at row
    // Dr = s - j rc:
    //   T(Dr) = min{ c : Z(c) >= Y }, Z(c) = #{c' < c : z(c')},
    //   z(c) = hi(c) ? row(c) < Dr : row(c) >= Dr,  Y = #{lo rows >= Dr},
    // and delta(Dr, T) = Z(T) - Y = 0.  With the row bucket of a column,
    // b(c) = (s-1-row(c)) / rc, z_j(c) = lo ? b(c) < j : b(c) >= j, so one
    // wave finds all of a node's boundaries in two sweeps over its columns
    // (64 per round, conflict-free): the histogram of the round's buckets
    // (lo and hi, wave-local LDS counters) and a scan over the buckets give
    // every boundary's count at once; lane L keeps boundaries 2L and 2L+1.
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1;
    if (nch > 1) {
        int *hlo = hist + wave * 2 * kMaxChunks, *hhi = hlo + kMaxChunks;   // this wave's counters
        const int j0 = 2 * lane, j1 = 2 * lane + 1;
        // Bucket counts of the columns [c_begin, c_end) of a node, as seen
        // from boundaries j0 and j1: lo columns with bucket < j (lp) and
        // columns with z_j (cnt).  The wave's LDS accesses run in order; the
        // fences keep the compiler from moving counter stores, adds and loads
        // across each other.
        auto range_counts = [&](const UT *E, int v0, int c_begin, int c_end, int &lp0, int &lp1, int &cnt0,
                                int &cnt1) {
            hlo[j0] = 0;
            hlo[j1] = 0;
            hhi[j0] = 0;
            hhi[j1] = 0;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            for (int c0 = c_begin; c0 < c_end; c0 += 64) {
                const uint16_t e = E[v0 + c0 + lane];
                const int bk = (s - 1 - (e & kIdx)) >> rc_log;
                atomicAdd((e & kLoFlag) ? &hlo[bk] : &hhi[bk], 1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            const int l0 = __hip_atomic_load(&hlo[j0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const int l1 = __hip_atomic_load(&hlo[j1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const int g0 = __hip_atomic_load(&hhi[j0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const int g1 = __hip_atomic_load(&hhi[j1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const int il = wave_incl_scan(l0 + l1), ih = wave_incl_scan(g0 + g1);
            const int hi_total = __builtin_amdgcn_readlane(ih, 63);
            lp0 = il - l0 - l1;
            lp1 = lp0 + l0;
            const int hp0 = ih - g0 - g1, hp1 = hp0 + g0;   // hi columns with bucket < j
            cnt0 = lp0 + (hi_total - hp0);
            cnt1 = lp1 + (hi_total - hp1);
        };
        // Running Z_j over [c_begin, c_end) in rounds of 64 columns for the
        // open boundaries until Z_j reaches Y_j; T = the column where it does.
        auto find_T = [&](const UT *E, int v0, int c_begin, int c_end, int Y0, int Y1, int Z0, int Z1, bool open0,
                          bool open1, int &T0, int &T1) {
            for (int c0 = c_begin; c0 < c_end && __ballot(open0 || open1); c0 += 64) {
                hlo[j0] = 0;
                hlo[j1] = 0;
                hhi[j0] = 0;
                hhi[j1] = 0;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const uint16_t e = E[v0 + c0 + lane];
                const bool lo = (e & kLoFlag) != 0;
                const int bk = (s - 1 - (e & kIdx)) >> rc_log;
                atomicAdd(lo ? &hlo[bk] : &hhi[bk], 1);
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const int l0 = __hip_atomic_load(&hlo[j0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const int l1 = __hip_atomic_load(&hlo[j1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const int g0 = __hip_atomic_load(&hhi[j0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const int g1 = __hip_atomic_load(&hhi[j1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                const int il = wave_incl_scan(l0 + l1), ih = wave_incl_scan(g0 + g1);
                const int hi_total = __builtin_amdgcn_readlane(ih, 63);
                const int lp0 = il - l0 - l1, lp1 = lp0 + l0;      // lo columns with bucket < j
                const int hp0 = ih - g0 - g1, hp1 = hp0 + g0;      // hi columns with bucket < j
                const int cnt0 = lp0 + (hi_total - hp0), cnt1 = lp1 + (hi_total - hp1);
                const bool cross0 = open0 && Z0 + cnt0 >= Y0, cross1 = open1 && Z1 + cnt1 >= Y1;
                if (open0 && !cross0) Z0 += cnt0;
                if (open1 && !cross1) Z1 += cnt1;
                // boundaries reaching Y in this round: find the exact column, one at a time
                const unsigned long long todo = __ballot(cross0), todo1 = __ballot(cross1);
                for (int which = 0; which < 2; ++which) {
                    unsigned long long t = which ? todo1 : todo;
                    while (t) {
                        const int l = __ffsll((long long)t) - 1;
                        t &= t - 1;
                        const int jj = 2 * l + which;
                        const int need = __shfl(which ? Y1 - Z1 : Y0 - Z0, l);   // 1-based rank in this round
                        const bool z = lo ? bk < jj : bk >= jj;
                        const unsigned long long zm = __ballot(z);
                        const bool hit = z && __popcll(zm & lt) == need - 1;
                        const int pos = __ffsll((long long)__ballot(hit));  // 1-based lane of the column
                        if (lane == l) {
                            if (which) T1 = c0 + pos; else T0 = c0 + pos;
                        }
                    }
                }
                if (cross0) open0 = false;
                if (cross1) open1 = false;
            }
        };
        // entries: enter row Dr-1, delta(Dr-1, T) = -[row Dr-1 counted at T]
        auto entry = [&](int node, int v0, int j, int T) {
            const int Dr = s - j * rc;
            const uint16_t e = U[v0 + Dr - 1];
            const int col = e & kIdx;
            const int ch = node * nch + j;
            bT[ch] = T;
            bD[ch] = -((e & kLoFlag) ? (col >= T) : (col < T));
        };
        constexpr int kWaves = kThreads / 64;
        const int nodes = K / s;
        // dir 0: row lines from the columns' entries (CI); dir 1 (nodes of
        // kColLineMinS rows or more): column lines from the rows' entries
        // (U), the column line c = i rc being boundary nch - i, T = rho + 1
        for (int dir = 0; dir < (s >= kColLineMinS ? 2 : 1); ++dir) {
            const UT *E = dir ? U : CI;
            auto store = [&](int node, int v0, int j, int T) {
                if (dir) cT[node * nch + j] = T;
                else entry(node, v0, j, T);
            };
            if (dir) __syncthreads();   // the row direction's counters are read
            if (nodes >= kWaves || K < 64 * kWaves) {
                // one wave per node: Y_j from the node's counts, then one sweep
                for (int node = wave; node < nodes; node += kWaves) {
                    const int v0 = node * s;
                    int Y0, Y1, c0_, c1_;
                    range_counts(E, v0, 0, s, Y0, Y1, c0_, c1_);
                    int T0 = 0, T1 = 0;
                    find_T(E, v0, 0, s, Y0, Y1, 0, 0, j0 >= 1 && j0 < nch && Y0 > 0, j1 < nch && Y1 > 0, T0, T1);
                    if (j0 >= 1 && j0 < nch) store(node, v0, j0, T0);
                    if (j1 < nch) store(node, v0, j1, T1);
                }
            } else {
                // W = kWaves / nodes waves per node, each on s / W consecutive
                // entries: every wave counts its segment, the segments' counts
                // give Y_j and each segment's starting Z_j, and only the wave in
                // whose segment Z_j reaches Y_j sweeps for T_j.
                const int W = kWaves / nodes, node = wave / W, seg = wave % W, seg_len = s / W;
                const int v0 = node * s, cb = seg * seg_len;
                int lp0, lp1, cnt0, cnt1;
                range_counts(E, v0, cb, cb + seg_len, lp0, lp1, cnt0, cnt1);
                hlo[j0] = cnt0;
                hlo[j1] = cnt1;
                hhi[j0] = lp0;
                hhi[j1] = lp1;
                __syncthreads();
                int Y0 = 0, Y1 = 0, Zs0 = 0, Zs1 = 0;
                for (int w2 = 0; w2 < W; ++w2) {
                    const int *other = hist + (node * W + w2) * 2 * kMaxChunks;
                    Y0 += other[kMaxChunks + j0];
                    Y1 += other[kMaxChunks + j1];
                    if (w2 < seg) {
                        Zs0 += other[j0];
                        Zs1 += other[j1];
                    }
                }
                __syncthreads();   // every wave has read the counts before they are reused
                const bool mine0 = j0 >= 1 && j0 < nch && Y0 > 0 && Zs0 < Y0 && Y0 <= Zs0 + cnt0;
                const bool mine1 = j1 < nch && Y1 > 0 && Zs1 < Y1 && Y1 <= Zs1 + cnt1;
                int T0 = 0, T1 = 0;
                find_T(E, v0, cb, cb + seg_len, Y0, Y1, Zs0, Zs1, mine0, mine1, T0, T1);
                if (j0 >= 1 && j0 < nch && (mine0 || (Y0 == 0 && seg == 0))) store(node, v0, j0, T0);
                if (j1 < nch && (mine1 || (Y1 == 0 && seg == 0))) store(node, v0, j1, T1);
            }
        }
    }
    __syncthreads();
    // The walk (the reference's ant passage, :123-189) with the four entries
    // it reads kept in registers: cl = CI[dc-1], cc = CI[dc], ud = U[dr],
    // ur = U[dr-1].  A step right only needs the new CI[dc] and a step down
    // only the new U[dr-1], so every step is one LDS read from a selected
    // address, and the step is written with selects so that no lane of the
    // wave branches on the direction.
    // counted(e, r) = lo ? idx >= r : idx < r for an entry e = lo << 15 | idx:
    // bit 15 of e - r, for idx in [0, 32768) and r in [0, 32768].
    auto counted = [](uint32_t e, int r) { return (int)(((e - (uint32_t)r) >> 15) & 1u); };
    // Pieces: chunk j (q = j < nch) starts at its row line; in nodes of
    // kColLineMinS rows or more the walks are also cut at the column lines
    // c = i rc (q = nch - 1 + i), so every piece stays in one rc x rc box
    // and takes at most 2 rc + 1 steps, whatever the permutations look like
    // (a strip permutation's long straight runs would otherwise fall to a
    // few chunks).  Smaller nodes walk whole row chunks.
    const bool boxes = s >= kColLineMinS && nch > 1;
    const int per_node = boxes ? 2 * nch - 1 : nch, pieces = (K / s) * per_node;
    for (int pc = threadIdx.x; pc < pieces; pc += kThreads) {
        const int node = pc / per_node, q = pc % per_node, v0 = node * s;
        int dr, dc, delta, r_lo, c_hi;
        if (q < nch) {
            const int j = q, r_hi = s - j * rc;
            r_lo = r_hi - rc;
            dr = r_hi - 1;
            if (j == 0) {
                // T(s) = 0, delta(s, 0) = 0; entering row s-1 subtracts it when it is lo
                dc = 0;
                delta = (U[v0 + s - 1] & kLoFlag) ? -1 : 0;
            } else {
                dc = min(max(bT[node * nch + j], 0), s);
                delta = bD[node * nch + j];
            }
            const int cb = dc == 0 ? 0 : (dc - 1) >> rc_log;
            c_hi = !boxes || cb >= nch - 1 ? s + 1 : (cb + 1) * rc + 1;
        } else {
            // stepping right over column line c = i rc at row rho: delta = -1 + [CI(c) counted at rho]
            const int i = q - nch + 1;
            const int Tc = min(max(cT[node * nch + nch - i], 0), s);
            if (Tc == 0) continue;   // the ant never reaches this column line
            dr = Tc - 1;
            dc = i * rc + 1;
            delta = -1 + counted(CI[v0 + i * rc], dr);
            r_lo = s - (((s - 1 - dr) >> rc_log) + 1) * rc;
            c_hi = i == nch - 1 ? s + 1 : (i + 1) * rc + 1;
        }
        // entries outside the node (dc = 0 or s, dr = 0) are read clamped and never used
        uint32_t cl = CI[v0 + max(dc - 1, 0)], cc = CI[v0 + min(dc, s - 1)];
        uint32_t ud = U[v0 + dr], ur = U[v0 + max(dr - 1, 0)];
        while (dr >= r_lo && dc < c_hi) {
            const int d_left = delta - counted(cl, dr);
            const int d_down = delta + counted(ud, dc);
            if (dc > 0 && d_left < 0 && d_down > 0) Rn[v0 + dr] = dc - 1;
            // delta(dr, s) >= 0 always; the dc < s test only bounds the walk
            const bool right = delta < 0 && dc < s;
            const int d_right = counted(cc, dr), d_up = dr > 0 ? counted(ur, dc) : 0;
            delta = right ? delta + d_right : delta - d_up;
            dc += right ? 1 : 0;
            dr -= right ? 0 : 1;
            cl = right ? cc : cl;
            ud = right ? ud : ur;
            const uint32_t nx = right ? CI[v0 + min(dc, s - 1)] : U[v0 + max(dr - 1, 0)];
            cc = right ? nx : cc;
            ur = right ? ur : nx;
        }
    }
    __syncthreads();
}

// One depth of the top-down split with the keys in LDS, for P (waves
// 0 .. kSideW-1) and Q (the other half) at once: a stable partition of every
// aligned segment of size s by bit `bit` (zeros first), the ranks from wave
// ballots; par[pos] = old position.  Each wave counts the ones of its 64-key
// groups (gc, one word per group); after one barrier every wave scans all
// of its side's group counts itself (K <= 4096: at most 64 groups, one DPP
// scan), so a depth costs two barriers.  A wave takes the groups w, w +
// kSideW, ..., U at a time, with every load of the U groups issued before
// the first ballot, so their LDS latencies overlap (the compiler does not
// unroll a loop around a ballot by itself).  K >= 64.
constexpr int kSideW = kThreads / 128;
static_assert(kLdsCoreK <= 64 * 64, "one DPP scan covers a side's group counts");
template <int U>
__device__ __forceinline__ void split_count_groups(const uint16_t *key, int bit, int32_t *gc, int g0, int lane)
{
    uint16_t kv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) kv[u] = key[(g0 + u * kSideW) * 64 + lane];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const unsigned long long m = __ballot((kv[u] >> bit) & 1);
        if (lane == 0) gc[g0 + u * kSideW] = __popcll(m);
    }
}

// excl = exclusive prefix of the group counts at lane g, total = all ones
template <int U>
__device__ __forceinline__ void split_place_groups(const uint16_t *key, uint16_t *key_new, uint16_t *par, int s,
                                                   int bit, int excl, int total, int ngroups, int g0, int lane)
{
    const unsigned long long lt = (1ull << lane) - 1;
    uint16_t kv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) kv[u] = key[(g0 + u * kSideW) * 64 + lane];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int g = g0 + u * kSideW, x = g * 64 + lane;
        const int b = (kv[u] >> bit) & 1;
        const unsigned long long m = __ballot(b);
        const int seg = x & ~(s - 1);
        int before, ones;
        if (s >= 64) {
            // the segment's groups are gs .. ge-1; prefixes read from the scan
            const int gs = seg >> 6, ge = gs + (s >> 6);
            const int p_g = __builtin_amdgcn_readlane(excl, g), p_s = __builtin_amdgcn_readlane(excl, gs);
            const int p_e = ge < ngroups ? __builtin_amdgcn_readlane(excl, ge) : total;
            before = p_g - p_s + __popcll(m & lt);
            ones = p_e - p_s;
        } else {
            const unsigned long long segmask = ((1ull << s) - 1) << (lane & ~(s - 1));
            before = __popcll(m & lt & segmask);
            ones = __popcll(m & segmask);
        }
        const int pos = b ? seg + (s - ones) + before : seg + (x - seg - before);
        key_new[pos] = kv[u];
        par[pos] = (uint16_t)x;
    }
}

__device__ __forceinline__ void split_depth_pq(const uint16_t *keyP, uint16_t *keyP_new, uint16_t *parP,
                                               const uint16_t *keyQ, uint16_t *keyQ_new, uint16_t *parQ,
                                               int32_t *gcP, int32_t *gcQ, int K, int s, int bit)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool q_side = wave >= kSideW;
    const int w = q_side ? wave - kSideW : wave;
    const uint16_t *key = q_side ? keyQ : keyP;
    uint16_t *key_new = q_side ? keyQ_new : keyP_new, *par = q_side ? parQ : parP;
    int32_t *gc = q_side ? gcQ : gcP;
    const int ngroups = K >> 6;
    int g = w;
    for (; g + 3 * kSideW < ngroups; g += 4 * kSideW) split_count_groups<4>(key, bit, gc, g, lane);
    for (; g < ngroups; g += kSideW) split_count_groups<1>(key, bit, gc, g, lane);
    __syncthreads();
    const int c = lane < ngroups ? gc[lane] : 0;
    const int incl = wave_incl_scan(c);
    const int excl = incl - c, total = __builtin_amdgcn_readlane(incl, 63);
    g = w;
    for (; g + 3 * kSideW < ngroups; g += 4 * kSideW)
        split_place_groups<4>(key, key_new, par, s, bit, excl, total, ngroups, g, lane);
    for (; g < ngroups; g += kSideW) split_place_groups<1>(key, key_new, par, s, bit, excl, total, ngroups, g, lane);
    __syncthreads();
}

// The sticky product with keys, both products, the parents of every depth
// and the chunk entries in LDS for cores up to K = kLdsCoreK, one workgroup
// per core; the dynamic LDS is laid out as lds_core_bytes(L) says.  P holds
// row -> shared column; Q holds shared row -> column, or (QKEY) column ->
// shared row, the keys themselves.  Values are read modulo K (a leaf of a
// big core reads its slice of the big core's keys).
template <bool QKEY>
__device__ __forceinline__ void core_lds_body(const int32_t *P, const int32_t *Q, int32_t *Rout, int k, int L,
                                              uint16_t *lds, int32_t *bT, int32_t *bD, int32_t *cT, int *hist)
{
    const int K = 1 << L, mask = K - 1;
    uint16_t *keyP = lds, *keyP2 = lds + K;                       // top-down
    uint16_t *keyQ = lds + 2 * (size_t)K, *keyQ2 = lds + 3 * (size_t)K;
    uint16_t *U = lds, *CI = lds + K;                             // bottom-up (reuses the key space)
    uint16_t *R = lds + 2 * (size_t)K, *Rn = lds + 3 * (size_t)K;
    int32_t *gcP = (int32_t *)(lds + 4 * (size_t)K), *gcQ = gcP + (K / 64 + 1);
    uint16_t *parP = lds + 4 * (size_t)K + 4 * ((size_t)K / 64 + 1);   // L * K
    uint16_t *parQ = parP + (size_t)L * K;                              // L * K

    // P ordered by row (key = shared column), Q ordered by column (key = shared row)
    for (int x = threadIdx.x; x < K; x += kThreads) keyP[x] = (uint16_t)(x < k ? (P[x] & mask) : x);
    if (QKEY) {
        for (int x = threadIdx.x; x < K; x += kThreads) keyQ[x] = (uint16_t)(Q[x] & mask);
    } else {
        for (int x = threadIdx.x; x < K; x += kThreads)
            if (x >= k) keyQ[x] = (uint16_t)x;
        for (int r = threadIdx.x; r < k; r += kThreads)
            if ((uint32_t)Q[r] < (uint32_t)k) keyQ[Q[r]] = (uint16_t)r;
    }
    __syncthreads();
    for (int d = 0; d < L; ++d) {
        split_depth_pq(keyP, keyP2, parP + (size_t)d * K, keyQ, keyQ2, parQ + (size_t)d * K, gcP, gcQ, K, K >> d,
                       L - d - 1);
        uint16_t *t = keyP; keyP = keyP2; keyP2 = t;
        t = keyQ; keyQ = keyQ2; keyQ2 = t;
    }
    for (int x = threadIdx.x; x < K; x += kThreads) R[x] = 0;  // depth L: 1x1 nodes
    __syncthreads();
    for (int d = L - 1; d >= 0; --d) {
        bottom_up_depth(U, CI, R, Rn, parP + (size_t)d * K, parQ + (size_t)d * K, K, K >> d, bT, bD, cT, hist);
        uint16_t *tmp = R; R = Rn; Rn = tmp;
    }
    for (int x = threadIdx.x; x < k; x += kThreads) Rout[x] = R[x];
}

__global__ __launch_bounds__(kThreads) void braid_core_lds_kernel(CoreBatch batch)
{
    const BraidCoreDesc D = batch.d[blockIdx.x];
    extern __shared__ uint16_t lds[];
    __shared__ int32_t bT[kLdsCoreK / kChunkRows], bD[kLdsCoreK / kChunkRows], cT[kLdsCoreK / kChunkRows];
    __shared__ int hist[(kThreads / 64) * 2 * kMaxChunks];
    core_lds_body<false>(D.P, D.Q, D.R, D.k, D.L, lds, bT, bD, cT, hist);
}

// ---------------------------------------------------------------- big cores
// Cores above kLdsCoreK run the same recursion spread over the whole GPU
// (the one-workgroup core walks its top depths in the same box pieces):
//  * the top d0 = L - L0 depths of the top-down split are device-wide
//    stable partitions (count, scan, scatter per depth);
//  * the 2^d0 nodes of depth d0 are independent sticky products of 2^L0
//    points each, one workgroup per node (core_lds_body on their keys);
//  * bottom-up, each top depth lifts the children's products into node
//    coordinates (inverse_mapping, src/braid_multiplication.hpp:191-202)
//    and runs the ant passage (:123-189) cut into pieces that a lane each
//    walks in LDS.  The ant leaves row r at T(r) = min{c : delta(r, c) >= 0}
//    (delta non-decreasing in r and c), so its path is fixed by where it
//    crosses a grid of lines every g rows and every g columns of a node:
//      - row line Dr (= s - j g): enters row Dr-1 at column T(Dr), with
//        delta(Dr, T) = 0;
//      - column line c = i g: the ant steps right from (rho, c) to
//        (rho, c+1) at rho + 1 = min{r : delta(r, c) >= 0}, with
//        delta(rho, c) = -1.
//    Both are "first position where a running count reaches a target" over
//    the node's columns (row lines) or rows (column lines): with bucket
//    b(e) = (s-1-idx(e)) >> log2 g of an entry and boundary j,
//        z_j(e) = lo(e) ? b(e) < j : b(e) >= j,  Y_j = #{lo entries: b < j},
//        T_j = min{x : #{x' < x : z_j(e_x')} >= Y_j},
//    the column line c = i g being boundary j = s/g - i over the rows
//    (T_j = rho + 1).  Block histograms (big_hist_kernel), a search over
//    the blocks per boundary and a sweep of the one block where the count
//    crosses (big_sweep_kernel) find every T_j.  A piece starts at a
//    crossing and ends at the next one, so it stays inside one g x g box and
//    takes at most 2g + 1 steps (big_walk_kernel).
// Entries of U (row -> column) and CI (column -> row) are node-local
// indices with the lo flag in bit 31.
constexpr int kBigBatch = 32;          // products per big-core launch
constexpr int kSplitTile = 2048;       // elements per block of a device-wide partition
constexpr int kMaxLines = 8192;        // grid lines per node and direction
// Box side g of a node's grid: kBoxGSmall for nodes up to 2^kBoxSmallMaxL
// points, kBoxG above (nodes above kMaxLines * g take g = s / kMaxLines).
// Smaller boxes halve the walks and double the lines whose crossings the
// sweeps find; the sweeps are cheap for small nodes.  Same box: the Zipf
// pair's merges (nodes up to 2^14) 1.60 ms with boxes of 32 against 1.78
// with 64; a k = 2^20 product 1.60 ms with 32 everywhere against 1.47.
#ifndef LCS_BIG_BOX_G
#define LCS_BIG_BOX_G 64
#endif
#ifndef LCS_BIG_BOX_G_SMALL
#define LCS_BIG_BOX_G_SMALL 32
#endif
#ifndef LCS_BIG_BOX_SMALL_MAX_L
#define LCS_BIG_BOX_SMALL_MAX_L 14
#endif
constexpr int kBoxG = LCS_BIG_BOX_G, kBoxGSmall = LCS_BIG_BOX_G_SMALL, kBoxSmallMaxL = LCS_BIG_BOX_SMALL_MAX_L;
static_assert(kBoxG >= 16 && kBoxG <= 512 && (kBoxG & (kBoxG - 1)) == 0, "box side");
static_assert(kBoxGSmall >= 16 && kBoxGSmall <= 512 && (kBoxGSmall & (kBoxGSmall - 1)) == 0, "box side");
constexpr uint32_t kLo31 = 0x80000000u;

struct BigCore {
    const int32_t *P, *Q;   // k: row -> shared column, shared row -> column
    int32_t *R;             // k: row -> column of the product
    int32_t *ws;            // lcsb_big_ws_words(k, L0) words
    int32_t k;
};
// grid of one top depth (node size s): lines every g = s / nb
struct BigDepth {
    int32_t s, nb, glog, w, nsegs;   // w entries per histogram block, nsegs per node
};

// Histogram blocks of w entries: every open boundary sweeps (on average)
// half a block, nb w / 2 entries per node and direction, while the
// blocks' prefixes cost about 2 (s / w)(nb + 1) words written and read;
// w = 2 sqrt(s) balances the two (clamped to [2^LCS_BIG_MIN_W_LOG, 8192],
// at most s = 2^e).
#ifndef LCS_BIG_MIN_W_LOG
#define LCS_BIG_MIN_W_LOG 9
#endif
static_assert(LCS_BIG_MIN_W_LOG >= 7 && LCS_BIG_MIN_W_LOG <= 11, "histogram block width");
__host__ __device__ constexpr int big_w_log(int e)
{
    const int wl = 1 + (e + 1) / 2;
    return wl < LCS_BIG_MIN_W_LOG ? LCS_BIG_MIN_W_LOG : (wl > 13 ? 13 : wl);
}
constexpr bool big_blocks_fit(int max_blocks)
{
    for (int e = 11; e <= kMaxCoreL; ++e)
        if ((1 << (e - (big_w_log(e) < e ? big_w_log(e) : e))) > max_blocks) return false;
    return true;
}

__host__ __device__ inline BigDepth big_depth(int L, int d)
{
    BigDepth D;
    D.s = 1 << (L - d);
    const int g = L - d <= kBoxSmallMaxL ? kBoxGSmall : kBoxG;
    D.nb = D.s / g < kMaxLines ? D.s / g : kMaxLines;
    D.glog = __builtin_ctz((unsigned)(D.s / D.nb));
    D.w = 1 << big_w_log(L - d);
    if (D.w > D.s) D.w = D.s;
    D.nsegs = D.s / D.w;
    return D;
}

// word offsets inside one product's workspace
struct BigLayout {
    size_t keyP, keyQ, keyP2, keyQ2, parP, parQ, R, Rn, U, CI, cnt, hist, T, total;
};

__host__ __device__ inline BigLayout big_layout(int L, int L0)
{
    const size_t K = (size_t)1 << L;
    const int d0 = L - L0;
    BigLayout B;
    size_t o = 0;
    B.keyP = o; o += K;
    B.keyQ = o; o += K;
    B.keyP2 = o; o += K;
    B.keyQ2 = o; o += K;
    B.parP = o; o += (size_t)d0 * K;
    B.parQ = o; o += (size_t)d0 * K;
    B.R = o; o += K;
    B.Rn = o; o += K;
    B.U = o; o += K;
    B.CI = o; o += K;
    B.cnt = o; o += 3 * 2 * (K / kSplitTile);   // tile counts of P and Q, three depths in rotation
    size_t hist = 0, lines = 0;
    for (int d = 0; d < d0; ++d) {
        const BigDepth D = big_depth(L, d);
        const size_t nodes = K / (size_t)D.s;
        const size_t h = 2 * nodes * (size_t)D.nsegs * (size_t)(D.nb + 1);
        hist = h > hist ? h : hist;
        lines = nodes * (size_t)D.nb > lines ? nodes * (size_t)D.nb : lines;
    }
    B.hist = o; o += hist;
    B.T = o; o += 2 * lines;
    B.total = o;
    return B;
}

// The launch arguments of the big-core kernels: the products, and the
// workspace layout and the grid of every top depth, computed once on the
// host (per wave they were a loop of 64-bit divisions on the scalar unit).
struct BigBatch {
    BigCore c[kBigBatch];
    int32_t L, L0;
    BigLayout Y;
    BigDepth D[kMaxCoreL];
};

__device__ __forceinline__ uint32_t counted31(uint32_t e, int r) { return ((e - (uint32_t)r) >> 31) & 1u; }

__global__ void big_init_kernel(BigBatch B)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int K = 1 << B.L, x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= K) return;
    int32_t *keyP = C.ws + Y.keyP, *keyQ = C.ws + Y.keyQ;
    keyP[x] = x < C.k ? C.P[x] : x;
    if (x >= C.k) keyQ[x] = x;
    else if ((uint32_t)C.Q[x] < (uint32_t)C.k) keyQ[C.Q[x]] = x;
    // the tile counts of depths 1 and 2 are accumulated by the scatters
    // before them and start at zero
    const int ntiles = K / kSplitTile;
    if (x < 2 * 2 * ntiles) C.ws[Y.cnt + 2 * (size_t)ntiles + x] = 0;
}

// Depth-0 ones of the top bit per tile of kSplitTile keys (y = 0: P keys,
// 1: Q keys), into count buffer 0; every later depth's counts are
// accumulated by the scatter of the depth before.
__global__ __launch_bounds__(256) void big_split_count_kernel(BigBatch B, int d)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int K = 1 << B.L, bit = B.L - d - 1, y = blockIdx.y;
    const int32_t *key = C.ws + (y ? ((d & 1) ? Y.keyQ2 : Y.keyQ) : ((d & 1) ? Y.keyP2 : Y.keyP));
    const int base = blockIdx.x * kSplitTile + threadIdx.x * 8;
    int c = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) c += (key[base + q] >> bit) & 1;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    __shared__ int red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0)
        C.ws[Y.cnt + ((size_t)(d % 3) * 2 + y) * (K / kSplitTile) + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Stable partition of every segment of size s by bit `bit` (zeros first);
// par[pos] = old position.  The block's own tile count prefix and its
// segment's total ones come from the tile counts of this depth (count
// buffer d % 3), summed by the block itself.  When another device-wide
// depth follows (NEXT), the block also counts the next bit's ones of the
// keys it places per destination tile -- its zeros and its ones each land
// in a run of at most kSplitTile positions, so at most two tiles each --
// and adds them to buffer (d + 1) % 3, and clears its own tile in buffer
// (d + 2) % 3 for the depth after that.
template <bool NEXT>
__global__ __launch_bounds__(256) void big_split_scatter_kernel(BigBatch B, int d)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int K = 1 << B.L, s = K >> d, bit = B.L - d - 1, y = blockIdx.y;
    const int ntiles = K / kSplitTile;
    const int32_t *key = C.ws + (y ? ((d & 1) ? Y.keyQ2 : Y.keyQ) : ((d & 1) ? Y.keyP2 : Y.keyP));
    int32_t *key_new = C.ws + (y ? ((d & 1) ? Y.keyQ : Y.keyQ2) : ((d & 1) ? Y.keyP : Y.keyP2));
    int32_t *par = C.ws + (y ? Y.parQ : Y.parP) + (size_t)d * K;
    const int32_t *cnt = C.ws + Y.cnt + ((size_t)(d % 3) * 2 + y) * ntiles;
    const int base = blockIdx.x * kSplitTile + threadIdx.x * 8;
    int32_t kv[8];
    int c = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        kv[q] = key[base + q];
        c += (kv[q] >> bit) & 1;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the segment's tiles t0 .. t0+nt-1; ones in the tiles before this one and in all of them
    const int seg = (blockIdx.x * kSplitTile) & ~(s - 1);
    const int t0 = seg / kSplitTile, nt = s / kSplitTile, mine = blockIdx.x - t0;
    int pre = 0, all = 0;
    for (int i = threadIdx.x; i < nt; i += 256) {
        const int v = cnt[t0 + i];
        all += v;
        pre += i < mine ? v : 0;
    }
    // exclusive prefix of the thread counts over the block, and the block sums
    const int incl = wave_incl_scan(c);
    for (int off = 32; off > 0; off >>= 1) {
        pre += __shfl_xor(pre, off);
        all += __shfl_xor(all, off);
    }
    __shared__ int wsum[4], wpre[4], wall[4];
    __shared__ int nz[2], no[2];
    if (lane == 63) wsum[wave] = incl;
    if (lane == 0) {
        wpre[wave] = pre;
        wall[wave] = all;
    }
    if (NEXT && threadIdx.x < 2) {
        nz[threadIdx.x] = 0;
        no[threadIdx.x] = 0;
    }
    __syncthreads();
    int before = incl - c;
    for (int w2 = 0; w2 < wave; ++w2) before += wsum[w2];
    const int before_tiles = wpre[0] + wpre[1] + wpre[2] + wpre[3];
    const int ones_total = wall[0] + wall[1] + wall[2] + wall[3];
    before += before_tiles;
    const int zeros_total = s - ones_total;
    // first destinations of the block's zeros and ones (their runs' first tiles)
    const int z0 = blockIdx.x * kSplitTile - before_tiles, o0 = seg + zeros_total + before_tiles;
    const int zt = z0 / kSplitTile, ot = o0 / kSplitTile;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int x = base + q;
        const int b = (kv[q] >> bit) & 1;
        const int pos = b ? seg + zeros_total + before : seg + (x - seg - before);
        before += b;
        key_new[pos] = kv[q];
        par[pos] = x;
        if (NEXT && ((kv[q] >> (bit - 1)) & 1))
            atomicAdd(b ? &no[pos / kSplitTile - ot] : &nz[pos / kSplitTile - zt], 1);
    }
    if (NEXT) {
        int32_t *next = C.ws + Y.cnt + ((size_t)((d + 1) % 3) * 2 + y) * ntiles;
        __syncthreads();
        if (threadIdx.x < 4) {
            const int v = threadIdx.x < 2 ? nz[threadIdx.x] : no[threadIdx.x - 2];
            const int tile = (threadIdx.x < 2 ? zt : ot) + (threadIdx.x & 1);
            if (v) atomicAdd(&next[tile], v);
        }
        if (threadIdx.x == 0) C.ws[Y.cnt + ((size_t)((d + 2) % 3) * 2 + y) * ntiles + blockIdx.x] = 0;
    }
}

// the 2^d0 independent nodes of depth d0, one workgroup each
__global__ __launch_bounds__(kThreads) void big_leaf_kernel(BigBatch B)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int d0 = B.L - B.L0, s0 = 1 << B.L0;
    const size_t v0 = (size_t)blockIdx.x * s0;
    const int32_t *keyP = C.ws + ((d0 & 1) ? Y.keyP2 : Y.keyP) + v0;
    const int32_t *keyQ = C.ws + ((d0 & 1) ? Y.keyQ2 : Y.keyQ) + v0;
    extern __shared__ uint16_t lds[];
    __shared__ int32_t bT[kLdsCoreK / kChunkRows], bD[kLdsCoreK / kChunkRows], cT[kLdsCoreK / kChunkRows];
    __shared__ int hist[(kThreads / 64) * 2 * kMaxChunks];
    // depth d0 of the big core's own recursion: the products go to R
    core_lds_body<true>(keyP, keyQ, C.ws + Y.R + v0, s0, B.L0, lds, bT, bD, cT, hist);
}

// lift the children's products into node coordinates (node size s).  The
// children's products are in R (flip = 0) or Rn (flip = 1), the node's go
// to the other one.
__global__ void big_lift_kernel(BigBatch B, int d, int flip)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int K = 1 << B.L, s = K >> d, h = s >> 1;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= K) return;
    const int32_t *pp = C.ws + Y.parP + (size_t)d * K, *pq = C.ws + Y.parQ + (size_t)d * K;
    const int32_t *R = C.ws + (flip ? Y.Rn : Y.R);
    int32_t *Rn = C.ws + (flip ? Y.R : Y.Rn);
    uint32_t *U = (uint32_t *)(C.ws + Y.U), *CI = (uint32_t *)(C.ws + Y.CI);
    const int p_pos = pp[x];
    const int cseg = x & ~(h - 1);
    const int q_pos = pq[cseg + R[x]];
    const int v0 = p_pos & ~(s - 1);
    const uint32_t lo = (x & h) ? 0u : kLo31;
    const int lr = p_pos - v0, lc = q_pos - v0;
    U[p_pos] = (uint32_t)lc | lo;
    CI[v0 + lc] = (uint32_t)lr | lo;
    Rn[p_pos] = lc;
}

// Per histogram block (w entries of CI for y = 0, of U for y = 1): packed
// exclusive prefix over the buckets of (lo count | hi count << 16) for each
// boundary j < nb, and the packed totals at [nb].
__global__ __launch_bounds__(256) void big_hist_kernel(BigBatch B, int d)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const BigDepth &D = B.D[d];
    const int K = 1 << B.L, y = blockIdx.y;
    const uint32_t *E = (const uint32_t *)(C.ws + (y ? Y.U : Y.CI));
    __shared__ uint32_t h[kMaxLines];
    __shared__ uint32_t sums[4];
    for (int j = threadIdx.x; j < D.nb; j += 256) h[j] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * D.w;
    for (int i = threadIdx.x; i < D.w; i += 256) {
        const uint32_t e = E[base + i];
        const int bk = (D.s - 1 - (int)(e & ~kLo31)) >> D.glog;
        atomicAdd(&h[bk], (e & kLo31) ? 1u : 0x10000u);
    }
    __syncthreads();
    // exclusive prefix in place: each of nw waves scans its nb / nw
    // consecutive buckets 64 at a time (DPP scan plus a carry; the packed
    // halves stay below 2^16, so they add without carrying into each
    // other), then the prefix leaves with consecutive threads on
    // consecutive buckets plus the totals of the waves before
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = D.nb >= 256 ? 4 : 1, chunk = D.nb / nw;   // nb is a power of two
    if (wave < nw) {
        uint32_t carry = 0;
        const int jb = wave * chunk, je = jb + chunk;
        for (int j0 = jb; j0 < je; j0 += 64) {
            const int j = j0 + lane;
            const uint32_t c = j < je ? h[j] : 0u;
            const uint32_t incl = (uint32_t)wave_incl_scan((int)c);
            if (j < je) h[j] = carry + incl - c;
            carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
        }
        if (lane == 0) sums[wave] = carry;
    }
    __syncthreads();
    // per (y, node): the nsegs blocks' prefix rows of nb words (each block
    // writes its own row, coalesced), then the nsegs block totals in a row
    // of their own (read coalesced by the search)
    const int node = blockIdx.x / D.nsegs, g = blockIdx.x % D.nsegs, nodes = K / D.s;
    uint32_t *rows = (uint32_t *)(C.ws + Y.hist) + ((size_t)y * nodes + node) * D.nsegs * (size_t)(D.nb + 1);
    uint32_t *out = rows + (size_t)g * D.nb;
    const uint32_t w1 = sums[0], w2 = w1 + (nw > 1 ? sums[1] : 0u), w3 = w2 + (nw > 2 ? sums[2] : 0u);
    for (int j = threadIdx.x; j < D.nb; j += 256) {
        const int wj = j / chunk;
        out[j] = h[j] + (wj == 0 ? 0u : wj == 1 ? w1 : wj == 2 ? w2 : w3);
    }
    if (threadIdx.x == 0) rows[(size_t)D.nsegs * D.nb + g] = nw > 1 ? w3 + sums[3] : w1;
}

// The block of a node where the running count of z_j entries reaches Y_j.
// Hj[g * nb] = block g's prefix at j (lo | hi << 16 entries with bucket <
// j), Ht[g] = its totals: Y_j sums the lo parts; the block holds lo(P) +
// hi(total) - hi(P) z_j entries.  Lane l holds blocks l, l + 64, ... (R
// rounds, 64 R >= nsegs), all loaded in one batch from clamped (valid)
// blocks and masked afterwards; the rounds' wave scans find the first block
// where Z (z_j entries before it) plus its own count reaches Y_j.
template <int R>
__device__ __forceinline__ void block_search(const uint32_t *Hj, const uint32_t *Ht, int nb, int nsegs, int lane,
                                             int &Yj, int &Z, int &seg)
{
    uint32_t Pv[R], Tv[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int g = min(64 * u + lane, nsegs - 1);
        Pv[u] = Hj[(size_t)g * nb];
        Tv[u] = Ht[g];
    }
    int lo_sum = 0;
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const bool mine = 64 * u + lane < nsegs;
        Pv[u] = mine ? Pv[u] : 0u;
        Tv[u] = mine ? Tv[u] : 0u;
        lo_sum += (int)(Pv[u] & 0xffffu);
    }
    Yj = lo_sum;
    for (int off = 32; off > 0; off >>= 1) Yj += __shfl_xor(Yj, off);
    Z = 0;
    seg = nsegs - 1;
    bool found = false;   // uniform
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int c = (int)(Pv[u] & 0xffffu) + (int)(Tv[u] >> 16) - (int)(Pv[u] >> 16);
        const int incl = wave_incl_scan(c);
        const unsigned long long reach = found ? 0ull : __ballot(Z + incl >= Yj);
        if (reach) {
            const int l = __ffsll((long long)reach) - 1;
            seg = min(64 * u + l, nsegs - 1);
            Z += __shfl(incl - c, l);
            found = true;
        } else if (!found) {
            Z += __builtin_amdgcn_readlane(incl, 63);
        }
    }
}

// The entry of a block of w entries (w >= 64 RS, a multiple of it) where
// the count of z_j entries reaches `need`: the block is read RS x 64
// entries at a time, all loads issued before the first ballot, so one
// memory latency covers that much of it.  Returns the entry's position
// + 1 in the block (0 if the count never gets there).
template <int RS>
__device__ __forceinline__ int sweep_block(const uint32_t *E, int w, int s, int glog, int j, int need, int lane)
{
    const unsigned long long lt = (1ull << lane) - 1;
    for (int r0 = 0; r0 < w; r0 += 64 * RS) {
        uint32_t ev[RS];
#pragma unroll
        for (int u = 0; u < RS; ++u) ev[u] = E[r0 + 64 * u + lane];
#pragma unroll
        for (int u = 0; u < RS; ++u) {
            const uint32_t e = ev[u];
            const int bk = (s - 1 - (int)(e & ~kLo31)) >> glog;
            const bool z = (e & kLo31) ? bk < j : bk >= j;
            const unsigned long long ball = __ballot(z);
            const int cnt = __popcll(ball);
            if (need <= cnt) {
                const bool hit = z && __popcll(ball & lt) == need - 1;
                return r0 + 64 * u + __ffsll((long long)__ballot(hit));
            }
            need -= cnt;
        }
    }
    return 0;
}

// One wave per (y, node, boundary j): Y_j and the block where the running
// count reaches it (block_search, at most kSegRounds x 64 blocks per node),
// then a sweep of that block (sweep_block) for the entry where the count
// reaches the rank still needed; T = that entry's position + 1.
// Boundaries with j = 0 or Y_j = 0 get T = 0.
constexpr int kSegRounds = 16;
static_assert(big_blocks_fit(64 * kSegRounds), "histogram blocks per node");
#ifndef LCS_SWEEP_WAVES
#define LCS_SWEEP_WAVES 4
#endif
constexpr int kSweepWaves = LCS_SWEEP_WAVES;   // waves (boundaries) per workgroup
__global__ __launch_bounds__(64 * kSweepWaves) void big_sweep_kernel(BigBatch B, int d)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const BigDepth &D = B.D[d];
    const int K = 1 << B.L, nodes = K / D.s, y = blockIdx.y;
    // the wave's boundary, made scalar so that its addresses stay in SGPRs
    const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * kSweepWaves + (threadIdx.x >> 6)), lane = threadIdx.x & 63;
    if (t >= nodes * D.nb) return;
    const int node = t / D.nb, j = t % D.nb;
    int32_t *T = C.ws + Y.T + (size_t)y * nodes * D.nb + t;
    if (j == 0) {
        if (lane == 0) *T = 0;
        return;
    }
    const uint32_t *H = (const uint32_t *)(C.ws + Y.hist) + ((size_t)y * nodes + node) * D.nsegs * (size_t)(D.nb + 1);
    const uint32_t *Hj = H + j, *Ht = H + (size_t)D.nsegs * D.nb;
    int Yj, Z, seg;
    if (D.nsegs <= 64) block_search<1>(Hj, Ht, D.nb, D.nsegs, lane, Yj, Z, seg);
    else if (D.nsegs <= 128) block_search<2>(Hj, Ht, D.nb, D.nsegs, lane, Yj, Z, seg);
    else if (D.nsegs <= 256) block_search<4>(Hj, Ht, D.nb, D.nsegs, lane, Yj, Z, seg);
    else if (D.nsegs <= 512) block_search<8>(Hj, Ht, D.nb, D.nsegs, lane, Yj, Z, seg);
    else block_search<kSegRounds>(Hj, Ht, D.nb, D.nsegs, lane, Yj, Z, seg);
    if (Yj == 0) {
        if (lane == 0) *T = 0;
        return;
    }
    const uint32_t *E = (const uint32_t *)(C.ws + (y ? Y.U : Y.CI)) + (size_t)node * D.s + (size_t)seg * D.w;
    // w is a power of two >= 128: the whole block in one batch up to 1024 entries
    const int need = Yj - Z;
    const int pos = D.w <= 128   ? sweep_block<2>(E, D.w, D.s, D.glog, j, need, lane)
                    : D.w <= 256 ? sweep_block<4>(E, D.w, D.s, D.glog, j, need, lane)
                    : D.w <= 512 ? sweep_block<8>(E, D.w, D.s, D.glog, j, need, lane)
                                 : sweep_block<16>(E, D.w, D.s, D.glog, j, need, lane);
    if (lane == 0 && pos) *T = seg * D.w + pos;
}

// A box entry relative to the box: idx clamped to [base, top] minus base,
// the lo flag in bit 15.  For r in (base, top], counted31(e, r) of the
// absolute entry equals counted16 of this one at r - base.
__device__ __forceinline__ uint16_t box_entry(uint32_t e, int base, int top)
{
    const int idx = (int)(e & ~kLo31);
    return (uint16_t)(((e & kLo31) ? 0x8000u : 0u) | (uint32_t)(min(max(idx, base), top) - base));
}
__device__ __forceinline__ int counted16(uint32_t e, int r) { return (int)(((e - (uint32_t)r) >> 15) & 1u); }

// One lane per piece of the ant's path.  Piece q < nb of a node starts at
// row line j = q (q = 0: the start of the passage at (s-1, 0)), q >= nb at
// column line i = q - nb + 1; it stays in one g x g box.  Each lane stages
// its box -- rows r_lo-1 .. r_hi-1 of U and columns c_lo-1 .. c_hi-1 of CI,
// clamped into the node -- in LDS as 16-bit entries relative to the box, then
// walks it with the reference's incremental updates, placing the good
// points.  In box coordinates R = dr - r_lo + 1 (1 .. g while in the box)
// and X = dc - c_lo + 1 (1 .. nc - 1 before the exit line): U(dr) = bu[R],
// U(dr-1) = bu[R-1], CI(dc) = bc[X], CI(dc-1) = bc[X-1].
constexpr int kMaxBoxG = 512;   // g at the largest node (2^22 / kMaxLines)
constexpr size_t walk_lds_bytes(int g) { return (size_t)64 * (2 * g + 3) * sizeof(uint16_t); }
__global__ __launch_bounds__(64) void big_walk_kernel(BigBatch B, int d, int flip)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const BigDepth &D = B.D[d];
    const int K = 1 << B.L, s = D.s, nodes = K / s, g = 1 << D.glog, nb = D.nb;
    const int pieces = 2 * nb - 1;
    const int t = blockIdx.x * 64 + threadIdx.x;
    if (t >= nodes * pieces) return;
    const int node = t / pieces, q = t % pieces;
    const size_t v0 = (size_t)node * s;
    const uint32_t *U = (const uint32_t *)(C.ws + Y.U) + v0, *CI = (const uint32_t *)(C.ws + Y.CI) + v0;
    int32_t *Rn = C.ws + (flip ? Y.R : Y.Rn) + v0;
    const int32_t *Trow = C.ws + Y.T + (size_t)node * nb;
    const int32_t *Tcol = C.ws + Y.T + (size_t)nodes * nb + (size_t)node * nb;
    // T lies in [0, s]; the clamps only keep every access inside the node
    // after an aborted comb fed undefined input
    int dr, dc, r_lo, cb;
    if (q < nb) {
        const int Dr = s - q * g;
        dc = q == 0 ? 0 : min(max(Trow[q], 0), s);
        dr = Dr - 1;
        r_lo = Dr - g;
        cb = dc == 0 ? 0 : (dc - 1) >> D.glog;
    } else {
        const int i = q - nb + 1;
        const int Tc = min(max(Tcol[nb - i], 0), s);
        if (Tc == 0) return;   // the ant never reaches this column line
        dr = Tc - 1;
        dc = i * g + 1;
        r_lo = s - (((s - 1 - dr) >> D.glog) + 1) * g;
        cb = i;
    }
    const int c_lo = cb ? cb * g + 1 : 0, c_hi = cb == nb - 1 ? s + 1 : (cb + 1) * g + 1;
    const int nc = c_hi - c_lo + 1;   // at most g + 2
    extern __shared__ uint16_t walk_lds[];
    uint16_t *bu = walk_lds + (size_t)threadIdx.x * (2 * g + 3), *bc = bu + (g + 1);
    // U entries hold columns, compared with dc in [c_lo, c_hi - 1]; CI
    // entries hold rows, compared with dr in [r_lo, r_lo + g - 1]
#pragma unroll 32
    for (int i = 0; i <= g; ++i) bu[i] = box_entry(U[min(max(r_lo - 1 + i, 0), s - 1)], c_lo - 1, c_hi);
#pragma unroll 32
    for (int i = 0; i < nc; ++i) bc[i] = box_entry(CI[min(max(c_lo - 1 + i, 0), s - 1)], r_lo - 1, r_lo + g);
    // entering row Dr-1 at T: delta = -[U(Dr-1) counted at T]; stepping right
    // over column line c = i g at row rho: delta = -1 + [CI(c) counted at rho]
    int R = dr - r_lo + 1, X = dc - c_lo + 1;
    const int Xs = s - c_lo + 1;   // X < Xs <=> dc < s
    int delta = q < nb ? -counted16(bu[R], X) : -1 + counted16(bc[0], R);
    uint32_t cl = bc[X - 1], cc = bc[X], ud = bu[R], ur = bu[R - 1];
    // the lanes of a wave take different directions, so each step is written
    // with selects and reads one LDS entry from a selected index: the new
    // CI(dc) after a step right, U(dr-1) after a step down
    while (R >= 1 && X < nc) {
        const int d_left = delta - counted16(cl, R);
        const int d_down = delta + counted16(ud, X);
        if (X + c_lo > 1 && d_left < 0 && d_down > 0) Rn[R + r_lo - 1] = X + c_lo - 2;
        const bool right = delta < 0 && X < Xs;
        const int d_right = counted16(cc, R), d_up = R + r_lo > 1 ? counted16(ur, X) : 0;
        delta = right ? delta + d_right : delta - d_up;
        X += right ? 1 : 0;
        R -= right ? 0 : 1;
        cl = right ? cc : cl;
        ud = right ? ud : ur;
        const uint32_t nx = bu[right ? g + 1 + min(X, nc - 1) : max(R - 1, 0)];
        cc = right ? nx : cc;
        ur = right ? ur : nx;
    }
}

// the product of the last depth is in R (flip = 0) or Rn (flip = 1)
__global__ void big_out_kernel(BigBatch B, int flip)
{
    const BigCore &C = B.c[blockIdx.z];
    const BigLayout &Y = B.Y;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < C.k) C.R[x] = C.ws[(flip ? Y.Rn : Y.R) + x];
}

// ---------------------------------------------------------------- slicing
// Device-wide stable compaction in three kernels: per-block counts, a scan
// of the counts in one block, per-block scatter.  The products of one merge
// level go through each kernel in one launch: blockIdx.y picks the slice
// (two per product, p's rows at even y and q's columns at odd y).
constexpr int kScanBlock = 256, kScanPer = 8, kScanTile = kScanBlock * kScanPer;
constexpr int kMulBatch = 32;   // products per slicing / assembly launch

struct SliceSpec {
    // selects element i of `src` (size n) when lo <= src[i] < hi; per-block
    // counts (then their exclusive scan) in `counts`; records, for slice j of
    // the selected elements (j < cap):
    //   even slice (vertical slice of p, rows in order):   map[j] = i, val[j] = src[i] - lo
    //   odd slice (horizontal slice of q, cols in order):  map[j] = i (a column), val[src[i] - lo] = j
    const int32_t *src;
    int32_t *counts, *map, *val;
    int32_t n, lo, hi, cap;
};
struct SliceBatch {
    SliceSpec s[2 * kMulBatch];
};

__device__ __forceinline__ bool sel(const SliceSpec &S, int i) { const int v = S.src[i]; return v >= S.lo && v < S.hi; }

__global__ void slice_count_kernel(SliceBatch B)
{
    const SliceSpec &S = B.s[blockIdx.y];
    if ((int64_t)blockIdx.x * kScanTile >= S.n) return;   // past this slice's blocks (whole block)
    __shared__ int32_t red[kScanBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int c = 0;
    for (int q = 0; q < kScanPer; ++q) c += (base + q < S.n) && sel(S, base + q);
    red[threadIdx.x] = c;
    __syncthreads();
    for (int off = kScanBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) S.counts[blockIdx.x] = red[0];
}

__global__ void scan_counts_kernel(SliceBatch B)
{
    // exclusive scan of slice blockIdx.x's block counts in place, one block of 1024
    const SliceSpec &S = B.s[blockIdx.x];
    int32_t *counts = S.counts;
    const int nblocks = (S.n + kScanTile - 1) / kScanTile;
    __shared__ int32_t s[1024];
    const int t = threadIdx.x;
    const int per = (nblocks + 1023) / 1024;
    const int lo = min(t * per, nblocks), hi = min(lo + per, nblocks);
    int run = 0;
    for (int i = lo; i < hi; ++i) run += counts[i];
    s[t] = run;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int acc = s[t] - run;
    for (int i = lo; i < hi; ++i) { const int c = counts[i]; counts[i] = acc; acc += c; }
}

__global__ void slice_scatter_kernel(SliceBatch B)
{
    const SliceSpec &S = B.s[blockIdx.y];
    if ((int64_t)blockIdx.x * kScanTile >= S.n) return;   // past this slice's blocks (whole block)
    const bool rows = (blockIdx.y & 1) == 0;
    __shared__ int32_t s[kScanBlock];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanPer;
    int c = 0;
    for (int q = 0; q < kScanPer; ++q) c += (base + q < S.n) && sel(S, base + q);
    s[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        int v = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int j = S.counts[blockIdx.x] + s[threadIdx.x] - c;
    for (int q = 0; q < kScanPer; ++q) {
        const int i = base + q;
        if (i < S.n && sel(S, i) && j < S.cap) {
            S.map[j] = i;
            if (rows) S.val[j] = S.src[i] - S.lo;
            else S.val[S.src[i] - S.lo] = j;
            ++j;
        }
    }
}

struct InvertBatch {
    struct {
        const int32_t *r2c;
        int32_t *c2r;
        int32_t n;
    } s[kMulBatch];
};

__global__ void invert_kernel(InvertBatch B)
{
    const auto &S = B.s[blockIdx.y];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // no wrap near 2^31
    if (i < S.n && (uint32_t)S.r2c[i] < (uint32_t)S.n) S.c2r[S.r2c[i]] = (int32_t)i;
}

struct AssembleBatch {
    AssembleArgs a[kMulBatch];
};

// Writes the product (row -> col) from the core result and the untouched
// points of p and q (braid_multiplication.hpp:353-401).
__global__ void assemble_kernel(AssembleBatch B)
{
    const AssembleArgs &A = B.a[blockIdx.y];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // no wrap near 2^31
    const int k = A.k, ps = A.p_size, qs = A.q_size;
    // the range tests only matter after an aborted comb fed undefined input
    if (i < k && (uint32_t)A.core[i] < (uint32_t)k) {
        const int row = A.map_row[i], col = A.map_col[A.core[i]];
        if ((uint32_t)row < (uint32_t)ps) {
            if (A.row_glue) A.out[qs - k + row] = col;
            else A.out[row] = col + ps - k;
        }
    }
    if (i < ps) {
        const int c = A.p[i];
        if (A.row_glue) { if (c >= k) A.out[i + qs - k] = c + qs - k; }
        else { if (c < ps - k) A.out[i] = c; }
    }
    if (i < qs) {
        const int c = A.q[i];
        if (A.row_glue) { if (i < qs - k) A.out[i] = c; }
        else { if (i >= k) A.out[i - k + ps] = c - k + ps; }
    }
}

}  // namespace lcsb

using namespace lcsb;

// Cores of 2^L points with L >= kBigMinL run over the whole GPU, with
// leaves of at most 2^kBigLeafMaxL points (at least 2^10: the device-wide
// partitions work on tiles of 2048 keys inside nodes of 2^(leaf+1); at most
// 2^kLdsCoreL: a leaf is one workgroup's LDS core).
// Measured on the Zipf pair's five merge levels (k = 2^14, 16 .. 1 products
// per level): one workgroup per core 15.6 ms for the whole pair; whole-GPU
// cores with leaves of 2^10 / 2^11 / 2^12 points 12.5 / 12.3 / 12.5 ms.
#ifndef LCS_BIG_MIN_L
#define LCS_BIG_MIN_L 12
#endif
#ifndef LCS_BIG_LEAF_L
#define LCS_BIG_LEAF_L 11
#endif
// cores of 2^16 points and more take leaves of up to 2^LCS_BIG_LEAF_L_LARGE
#ifndef LCS_BIG_LEAF_L_LARGE
#define LCS_BIG_LEAF_L_LARGE 11
#endif
constexpr int kBigMinL = LCS_BIG_MIN_L;
constexpr int kBigLeafMaxL = LCS_BIG_LEAF_L;
constexpr int kBigLeafLargeL = LCS_BIG_LEAF_L_LARGE;
static_assert(kBigMinL >= 11 && kBigMinL <= kLdsCoreL + 1 && kBigLeafMaxL >= 10 && kBigLeafMaxL <= kLdsCoreL &&
                  kBigLeafLargeL >= 10 && kBigLeafLargeL <= kLdsCoreL,
              "big-core leaf sizes");
static int big_leaf_L(int L)
{
    const int lim = L >= 16 ? kBigLeafLargeL : kBigLeafMaxL;
    return L - 1 < lim ? L - 1 : lim;
}
static bool is_big(int L) { return L >= kBigMinL; }

size_t lcsb_core_scratch_words(int32_t k)
{
    const int L = lcsb_core_L(k);
    if (!is_big(L)) return 0;   // the one-workgroup core keeps everything in LDS
    return big_layout(L, big_leaf_L(L)).total;
}

// Padded core size 2^L: at least 64 (the LDS kernel ranks keys by wave
// ballots over groups of 64); padding adds fixed points only.
int lcsb_core_L(int32_t k)
{
    int L = 6;
    while ((1 << L) < k) ++L;
    return L;
}

// The big-core sequence for up to kBigBatch products of one padded size.
static int launch_big(const BigBatch &B, int cnt, hipStream_t stream)
{
    const int L = B.L, L0 = B.L0, d0 = L - L0, K = 1 << L;
    const dim3 lin((unsigned)(K / 256), 1, (unsigned)cnt), tiles((unsigned)(K / kSplitTile), 2, (unsigned)cnt);
    hipLaunchKernelGGL(big_init_kernel, lin, dim3(256), 0, stream, B);
    hipLaunchKernelGGL(big_split_count_kernel, tiles, dim3(256), 0, stream, B, 0);
    for (int d = 0; d < d0; ++d) {
        if (d + 1 < d0) hipLaunchKernelGGL(big_split_scatter_kernel<true>, tiles, dim3(256), 0, stream, B, d);
        else hipLaunchKernelGGL(big_split_scatter_kernel<false>, tiles, dim3(256), 0, stream, B, d);
    }
    hipLaunchKernelGGL(big_leaf_kernel, dim3(1u << d0, 1, (unsigned)cnt), dim3(kThreads), lds_core_bytes(L0),
                       stream, B);
    int flip = 0;
    for (int d = d0 - 1; d >= 0; --d) {
        const BigDepth D = big_depth(L, d);
        const int lines = (K / D.s) * D.nb, pieces = (K / D.s) * (2 * D.nb - 1);
        hipLaunchKernelGGL(big_lift_kernel, lin, dim3(256), 0, stream, B, d, flip);
        hipLaunchKernelGGL(big_hist_kernel, dim3((unsigned)(K / D.w), 2, (unsigned)cnt), dim3(256), 0, stream, B, d);
        hipLaunchKernelGGL(big_sweep_kernel, dim3((unsigned)((lines + kSweepWaves - 1) / kSweepWaves), 2, (unsigned)cnt),
                           dim3(64 * kSweepWaves), 0, stream,
                           B, d);
        hipLaunchKernelGGL(big_walk_kernel, dim3((unsigned)((pieces + 63) / 64), 1, (unsigned)cnt), dim3(64),
                           walk_lds_bytes(1 << D.glog), stream, B, d, flip);
        flip ^= 1;
    }
    hipLaunchKernelGGL(big_out_kernel, lin, dim3(256), 0, stream, B, flip);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

// The dynamic LDS limit of the core kernels, set once per process (before
// any capture: it is not a stream operation).
int lcsb_init(void)
{
    static std::once_flag once;
    static int rc = 0;
    std::call_once(once, [] {
        if (hipFuncSetAttribute((const void *)braid_core_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_core_bytes(kLdsCoreL)) != hipSuccess ||
            hipFuncSetAttribute((const void *)big_leaf_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_core_bytes(kLdsCoreL)) != hipSuccess ||
            hipFuncSetAttribute((const void *)big_walk_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)walk_lds_bytes(kMaxBoxG)) != hipSuccess)
            rc = -5;
    });
    return rc;
}

int lcsb_launch_core(const BraidCoreDesc *descs, int count, hipStream_t stream)
{
    if (lcsb_init()) return -5;
    // small cores: one workgroup each, descriptors by value in the arguments
    std::vector<BraidCoreDesc> small;
    std::vector<BraidCoreDesc> big;
    for (int i = 0; i < count; ++i) {
        if (descs[i].L > kMaxCoreL) return -34;
        (is_big(descs[i].L) ? big : small).push_back(descs[i]);
    }
    for (size_t base = 0; base < small.size(); base += kCoreBatch) {
        CoreBatch batch{};
        const int cnt = (int)std::min<size_t>(small.size() - base, kCoreBatch);
        int max_L = 0;
        for (int i = 0; i < cnt; ++i) {
            batch.d[i] = small[base + (size_t)i];
            max_L = std::max(max_L, (int)batch.d[i].L);
        }
        if (max_L > kLdsCoreL) return -34;
        hipLaunchKernelGGL(braid_core_lds_kernel, dim3(cnt), dim3(kThreads), lds_core_bytes(max_L), stream, batch);
    }
    // big cores: products of one padded size share each launch of the sequence
    std::sort(big.begin(), big.end(), [](const BraidCoreDesc &x, const BraidCoreDesc &y) { return x.L < y.L; });
    for (size_t i = 0; i < big.size();) {
        BigBatch B{};
        B.L = big[i].L;
        B.L0 = big_leaf_L(B.L);
        B.Y = big_layout(B.L, B.L0);
        for (int d = 0; d < B.L - B.L0; ++d) B.D[d] = big_depth(B.L, d);
        int cnt = 0;
        while (i < big.size() && big[i].L == B.L && cnt < kBigBatch) {
            B.c[cnt++] = BigCore{big[i].P, big[i].Q, big[i].R, big[i].scratch, big[i].k};
            ++i;
        }
        const int rc = launch_big(B, cnt, stream);
        if (rc) return rc;
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

// staggered_multiply for `count` independent products on device pointers.
// Their slicing, cores and assembly each go to one launch per kind (in
// groups of kMulBatch products for the slicing and assembly), so the
// products of a merge-tree level run side by side.  specs[i].ws holds
// lcsb_multiply_ws_words(p_size, q_size, k) words.
int lcsb_multiply_many_device(const MulSpec *specs, int count, hipStream_t stream)
{
    std::vector<BraidCoreDesc> cores;
    cores.reserve((size_t)count);
    auto nblocks = [](int32_t n) { return (n + kScanTile - 1) / kScanTile; };
    for (int base = 0; base < count; base += kMulBatch) {
        const int cnt = count - base < kMulBatch ? count - base : kMulBatch;
        InvertBatch ib{};
        SliceBatch sb{};
        int ni = 0, inv_blocks = 0, slice_blocks = 0;
        for (int i = 0; i < cnt; ++i) {
            const MulSpec &M = specs[base + i];
            const int32_t p_size = M.p_size, q_size = M.q_size, k = M.k;
            if (k == 0) continue;
            const int nmax = p_size > q_size ? p_size : q_size;
            const int nblocks_max = nblocks(nmax) + 1;
            int32_t *q_c2r = M.ws;                // q_size
            int32_t *map_row = q_c2r + q_size;    // k
            int32_t *map_col = map_row + k;       // k
            int32_t *Pred = map_col + k;          // k
            int32_t *Qred = Pred + k;             // k
            int32_t *core = Qred + k;             // k
            int32_t *counts_p = core + k;         // nblocks_max
            int32_t *counts_q = counts_p + nblocks_max;
            int32_t *core_ws = counts_q + nblocks_max;
            ib.s[ni].r2c = M.q;
            ib.s[ni].c2r = q_c2r;
            ib.s[ni].n = q_size;
            // vertical slice of p: rows whose column lies in the glued range
            sb.s[2 * ni] = SliceSpec{M.p, counts_p, map_row, Pred, p_size, M.row_glue ? 0 : p_size - k,
                                     M.row_glue ? k : p_size, k};
            // horizontal slice of q: columns whose row lies in the glued range
            sb.s[2 * ni + 1] = SliceSpec{q_c2r, counts_q, map_col, Qred, q_size, M.row_glue ? q_size - k : 0,
                                         M.row_glue ? q_size : k, k};
            ++ni;
            inv_blocks = std::max(inv_blocks, (q_size + 255) / 256);
            slice_blocks = std::max(slice_blocks, std::max(nblocks(p_size), nblocks(q_size)));
            BraidCoreDesc h{};
            h.P = Pred;
            h.Q = Qred;
            h.R = core;
            h.k = k;
            h.L = lcsb_core_L(k);
            h.scratch = core_ws;
            cores.push_back(h);
        }
        if (ni == 0) continue;
        hipLaunchKernelGGL(invert_kernel, dim3(inv_blocks, ni), dim3(256), 0, stream, ib);
        hipLaunchKernelGGL(slice_count_kernel, dim3(slice_blocks, 2 * ni), dim3(kScanBlock), 0, stream, sb);
        hipLaunchKernelGGL(scan_counts_kernel, dim3(2 * ni), dim3(1024), 0, stream, sb);
        hipLaunchKernelGGL(slice_scatter_kernel, dim3(slice_blocks, 2 * ni), dim3(kScanBlock), 0, stream, sb);
    }
    if (!cores.empty()) {
        const int rc = lcsb_launch_core(cores.data(), (int)cores.size(), stream);
        if (rc) return rc;
    }
    for (int base = 0; base < count; base += kMulBatch) {
        const int cnt = count - base < kMulBatch ? count - base : kMulBatch;
        AssembleBatch ab{};
        int blocks = 1;
        for (int i = 0; i < cnt; ++i) {
            const MulSpec &M = specs[base + i];
            int32_t *map_row = M.ws + M.q_size, *map_col = map_row + M.k, *core = map_col + 3 * (size_t)M.k;
            ab.a[i] = AssembleArgs{M.p, M.q, M.out, map_row, map_col, core, M.p_size, M.q_size, M.k, M.row_glue};
            const int n = std::max(std::max(M.p_size, M.q_size), M.k);
            blocks = std::max(blocks, (n + 255) / 256);
        }
        hipLaunchKernelGGL(assemble_kernel, dim3(blocks, cnt), dim3(256), 0, stream, ab);
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

// staggered_multiply on device pointers; ws must hold lcsb_multiply_ws_words(...)
int lcsb_multiply_device(const int32_t *p, int32_t p_size, const int32_t *q, int32_t q_size, int32_t k,
                         int32_t row_glue, int32_t *out, int32_t *ws, hipStream_t stream)
{
    const MulSpec M{p, p_size, q, q_size, k, row_glue, out, ws};
    return lcsb_multiply_many_device(&M, 1, stream);
}

size_t lcsb_multiply_ws_words(int32_t p_size, int32_t q_size, int32_t k)
{
    const int nmax = p_size > q_size ? p_size : q_size;
    const size_t nblocks_max = (size_t)(nmax + kScanTile - 1) / kScanTile + 1;
    return (size_t)q_size + 5 * (size_t)k + 2 * nblocks_max + lcsb_core_scratch_words(k > 0 ? k : 1);
}
