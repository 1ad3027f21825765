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
//    column T(r) where delta >= 0.  One wave finds T and delta at each chunk
//    boundary with a single pass over the node's columns; then one thread
//    per chunk walks it with the reference's incremental updates and places
//    the good points.
// The union of the children's points and the per-column view live in LDS
// (16-bit entries with the lo/hi flag in bit 15), so the walks run at LDS
// latency.  Cores up to K = 32768 take this path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "braid_kernels.h"

namespace lcsb {

constexpr int kThreads = 1024;
constexpr int kChunkRows = 32;
constexpr uint16_t kLoFlag = 0x8000u;
constexpr uint16_t kIdx = 0x7fffu;
// cores up to this size keep keys and products in LDS (braid_core_lds_kernel)
constexpr int kLdsCoreK = 16384;
// row chunks per node are at most 128 (chunk height max(32, s/128)); each wave
// keeps 2 x kMaxChunks bucket counters in LDS for the boundary sweeps
constexpr int kMaxChunks = 128;
constexpr size_t lds_core_bytes(int K) { return (size_t)4 * K * sizeof(uint16_t) + ((size_t)K / 64 + 1) * 4; }

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

// Exclusive prefix of the 0/1 flags of keys[0..K) (bit `bit`), written to
// S[0..K] (S[K] = total).  Each thread owns a contiguous run of elements.
__device__ void block_scan_bits(const int32_t *keys, int K, int bit, int32_t *S, int32_t *lds_sums)
{
    const int t = threadIdx.x;
    const int per = (K + kThreads - 1) / kThreads;
    const int lo = min(t * per, K), hi = min(lo + per, K);
    int cnt = 0;
    for (int x = lo; x < hi; ++x) cnt += (keys[x] >> bit) & 1;
    lds_sums[t] = cnt;
    __syncthreads();
    // Hillis-Steele over the 1024 per-thread sums
    for (int off = 1; off < kThreads; off <<= 1) {
        int v = t >= off ? lds_sums[t - off] : 0;
        __syncthreads();
        lds_sums[t] += v;
        __syncthreads();
    }
    int run = lds_sums[t] - cnt;  // exclusive
    for (int x = lo; x < hi; ++x) {
        S[x] = run;
        run += (keys[x] >> bit) & 1;
    }
    if (t == kThreads - 1) S[K] = lds_sums[kThreads - 1];
    __syncthreads();
}

// One depth of the top-down split: stable partition of every aligned segment
// of size s by bit `bit` of the key (zeros first).  par[pos] = old position.
__device__ void split_depth(const int32_t *key, int32_t *key_new, int32_t *par, int K, int s, int bit, int32_t *S,
                            int32_t *lds_sums)
{
    block_scan_bits(key, K, bit, S, lds_sums);
    for (int x = threadIdx.x; x < K; x += kThreads) {
        const int seg = x & ~(s - 1);
        const int b = (key[x] >> bit) & 1;
        const int ones_before = S[x] - S[seg];
        const int zeros_total = s - (S[seg + s] - S[seg]);
        const int pos = b ? seg + zeros_total + ones_before : seg + (x - seg - ones_before);
        key_new[pos] = key[x];
        par[pos] = x;
    }
    __syncthreads();
}

// One bottom-up depth (node size s) of the core: lift the children's
// products into node coordinates (the reference's inverse_mapping, :191-202)
// and run the chunked ant passage (:123-189) for every node of the depth.
// U, CI are in LDS; R (children's products) and Rn (this depth's) are
// either global int32 or LDS uint16 arrays; pp, pq map depth d+1 positions
// to depth d positions; bT, bD hold K/32 chunk entries.
template <typename UT, typename RT, typename RNT>
__device__ __forceinline__ void bottom_up_depth(UT *U, UT *CI, const RT *R, RNT *Rn, const int32_t *__restrict__ pp,
                                                const int32_t *__restrict__ pq, int K, int s, int32_t *bT,
                                                int32_t *bD, int *hist)
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
    // chunk height: 32 rows, or s/128 for big nodes, balancing the
    // boundary scans (O(s) each) against the walks (O(rows) each)
    const int rc = s < kChunkRows ? s : (s / 128 > kChunkRows ? s / 128 : kChunkRows);
    const int rc_log = __builtin_ctz(rc);   // rc and s are powers of two
    const int nch = s / rc;            // chunks per node
    const int total = K / rc;          // chunks at this depth
    // chunk entry points: chunk j of a node covers rows [s-(j+1)rc, s-j rc).
    // A node's first chunk (j = 0) enters at T = 0; its entry is computed in
    // the walk, so only chunks j > 0 (nodes of 64 rows or more, at most K/32
    // chunks) store entries in bT, bD.  For chunk j > 0 entering at row
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
        auto range_counts = [&](int v0, int c_begin, int c_end, int &lp0, int &lp1, int &cnt0, int &cnt1) {
            hlo[j0] = 0;
            hlo[j1] = 0;
            hhi[j0] = 0;
            hhi[j1] = 0;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            for (int c0 = c_begin; c0 < c_end; c0 += 64) {
                const uint16_t e = CI[v0 + c0 + lane];
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
        auto find_T = [&](int v0, int c_begin, int c_end, int Y0, int Y1, int Z0, int Z1, bool open0, bool open1,
                          int &T0, int &T1) {
            for (int c0 = c_begin; c0 < c_end && __ballot(open0 || open1); c0 += 64) {
                hlo[j0] = 0;
                hlo[j1] = 0;
                hhi[j0] = 0;
                hhi[j1] = 0;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                const uint16_t e = CI[v0 + c0 + lane];
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
        if (nodes >= kWaves || K < 64 * kWaves) {
            // one wave per node: Y_j from the node's counts, then one sweep
            for (int node = wave; node < nodes; node += kWaves) {
                const int v0 = node * s;
                int Y0, Y1, c0_, c1_;
                range_counts(v0, 0, s, Y0, Y1, c0_, c1_);
                int T0 = 0, T1 = 0;
                find_T(v0, 0, s, Y0, Y1, 0, 0, j0 >= 1 && j0 < nch && Y0 > 0, j1 < nch && Y1 > 0, T0, T1);
                if (j0 >= 1 && j0 < nch) entry(node, v0, j0, T0);
                if (j1 < nch) entry(node, v0, j1, T1);
            }
        } else {
            // W = kWaves / nodes waves per node, each on s / W consecutive
            // columns: every wave counts its segment, the segments' counts
            // give Y_j and each segment's starting Z_j, and only the wave in
            // whose segment Z_j reaches Y_j sweeps for T_j.
            const int W = kWaves / nodes, node = wave / W, seg = wave % W, seg_len = s / W;
            const int v0 = node * s, cb = seg * seg_len;
            int lp0, lp1, cnt0, cnt1;
            range_counts(v0, cb, cb + seg_len, lp0, lp1, cnt0, cnt1);
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
            find_T(v0, cb, cb + seg_len, Y0, Y1, Zs0, Zs1, mine0, mine1, T0, T1);
            if (j0 >= 1 && j0 < nch && (mine0 || (Y0 == 0 && seg == 0))) entry(node, v0, j0, T0);
            if (j1 < nch && (mine1 || (Y1 == 0 && seg == 0))) entry(node, v0, j1, T1);
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
    for (int ch = threadIdx.x; ch < total; ch += kThreads) {
        const int v0 = (ch / nch) * s, j = ch % nch;
        const int r_hi = s - j * rc, r_lo = r_hi - rc;
        int dr = r_hi - 1, dc, delta;
        if (j == 0) {
            // T(s) = 0, delta(s, 0) = 0; entering row s-1 subtracts it when it is lo
            dc = 0;
            delta = (U[v0 + s - 1] & kLoFlag) ? -1 : 0;
        } else {
            dc = bT[ch];
            delta = bD[ch];
        }
        // entries outside the node (dc = 0 or s, dr = 0) are read clamped and never used
        uint32_t cl = CI[v0 + max(dc - 1, 0)], cc = CI[v0 + min(dc, s - 1)];
        uint32_t ud = U[v0 + dr], ur = U[v0 + max(dr - 1, 0)];
        while (dr >= r_lo) {
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

// The sticky product of the k-permutations P (row -> shared column) and Q
// (shared row -> column) for every descriptor; one workgroup per product.
// Keys and products in global memory: cores up to K = 32768.
__global__ __launch_bounds__(kThreads) void braid_core_kernel(CoreBatch batch)
{
    const BraidCoreDesc D = batch.d[blockIdx.x];
    const int k = D.k, L = D.L, K = 1 << L;
    extern __shared__ uint16_t lds[];
    uint16_t *U = lds;            // node-local row -> node-local col | lo flag
    uint16_t *CI = lds + K;       // node-local col -> node-local row | lo flag
    __shared__ int32_t sums[kThreads];
    __shared__ int hist[(kThreads / 64) * 2 * kMaxChunks];
    int32_t *keyP = D.scratch, *keyQ = keyP + K, *keyP2 = keyQ + K, *keyQ2 = keyP2 + K;
    int32_t *S = keyQ2 + K;               // K+1
    int32_t *R = S + K + 1, *Rn = R + K;  // per depth products
    int32_t *bT = Rn + K, *bD = bT + K;   // per chunk entry column / delta
    int32_t *parP = bD + K;               // L * K
    int32_t *parQ = parP + (size_t)L * K; // L * K

    // depth 0: P ordered by row (key = shared column), Q by column (key = shared row)
    for (int x = threadIdx.x; x < K; x += kThreads) {
        keyP[x] = x < k ? D.P[x] : x;
        if (x >= k) keyQ[x] = x;
    }
    for (int r = threadIdx.x; r < k; r += kThreads)
        if ((uint32_t)D.Q[r] < (uint32_t)k) keyQ[D.Q[r]] = r;
    __syncthreads();
    for (int d = 0; d < L; ++d) {
        const int s = K >> d, bit = L - d - 1;
        split_depth(keyP, keyP2, parP + (size_t)d * K, K, s, bit, S, sums);
        split_depth(keyQ, keyQ2, parQ + (size_t)d * K, K, s, bit, S, sums);
        int32_t *t0 = keyP; keyP = keyP2; keyP2 = t0;
        int32_t *t1 = keyQ; keyQ = keyQ2; keyQ2 = t1;
    }
    for (int x = threadIdx.x; x < K; x += kThreads) R[x] = 0;  // depth L: 1x1 nodes
    __syncthreads();
    for (int d = L - 1; d >= 0; --d) {
        bottom_up_depth(U, CI, R, Rn, parP + (size_t)d * K, parQ + (size_t)d * K, K, K >> d, bT, bD, hist);
        int32_t *tmp = R; R = Rn; Rn = tmp;
    }
    for (int x = threadIdx.x; x < k; x += kThreads) D.R[x] = R[x];
}

// One depth of the top-down split with the keys in LDS: a stable partition
// of every aligned segment of size s by bit `bit` (zeros first), the ranks
// from wave ballots and one scan of the 64-key group counts.
// par[pos] = old position (global).  K >= 64.  A wave takes the groups
// wave, wave + 16, ...; U of them at a time, with every load of the U groups
// issued before the first ballot, so their LDS latencies overlap (the
// compiler does not unroll a loop around a ballot by itself).
template <int U>
__device__ __forceinline__ void split_count_groups(const uint16_t *key, int bit, int32_t *gpre, int g0, int lane)
{
    uint16_t kv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) kv[u] = key[(g0 + u * (kThreads / 64)) * 64 + lane];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const unsigned long long m = __ballot((kv[u] >> bit) & 1);
        if (lane == 0) gpre[g0 + u * (kThreads / 64)] = __popcll(m);
    }
}

template <int U>
__device__ __forceinline__ void split_place_groups(const uint16_t *key, uint16_t *key_new, int32_t *par, int s,
                                                   int bit, const int32_t *gpre, int g0, int lane)
{
    const unsigned long long lt = (1ull << lane) - 1;
    uint16_t kv[U];
    int gp[U], gs0[U], gs1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int g = g0 + u * (kThreads / 64), x = g * 64 + lane;
        kv[u] = key[x];
        if (s >= 64) {
            const int gs = (x & ~(s - 1)) >> 6;
            gp[u] = gpre[g];
            gs0[u] = gpre[gs];
            gs1[u] = gpre[gs + (s >> 6)];
        }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int g = g0 + u * (kThreads / 64), x = g * 64 + lane;
        const int b = (kv[u] >> bit) & 1;
        const unsigned long long m = __ballot(b);
        const int seg = x & ~(s - 1);
        int before, ones;
        if (s >= 64) {
            before = gp[u] - gs0[u] + __popcll(m & lt);
            ones = gs1[u] - gs0[u];
        } else {
            const unsigned long long segmask = ((1ull << s) - 1) << (lane & ~(s - 1));
            before = __popcll(m & lt & segmask);
            ones = __popcll(m & segmask);
        }
        const int pos = b ? seg + (s - ones) + before : seg + (x - seg - before);
        key_new[pos] = kv[u];
        par[pos] = x;
    }
}

__device__ void split_depth_lds(const uint16_t *key, uint16_t *key_new, int32_t *par, int K, int s, int bit,
                                int32_t *gpre)
{
    constexpr int kW = kThreads / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ngroups = K >> 6;
    int g = wave;
    for (; g + 3 * kW < ngroups; g += 4 * kW) split_count_groups<4>(key, bit, gpre, g, lane);
    for (; g < ngroups; g += kW) split_count_groups<1>(key, bit, gpre, g, lane);
    __syncthreads();
    if (wave == 0) {  // exclusive scan of the group counts, gpre[ngroups] = total
        int carry = 0;
        for (int base = 0; base < ngroups; base += 64) {
            const int c = base + lane < ngroups ? gpre[base + lane] : 0;
            const int incl = wave_incl_scan(c);
            if (base + lane < ngroups) gpre[base + lane] = carry + incl - c;
            carry += __builtin_amdgcn_readlane(incl, 63);
        }
        if (lane == 0) gpre[ngroups] = carry;
    }
    __syncthreads();
    g = wave;
    for (; g + 3 * kW < ngroups; g += 4 * kW) split_place_groups<4>(key, key_new, par, s, bit, gpre, g, lane);
    for (; g < ngroups; g += kW) split_place_groups<1>(key, key_new, par, s, bit, gpre, g, lane);
    __syncthreads();
}

// The same product with keys, both products and the chunk entries in LDS:
// cores up to K = kLdsCoreK.  LDS: 4K uint16 (keys / U, CI, R, Rn) plus
// K/64+1 group counts.
__global__ __launch_bounds__(kThreads) void braid_core_lds_kernel(CoreBatch batch)
{
    const BraidCoreDesc D = batch.d[blockIdx.x];
    const int k = D.k, L = D.L, K = 1 << L;
    extern __shared__ uint16_t lds[];
    __shared__ int32_t bT[kLdsCoreK / kChunkRows], bD[kLdsCoreK / kChunkRows];
    __shared__ int hist[(kThreads / 64) * 2 * kMaxChunks];
    uint16_t *key = lds, *key2 = lds + K;             // top-down (P, then Q)
    uint16_t *U = lds, *CI = lds + K;                 // bottom-up (reuses the key space)
    uint16_t *R = lds + 2 * (size_t)K, *Rn = lds + 3 * (size_t)K;
    int32_t *gpre = (int32_t *)(lds + 4 * (size_t)K);
    int32_t *parP = D.scratch;                        // L * K
    int32_t *parQ = parP + (size_t)L * K;             // L * K

    // P ordered by row (key = shared column)
    for (int x = threadIdx.x; x < K; x += kThreads) key[x] = (uint16_t)(x < k ? D.P[x] : x);
    __syncthreads();
    for (int d = 0; d < L; ++d) {
        split_depth_lds(key, key2, parP + (size_t)d * K, K, K >> d, L - d - 1, gpre);
        uint16_t *t = key; key = key2; key2 = t;
    }
    // Q ordered by column (key = shared row)
    key = lds;
    key2 = lds + K;
    for (int x = threadIdx.x; x < K; x += kThreads)
        if (x >= k) key[x] = (uint16_t)x;
    for (int r = threadIdx.x; r < k; r += kThreads)
        if ((uint32_t)D.Q[r] < (uint32_t)k) key[D.Q[r]] = (uint16_t)r;
    __syncthreads();
    for (int d = 0; d < L; ++d) {
        split_depth_lds(key, key2, parQ + (size_t)d * K, K, K >> d, L - d - 1, gpre);
        uint16_t *t = key; key = key2; key2 = t;
    }
    for (int x = threadIdx.x; x < K; x += kThreads) R[x] = 0;  // depth L: 1x1 nodes
    __syncthreads();
    for (int d = L - 1; d >= 0; --d) {
        bottom_up_depth(U, CI, R, Rn, parP + (size_t)d * K, parQ + (size_t)d * K, K, K >> d, bT, bD, hist);
        uint16_t *tmp = R; R = Rn; Rn = tmp;
    }
    for (int x = threadIdx.x; x < k; x += kThreads) D.R[x] = R[x];
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

size_t lcsb_core_scratch_words(int32_t k)
{
    const int L = lcsb_core_L(k);
    const size_t K = (size_t)1 << L;
    return (size_t)9 * K + 1 + (size_t)2 * L * K;
}

// Padded core size 2^L: at least 64 (the LDS kernel ranks keys by wave
// ballots over groups of 64); padding adds fixed points only.
int lcsb_core_L(int32_t k)
{
    int L = 6;
    while ((1 << L) < k) ++L;
    return L;
}

int lcsb_launch_core(const BraidCoreDesc *descs, int count, hipStream_t stream)
{
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)braid_core_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(2 * kMaxCoreK * sizeof(uint16_t))) != hipSuccess ||
            hipFuncSetAttribute((const void *)braid_core_lds_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_core_bytes(kLdsCoreK)) != hipSuccess)
            return -5;
        attr_set = true;
    }
    // descriptors travel by value in the kernel arguments
    for (int base = 0; base < count; base += kCoreBatch) {
        CoreBatch batch{};
        const int cnt = count - base < kCoreBatch ? count - base : kCoreBatch;
        int max_K = 1;
        for (int i = 0; i < cnt; ++i) {
            batch.d[i] = descs[base + i];
            if ((1 << batch.d[i].L) > max_K) max_K = 1 << batch.d[i].L;
        }
        if (max_K > kMaxCoreK) return -34;
        if (max_K <= kLdsCoreK)
            hipLaunchKernelGGL(braid_core_lds_kernel, dim3(cnt), dim3(kThreads), lds_core_bytes(max_K), stream, batch);
        else
            hipLaunchKernelGGL(braid_core_kernel, dim3(cnt), dim3(kThreads), (size_t)2 * max_K * sizeof(uint16_t),
                               stream, batch);
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
