// Semi-local score queries read from a seaweed permutation.
//
// For the permutation P of a (m) against b (n) in row_to_col form (row =
// strand id: 0..m-1 enter the rows, m..m+n-1 the columns; column = exit:
// 0..n-1 the bottom, n..n+m-1 the right side), every semi-local score is
//     H(i, j) = (j - i) - D(m + i, j),   D(r0, c0) = #{ r >= r0 : P[r] < c0 },
// read at a fixed offset per quadrant (query_kernels.h).  Two ways to count D:
//
//  * scan: one wave per query, each lane counts a strided share of the rows
//    (HBM/L2-bound, size*4 bytes per query; no set-up);
//  * wavelet matrix: levels L = bit length of size over the values P + 1
//    (so unset rows, -1, become 0).  Level k holds bit L-1-k of the values
//    in that level's order; the next level is a stable partition, zeros
//    first.  D(r0, c) walks the levels with two ranks each (O(L) dependent
//    16-byte loads per query).  The build is three passes per level: ballot
//    the bits into 64-row words with block-local zero counts, scan the block
//    totals, then finish the counts and scatter the values to the next level.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "query_kernels.h"

namespace lcsq {

constexpr int kBuildThreads = 256;            // 4 waves
constexpr int kWordsPerWave = 16;
constexpr int kWordsPerBlock = 4 * kWordsPerWave;   // 64 words = 4096 rows

__global__ void scan_kernel(const int32_t *perm, int32_t size, int kind, int32_t m, int32_t n, const int32_t *q0,
                            const int32_t *q1, int32_t nq, int32_t *out)
{
    const int wave = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= nq) return;
    Window w;
    if (!query_window(kind, m, n, q0[wave], q1[wave], w)) {
        if (lane == 0) out[wave] = -1;
        return;
    }
    int cnt = 0;
    if (w.r0 < size)
        for (int32_t r = (int32_t)w.r0 + lane; r < size; r += 64) cnt += (int64_t)perm[r] < w.c0;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0) out[wave] = w.score ? w.base - cnt : cnt;
}

// values + 1 into level 0's order; flags any value outside [-1, size)
__global__ void index_prep_kernel(const int32_t *perm, int32_t size, int32_t *A, int32_t *bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= size) return;
    const int32_t x = perm[i];
    if (x < -1 || x >= size) *bad = 1;
    A[i] = x + 1;
}

__global__ __launch_bounds__(kBuildThreads) void index_bits_kernel(const int32_t *A, int32_t size, int bit,
                                                                   RankWord *wk, int32_t W, uint32_t *blocksum)
{
    __shared__ uint32_t wave_total[4];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t w0 = (int64_t)blockIdx.x * kWordsPerBlock + wave * kWordsPerWave;
    unsigned long long mine = 0;   // lane j < 16 keeps word w0 + j
    uint32_t mine_before = 0, run = 0;
    for (int j = 0; j < kWordsPerWave; ++j) {
        const int64_t idx = (w0 + j) * 64 + lane;
        const bool valid = idx < size;
        const bool one = valid && ((A[valid ? idx : 0] >> bit) & 1);
        const unsigned long long ones = __ballot(one), live = __ballot(valid);
        if (lane == j) {
            mine = ones;
            mine_before = run;
        }
        run += (uint32_t)__popcll(live & ~ones);
    }
    if (lane == 0) wave_total[wave] = run;
    __syncthreads();
    uint32_t off = 0;
    for (int q = 0; q < wave; ++q) off += wave_total[q];
    if (lane < kWordsPerWave && w0 + lane < W) {
        RankWord r;
        r.bits = mine;
        r.zeros_before = off + mine_before;   // block-local until index_place_kernel
        r.pad = 0;
        wk[w0 + lane] = r;
    }
    if (threadIdx.x == 0) blocksum[blockIdx.x] = wave_total[0] + wave_total[1] + wave_total[2] + wave_total[3];
}

// exclusive scan of the block totals in place (one workgroup); the level's
// zero count goes to *zeros_k
__global__ __launch_bounds__(1024) void index_scan_kernel(uint32_t *blocksum, int32_t nb, uint32_t *zeros_k)
{
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int64_t b0 = (int64_t)t * per;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < nb) s += blocksum[b0 + i];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint32_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0;
    for (int i = 0; i < per; ++i)
        if (b0 + i < nb) {
            const uint32_t x = blocksum[b0 + i];
            blocksum[b0 + i] = run;
            run += x;
        }
    if (t == 1023) *zeros_k = part[1023];
}

// one thread per row: global zero counts per word, and (Anext != null) the
// stable partition of the level's values into the next level's order
__global__ __launch_bounds__(kBuildThreads) void index_place_kernel(const int32_t *A, int32_t size, int bit,
                                                                    RankWord *wk, int32_t W,
                                                                    const uint32_t *blockpre,
                                                                    const uint32_t *zeros_k, int32_t *Anext)
{
    const int64_t idx = (int64_t)blockIdx.x * kBuildThreads + threadIdx.x;
    const int64_t w = idx >> 6;
    if (w >= W) return;   // whole waves: W words cover 64*W rows
    const int lane = threadIdx.x & 63;
    const RankWord r = wk[w];
    const uint32_t zb = r.zeros_before + blockpre[w / kWordsPerBlock];
    if (lane == 0) wk[w].zeros_before = zb;
    if (Anext && idx < size) {
        const int32_t x = A[idx];
        const unsigned long long below = lane ? (r.bits << (64 - lane)) : 0ull;
        const uint32_t zeros_below = zb + (uint32_t)(lane - __popcll(below));
        const int64_t dest = ((x >> bit) & 1) ? (int64_t)*zeros_k + (idx - zeros_below) : (int64_t)zeros_below;
        Anext[dest] = x;
    }
}

__device__ inline uint32_t rank0(const RankWord *wk, uint32_t i)
{
    const RankWord r = wk[i >> 6];
    const int o = i & 63;
    const unsigned long long below = o ? ((~r.bits) << (64 - o)) : 0ull;
    return r.zeros_before + (uint32_t)__popcll(below);
}

__global__ void index_query_kernel(const RankWord *words, const uint32_t *zeros, int levels, int32_t size, int kind,
                                   int32_t m, int32_t n, const int32_t *q0, const int32_t *q1, int32_t nq,
                                   int32_t *out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    Window w;
    if (!query_window(kind, m, n, q0[q], q1[q], w)) {
        out[q] = -1;
        return;
    }
    const int64_t W = ((int64_t)size >> 6) + 1;
    uint32_t s = (uint32_t)(w.r0 < size ? w.r0 : size), e = (uint32_t)size;
    const int64_t c = w.c0 + 1;   // values are stored shifted by one
    uint32_t cnt = 0;
    if (s >= e || c <= 0) {
        cnt = 0;
    } else if (c >= ((int64_t)1 << levels)) {
        cnt = e - s;
    } else {
        for (int k = 0; k < levels && s < e; ++k) {
            const RankWord *wk = words + (size_t)k * W;
            const uint32_t s0 = rank0(wk, s), e0 = rank0(wk, e);
            if ((c >> (levels - 1 - k)) & 1) {
                cnt += e0 - s0;
                const uint32_t z = zeros[k];
                s = z + (s - s0);
                e = z + (e - e0);
            } else {
                s = s0;
                e = e0;
            }
        }
    }
    out[q] = w.score ? w.base - (int32_t)cnt : (int32_t)cnt;
}

}  // namespace lcsq

using namespace lcsq;

static int launch_status() { return hipGetLastError() == hipSuccess ? 0 : LCS_EIO; }

int lcsq_launch_scan(const int32_t *perm, int32_t size, int kind, int32_t m, int32_t n, const int32_t *q0,
                     const int32_t *q1, int32_t nq, int32_t *out, hipStream_t s)
{
    if (nq <= 0) return 0;
    const int blocks = (int)(((int64_t)nq * 64 + 255) / 256);
    hipLaunchKernelGGL(scan_kernel, dim3(blocks), dim3(256), 0, s, perm, size, kind, m, n, q0, q1, nq, out);
    return launch_status();
}

int lcsq_launch_index_build(const int32_t *perm, int32_t size, int levels, RankWord *words, uint32_t *zeros,
                            int32_t *scratch, hipStream_t s)
{
    if (size <= 0) return 0;
    const int32_t W = (size >> 6) + 1;
    const int32_t nb = (W + kWordsPerBlock - 1) / kWordsPerBlock;
    int32_t *A = scratch, *B = scratch + size;
    uint32_t *blocksum = (uint32_t *)(scratch + 2 * (int64_t)size);
    int32_t *bad = scratch + 2 * (int64_t)size + nb;
    hipLaunchKernelGGL(index_prep_kernel, dim3((unsigned)(((int64_t)size + 255) / 256)), dim3(256), 0, s, perm,
                       size, A, bad);
    const unsigned place_blocks = (unsigned)(((int64_t)W * 64 + kBuildThreads - 1) / kBuildThreads);
    for (int k = 0; k < levels; ++k) {
        const int bit = levels - 1 - k;
        RankWord *wk = words + (size_t)k * W;
        hipLaunchKernelGGL(index_bits_kernel, dim3(nb), dim3(kBuildThreads), 0, s, A, size, bit, wk, W, blocksum);
        hipLaunchKernelGGL(index_scan_kernel, dim3(1), dim3(1024), 0, s, blocksum, nb, zeros + k);
        hipLaunchKernelGGL(index_place_kernel, dim3(place_blocks), dim3(kBuildThreads), 0, s, A, size, bit, wk, W,
                           blocksum, zeros + k, k + 1 < levels ? B : nullptr);
        int32_t *t = A;
        A = B;
        B = t;
    }
    return launch_status();
}

int lcsq_launch_index_query(const RankWord *words, const uint32_t *zeros, int levels, int32_t size, int kind,
                            int32_t m, int32_t n, const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out,
                            hipStream_t s)
{
    if (nq <= 0) return 0;
    hipLaunchKernelGGL(index_query_kernel, dim3((unsigned)(((int64_t)nq + 255) / 256)), dim3(256), 0, s, words,
                       zeros, levels, size, kind, m, n, q0, q1, nq, out);
    return launch_status();
}
