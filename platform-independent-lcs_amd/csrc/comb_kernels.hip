// Seaweed (strand) combing kernels for CDNA4 / gfx950.
//
// Grid orientation used on the device: row r = 0..m-1 holds a[r] (the
// reference's reversed row i = m-1-r, src/lcs_reference.hpp:13-20), column c
// holds b[c].  Horizontal strands travel left to right along a row, vertical
// strands travel top to bottom (from row r-1 into row r, the reference's "from
// high i to low i", src/lcs_common.hpp:76-88).  Cell rule
// (src/lcs_common.hpp:7-24): swap iff a[r] == b[c] || h > v.  Strand ids are
// non-negative, so the signed compare equals the unsigned one and the swap is
//     h' = min(match ? ~0u : h, v),   v' = h ^ v ^ h'
// i.e. v_cmp_eq + v_min_u32 + v_cndmask + v_bitop3 (xor3): 4 VALU, no SALU.
//
// Band chain.  One wave64 owns a band of 64*TM consecutive rows; lane L keeps
// TM rows (symbols and h strands) in VGPRs.  At step t lane L combs column
// t-L, so the wave sweeps the band along anti-diagonals.  The v strand and the
// b symbol move one lane down per step by DPP wave_shr:1; lane 0 receives the
// band's inputs from two "conveyor" registers shifted by wave_shl:1, and the v
// strand leaving lane 63 is caught by a third conveyor.  Nothing touches LDS.
// Every 16 steps the wave publishes the 16 newest bottom-edge strands of its
// band to a ring in global memory as 64-bit (tag, strand) words stored with
// device-scope relaxed atomics; the band below polls for words carrying its
// tag, so no fence, flag or cache flush is needed.  Bands are handed out by an
// atomic ticket in launch order, so a wave only ever waits for a band that a
// resident wave already owns: any grid size makes progress.  Every wait is
// bounded in time; on expiry the kernel raises an abort flag that every wave
// checks, and all waves exit.
//
// Residuals (the reference's CPU fixup, src/lcs_grid.hpp:451-493) are handled
// in the kernel: lanes whose column is outside [0, n) skip the cells, rows
// beyond m in the last band never match and hold the smallest strand id, so
// they pass every v strand through unchanged.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cell_asm.h"
#include "comb_kernels.h"

namespace lcsk {

constexpr int kWave = 64;
constexpr int kDppWaveShl1 = 0x130;  // lane L reads lane L+1
constexpr int kDppWaveShr1 = 0x138;  // lane L reads lane L-1

__device__ __forceinline__ uint32_t shl1_zero(uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, kDppWaveShl1, 0xf, 0xf, true);
}
__device__ __forceinline__ int32_t shl1_zero_i(int32_t x)
{
    return __builtin_amdgcn_mov_dpp(x, kDppWaveShl1, 0xf, 0xf, true);
}
// lane L <- x[L-1]; lane 0 keeps `old`
__device__ __forceinline__ uint32_t shr1_keep(uint32_t old_, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old_, (int)x, kDppWaveShr1, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t shr1_keep_i(int32_t old_, int32_t x)
{
    return __builtin_amdgcn_update_dpp(old_, x, kDppWaveShr1, 0xf, 0xf, false);
}
// lane L <- x[L+1]; lane 63 keeps `old`
__device__ __forceinline__ uint32_t shl1_keep(uint32_t old_, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old_, (int)x, kDppWaveShl1, 0xf, 0xf, false);
}

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned long long *p, unsigned long long x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_flag(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void raise_abort(unsigned *p)
{
    __hip_atomic_store(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One cell.  `sb` is the column symbol; `valid` is false for padding rows.
template <bool ROWMASK>
__device__ __forceinline__ void cell(int32_t sa, int32_t sb, bool valid, uint32_t &h, uint32_t &v)
{
    bool match = (sa == sb);
    if (ROWMASK) match = match && valid;
    uint32_t hx = match ? 0xffffffffu : h;
    uint32_t hn = min(hx, v);
    v = __builtin_amdgcn_bitop3_b32(h, v, hn, 0x96);
    h = hn;
}

// All TM rows of one column for one lane.
template <int TM, bool ROWMASK>
__device__ __forceinline__ void column_cells(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                             int32_t sb, uint32_t &v)
{
    // Rows go through hand-scheduled blocks of up to 16 (cell_asm.h); one
    // block per column keeps the whole chain free of hazard nops.
    if constexpr (!ROWMASK && TM >= 4) {
        constexpr int R = TM >= 16 ? 16 : TM;
#pragma unroll
        for (int k = 0; k < TM; k += R) {
            uint32_t hn[R];
            if constexpr (R == 16) cell_block16(&a[k], &h[k], hn, sb, v);
            else if constexpr (R == 8) cell_block8(&a[k], &h[k], hn, sb, v);
            else cell_block4(&a[k], &h[k], hn, sb, v);
#pragma unroll
            for (int q = 0; q < R; ++q) h[k + q] = hn[q];
        }
    } else {
#pragma unroll
        for (int k = 0; k < TM; ++k) cell<ROWMASK>(a[k], sb, valid[k], h[k], v);
    }
}

// Steps t0 .. t0+C-1 of a band.  COLMASK: some lane's column leaves [0, n).
template <int TM, int C, bool ROWMASK, bool COLMASK>
__device__ __forceinline__ void band_steps(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                           uint32_t &v, int32_t &bq, uint32_t &O, uint32_t I, int32_t IB,
                                           int t0, int lane, int n)
{
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const uint32_t vin = I;
        I = shl1_zero(I);
        const int32_t bin = IB;
        IB = shl1_zero_i(IB);
        const uint32_t vn = shr1_keep(vin, v);   // lane 0 takes the band input
        O = shl1_keep(v, O);                     // lane 63's last output joins O
        v = vn;
        bq = shr1_keep_i(bin, bq);
        bool active = true;
        if (COLMASK) {
            const int j = t0 + s - lane;
            active = (j >= 0) && (j < n);
        }
        if (active) column_cells<TM, ROWMASK>(a, h, valid, bq, v);
    }
}

template <int TM, int C, bool ROWMASK>
__device__ bool band_pass(const ChainArgs &A, int g, int lane)
{
    const int m = A.m, n = A.n;
    const int r0 = g * (kWave * TM) + lane * TM;
    int32_t a[TM];
    uint32_t h[TM];
    bool valid[TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        const int r = r0 + k;
        valid[k] = ROWMASK ? (r < m) : true;
        a[k] = valid[k] ? A.a[r] : 0;
        h[k] = valid[k] ? (A.h_in ? (uint32_t)A.h_in[m - 1 - r] : (uint32_t)(m - 1 - r)) : 0u;
    }
    const bool first = (g == 0);
    const bool last = (g == A.nbands - 1);
    const uint32_t tag_in = (uint32_t)g;
    const unsigned long long tag_out = (unsigned long long)(g + 1) << 32;
    const unsigned long long *Ein = A.E + (size_t)(g & 1) * (size_t)n;
    unsigned long long *Eout = A.E + (size_t)((g + 1) & 1) * (size_t)n;
    const int nchunks = (n + kWave + C - 1) / C;

    // prefetch chunk 0
    int32_t pB = 0;
    unsigned long long pE = 0;
    {
        const int c = lane;
        const bool need = lane < C && c < n;
        if (need) pB = A.b[c];
        if (!first && need) pE = ld_relaxed(Ein + c);
    }
    uint32_t v = 0, O = 0;
    int32_t bq = 0;
    for (int k = 0; k < nchunks; ++k) {
        const int t0 = k * C;
        uint32_t I;
        const int32_t IB = pB;
        {
            const int c = t0 + lane;
            const bool need = lane < C && c < n;
            if (first) {
                I = A.v_in ? (need ? (uint32_t)A.v_in[c] : 0u) : (uint32_t)(m + c);
            } else {
                unsigned long long e = pE;
                unsigned long long t_start = 0;
                for (;;) {
                    const bool ok = !need || (uint32_t)(e >> 32) == tag_in;
                    if (__builtin_amdgcn_ballot_w64(!ok) == 0) break;
                    if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) return false;
                    const unsigned long long now = __builtin_amdgcn_s_memrealtime();
                    if (t_start == 0) t_start = now;
                    else if (now - t_start > A.timeout_ticks) { raise_abort(A.abort_flag); return false; }
                    __builtin_amdgcn_s_sleep(2);
                    if (!ok) e = ld_relaxed(Ein + c);
                }
                I = (uint32_t)e;
            }
        }
        {   // prefetch chunk k+1
            const int c = t0 + C + lane;
            const bool need = lane < C && c < n;
            pB = need ? A.b[c] : 0;
            if (!first) pE = need ? ld_relaxed(Ein + c) : 0ull;
        }
        if (t0 >= kWave - 1 && t0 + C <= n)
            band_steps<TM, C, ROWMASK, false>(a, h, valid, v, bq, O, I, IB, t0, lane, n);
        else
            band_steps<TM, C, ROWMASK, true>(a, h, valid, v, bq, O, I, IB, t0, lane, n);
        // O lanes [64-C, 64) hold the bottom-edge strands of columns
        // t0-64 .. t0-64+C-1 (captured one step late).
        const int col = t0 + C - 2 * kWave + lane;
        if (lane >= kWave - C && col >= 0 && col < n) {
            if (last) A.v_out[col] = (int32_t)O;
            else st_relaxed(Eout + col, tag_out | (unsigned long long)O);
        }
    }
#pragma unroll
    for (int k = 0; k < TM; ++k)
        if (valid[k]) A.h_out[m - 1 - (r0 + k)] = (int32_t)h[k];
    return true;
}

template <int TM>
__global__ __launch_bounds__(256) void comb_chain_kernel(ChainArgs A)
{
    constexpr int C = 16;
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = atomicAdd(A.ticket, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int)t >= A.nbands) break;
        if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) break;
        const int g = (int)t;
        const bool rowmask = (long long)(g + 1) * (kWave * TM) > (long long)A.m;
        const bool ok = rowmask ? band_pass<TM, C, true>(A, g, lane) : band_pass<TM, C, false>(A, g, lane);
        if (!ok) break;
    }
}

// Clears the ring tags and the ticket/abort words before a chain launch (a
// kernel, not hipMemsetAsync, so the pair can live inside a hipGraph).
__global__ void comb_prep_kernel(unsigned long long *E, size_t words, unsigned *ctrl, int nctrl)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) E[i] = 0ull;
    if (blockIdx.x == 0 && (int)threadIdx.x < nctrl) ctrl[threadIdx.x] = 0u;
}

// Identity strands for empty grids (m == 0 or n == 0).
__global__ void comb_identity_kernel(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) h_out[i] = i;
    if (i < n) v_out[i] = m + i;
}

}  // namespace lcsk

using namespace lcsk;

int lcsk_launch_chain(const ChainArgs &args, int tile_m, int grid_blocks, int block_threads, hipStream_t stream)
{
    const size_t words = 2ull * (size_t)args.n;
    int prep_blocks = (int)((words + 255) / 256);
    if (prep_blocks < 1) prep_blocks = 1;
    if (prep_blocks > 2048) prep_blocks = 2048;
    hipLaunchKernelGGL(comb_prep_kernel, dim3(prep_blocks), dim3(256), 0, stream, args.E, words, args.ticket, 2);
    switch (tile_m) {
    case 1: hipLaunchKernelGGL(comb_chain_kernel<1>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    case 2: hipLaunchKernelGGL(comb_chain_kernel<2>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    case 4: hipLaunchKernelGGL(comb_chain_kernel<4>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    case 8: hipLaunchKernelGGL(comb_chain_kernel<8>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    case 16: hipLaunchKernelGGL(comb_chain_kernel<16>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    case 32: hipLaunchKernelGGL(comb_chain_kernel<32>, dim3(grid_blocks), dim3(block_threads), 0, stream, args); break;
    default: return -22;
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_identity(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n, hipStream_t stream)
{
    const int mx = m > n ? m : n;
    if (mx == 0) return 0;
    hipLaunchKernelGGL(comb_identity_kernel, dim3((mx + 255) / 256), dim3(256), 0, stream, h_out, m, v_out, n);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
