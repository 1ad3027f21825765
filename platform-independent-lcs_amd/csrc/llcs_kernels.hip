// Score-only LCS length by bit-parallel combing on gfx950.
//
// The reference's binary bit-parallel LLCS (src/lcs_reference.hpp:135-211,
// disabled by #if 0) packs strands into machine words.  This kernel uses the
// word-parallel LCS recurrence over the row string a (Allison-Dix; Hyyro's
// form): with V an m-bit vector, initially all ones, and M[c] the bits i with
// a[i] == c, every symbol c of the column string does
//     U = V & M[c];   V = (V + U) | (V & ~M[c])
// and LCS = number of zero bits of V at the end.  The addition carries from
// row i to row i+1, which is the only coupling between rows.
//
// MI355X mapping: one wave64 per band of 2048 rows, one 32-bit word of V per
// lane (rows 32l..32l+31 of the band).  Lanes are skewed like the seaweed
// band chain: lane l takes column symbol k-l at step k, so the carry leaving
// lane l-1 at step k-1 is the carry entering lane l at step k.  Carries live
// as a 64-bit lane mask in SGPRs: one v_addc_co_u32 per step takes the
// shifted mask as carry-in and writes the new carry-out mask, and the shift
// and the band's own carry-in/carry-out bits are scalar work.  A step is one
// v_and, one v_addc_co, one 3-input bit op and two LDS reads per 2048 cells.
//
//  * Match masks: each band builds M[c][lane] for c in [0, sigma) in LDS
//    once; the all-zero row just below the table (row -1) serves the
//    padding symbols before and after the column string.
//  * Column symbols: per chunk of 32 steps the 95 symbols the lanes need are
//    staged in LDS as ready-made table addresses (c*256 + const - 4p), so a
//    lane's mask read is two dependent ds_reads with immediate offsets and no
//    address arithmetic.
//  * Band hand-off: lane 63's carry-out bits, 32 per chunk, go to the band
//    above as one tagged 64-bit word (tag << 32 | bits) stored with an
//    agent-scope relaxed atomic; the band above polls for its tag, bounded in
//    time, with the abort flag of the comb kernels.  Consumer chunk q needs
//    words q+1 and q+2 of its producer (the skew of 63 steps is not a
//    multiple of 32).
//  * Bands are handed out by an atomic ticket in order, so every wait is on a
//    band that a resident wave already holds.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "llcs_kernels.h"

namespace lcsl {

#define LCSL_GLOBAL __attribute__((address_space(1)))
typedef LCSL_GLOBAL unsigned long long gword;
typedef LCSL_GLOBAL const unsigned long long gcword;
typedef __attribute__((address_space(3))) uint32_t lds_u32;

constexpr int kStage = kChunk + kWave - 1;   // 95 symbols staged per chunk
constexpr int kStageWords = 96;
constexpr int kZeroRowWords = kWave;         // row -1 of the table, all zero
constexpr unsigned kNoTag = 0xffffffffu;

__device__ __forceinline__ unsigned long long ld_relaxed(gcword *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(gword *p, unsigned long long x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_flag(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every lane loads the same ring word; readfirstlane makes the tag test and
// the value provably wave-uniform, so every branch on them is a scalar branch
// and the lanes of a wave never part (a per-lane exit would let the loop
// around a band run with part of the wave).
__device__ __forceinline__ uint32_t tag_of(unsigned long long e) { return __builtin_amdgcn_readfirstlane((uint32_t)(e >> 32)); }
__device__ __forceinline__ uint32_t bits_of(unsigned long long e) { return __builtin_amdgcn_readfirstlane((uint32_t)e); }

// Waits until ring word `p` carries `tag` (bounded in time).  Returns the
// word's 32 bits, or raises the abort flag and clears `ok`.
__device__ __forceinline__ uint32_t wait_word(const LlcsArgs &A, gcword *p, unsigned long long e, uint32_t tag,
                                              bool &ok)
{
    if (tag_of(e) == tag) return bits_of(e);
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) {
            ok = false;
            return 0;
        }
        if (__builtin_amdgcn_s_memrealtime() - t0 > A.timeout_ticks) {
            __hip_atomic_store(A.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = false;
            return 0;
        }
        __builtin_amdgcn_s_sleep(1);
        e = ld_relaxed(p);
        if (tag_of(e) == tag) return bits_of(e);
    }
}

// One step of one lane's 32 rows: U = V & M; (s, carry-out mask) = V + U +
// carry-in mask; V = s | (V & ~M)  (v_bitop3 table 0xf4 over (s, V, M)).
// Three VALU issues; the masks are SGPR pairs (bit l = lane l).
__device__ __forceinline__ void step(uint32_t &V, uint32_t M, unsigned long long cin, unsigned long long &cout)
{
    uint32_t u, s, vn;
    asm("v_and_b32 %[u], %[v], %[m]\n\t"
        "v_addc_co_u32_e64 %[s], %[co], %[v], %[u], %[ci]\n\t"
        "v_bitop3_b32 %[vn], %[s], %[v], %[m] bitop3:0xf4"
        : [u] "=&v"(u), [s] "=&v"(s), [vn] "=v"(vn), [co] "=&s"(cout)
        : [v] "v"(V), [m] "v"(M), [ci] "s"(cin));
    V = vn;
}

// Column symbol j as a table row: [0, sigma) as is, -1 (the zero row) for
// padding (j outside the string); a symbol outside [0, sigma) sets `bad`.
__device__ __forceinline__ int32_t code_of(int32_t s, bool inside, int sigma, bool &bad)
{
    if (!inside) return -1;
    if ((uint32_t)s >= (uint32_t)sigma) {
        bad = true;
        return -1;
    }
    return s;
}

__device__ bool band_llcs(const LlcsArgs &A, const LlcsDesc &D, int g, int lane, lds_u32 *stage, lds_u32 *tab)
{
    const int sigma = A.sigma;
    const uint32_t tab_addr = (uint32_t)(uintptr_t)tab;

    // match masks of this lane's 32 rows (each lane owns column `lane`)
    for (int c = -1; c < sigma; ++c) tab[c * kWave + lane] = 0u;
    const int64_t r0 = ((int64_t)g * kWave + lane) * kWordRows;
    bool bad = false;
    for (int i = 0; i < kWordRows; ++i) {
        const int64_t r = r0 + i;
        if (r < D.nrows) {
            const int32_t s = D.rows[r];
            if ((uint32_t)s < (uint32_t)sigma)
                tab[s * kWave + lane] |= 1u << i;
            else
                bad = true;
        }
    }

    const int32_t ncols = D.ncols, nchunks = D.nchunks;
    gcword *rin = (gcword *)(D.ring + (size_t)(g & 1) * nchunks);
    gword *rout = (gword *)(D.ring + (size_t)((g + 1) & 1) * nchunks);
    const uint32_t tag_in = (uint32_t)g;
    const unsigned long long tag_out = (unsigned long long)(g + 1) << 32;
    const bool has_in = g > 0, has_out = g + 1 < D.nbands;

    // column symbols of chunk 0 (positions p = lane, lane + 64 of 95)
    auto load_sym = [&](int64_t j) -> int32_t { return (j >= 0 && j < ncols) ? D.cols[j] : 0; };
    int32_t s0 = load_sym((int64_t)lane - (kWave - 1));
    int32_t s1 = lane < kStage - kWave ? load_sym((int64_t)lane + 1) : 0;

    // carry words of the band below: chunk q needs words q+1 and q+2
    bool ok = true;
    uint32_t w_hi = 0;
    unsigned long long e_next = 0;
    if (has_in) {
        if (nchunks > 1) w_hi = wait_word(A, rin + 1, ld_relaxed(rin + 1), tag_in, ok);
        if (!ok) return false;
        if (nchunks > 2) e_next = ld_relaxed(rin + 2);
    }

    uint32_t V = 0xffffffffu;
    unsigned long long cout = 0;
    for (int q = 0; q < nchunks; ++q) {
        // stage this chunk's symbols as table addresses
        {
            const int64_t j0 = (int64_t)q * kChunk - (kWave - 1);
            const int64_t ja = j0 + lane, jb = j0 + kWave + lane;
            // table address of row c, lane l, minus 4t (t = p + l - 63): read
            // back with offset 4t by lane l at step t (p = t - l + 63)
            const int32_t ca = code_of(s0, ja >= 0 && ja < ncols, sigma, bad);
            stage[lane] = (uint32_t)((int32_t)tab_addr + ca * (kWave * 4) + (kWave - 1) * 4 - 4 * lane);
            if (lane < kStage - kWave) {
                const int32_t cb = code_of(s1, jb >= 0 && jb < ncols, sigma, bad);
                stage[kWave + lane] = (uint32_t)((int32_t)tab_addr + cb * (kWave * 4) + (kWave - 1) * 4 - 4 * (kWave + lane));
            }
            s0 = load_sym(ja + kChunk);
            s1 = lane < kStage - kWave ? load_sym(jb + kChunk) : 0;
        }
        unsigned long long win = 0;
        if (has_in) {
            uint32_t w_lo = 0;
            if (q + 2 < nchunks) {
                w_lo = wait_word(A, rin + q + 2, e_next, tag_in, ok);
                if (!ok) return false;
                if (q + 3 < nchunks) e_next = ld_relaxed(rin + q + 3);
            }
            win = ((unsigned long long)w_hi << 32) | w_lo;
            w_hi = w_lo;
        }
        // the chunk's 32 masks: staged addresses, then the table reads
        const lds_u32 *sp = stage + (kWave - 1) - lane;
        uint32_t M[kChunk];
#pragma unroll
        for (int t = 0; t < kChunk; ++t) M[t] = sp[t];
#pragma unroll
        for (int t = 0; t < kChunk; ++t) M[t] = *((const lds_u32 *)(uintptr_t)M[t] + t);
        uint32_t wout = 0;
#pragma unroll
        for (int t = 0; t < kChunk; ++t) {
            const unsigned long long cin = (cout << 1) | ((win >> (32 - t)) & 1ull);
            step(V, M[t], cin, cout);
            wout = (wout << 1) | (uint32_t)(cout >> 63);
        }
        if (has_out && lane == 0) st_relaxed(rout + q, tag_out | wout);
    }

    if (__any(bad) && lane == 0) *D.err = 1;
    const int64_t live = D.nrows - r0;
    const uint32_t valid = live >= kWordRows ? 0xffffffffu : live <= 0 ? 0u : ((1u << live) - 1u);
    int zeros = __popc(~V & valid);
    for (int off = 32; off > 0; off >>= 1) zeros += __shfl_xor(zeros, off);
    if (lane == 0) atomicAdd(D.out, zeros);
    return true;
}

__global__ __launch_bounds__(kWave) void llcs_chain_kernel(LlcsArgs A)
{
    extern __shared__ uint32_t lds_dyn[];
    lds_u32 *stage = (lds_u32 *)lds_dyn;
    lds_u32 *tab = stage + kStageWords + kZeroRowWords;
    const int lane = threadIdx.x;
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = atomicAdd(A.ticket, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int)t >= A.ntasks) break;
        if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) break;
        const int2 task = A.tasks[t];
        const LlcsDesc D = A.descs[task.x];
        if (!__builtin_amdgcn_readfirstlane((int)band_llcs(A, D, task.y, lane, stage, tab))) break;
    }
}

// grid (x: ring words, y: pair); block (0,0) clears the ticket and abort words
__global__ void llcs_prep_kernel(const LlcsDesc *descs, unsigned *ticket, unsigned *abort_flag)
{
    const LlcsDesc D = descs[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * (int64_t)D.nchunks) D.ring[i] = (unsigned long long)kNoTag << 32;
    if (i == 0) {
        *D.out = 0;
        *D.err = 0;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        *ticket = 0u;
        *abort_flag = 0u;
    }
}

__global__ void llcs_finish_kernel(const LlcsDesc *descs, int npairs)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const LlcsDesc D = descs[p];
    if (*D.err) *D.out = -1;
}

}  // namespace lcsl

int lcsl_launch(const lcsl::LlcsArgs &args, const lcsl::LlcsDesc *d_descs, int npairs, int max_ring_words,
                int grid_blocks, hipStream_t stream)
{
    using namespace lcsl;
    if (npairs <= 0) return 0;
    if (args.sigma < 1 || args.sigma > kMaxSigma) return -22;
    const unsigned px = (unsigned)((max_ring_words + 255) / 256);
    hipLaunchKernelGGL(llcs_prep_kernel, dim3(px > 0 ? px : 1, npairs), dim3(256), 0, stream, d_descs, args.ticket,
                       args.abort_flag);
    if (args.ntasks > 0 && grid_blocks > 0)
        hipLaunchKernelGGL(llcs_chain_kernel, dim3(grid_blocks), dim3(kWave), llcs_lds_bytes(args.sigma), stream,
                           args);
    hipLaunchKernelGGL(llcs_finish_kernel, dim3((npairs + 255) / 256), dim3(256), 0, stream, d_descs, npairs);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
