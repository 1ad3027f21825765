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
//    above as one tagged 64-bit word (tag << 32 | bits) that the comb wave
//    stores with an agent-scope relaxed atomic; the fetch wave of the band
//    above polls for its tag, bounded in time, with the abort flag of the
//    comb kernels.  A carry-out bit is
//    collected one step late, which aligns producer word c with symbols
//    32(c-2) .. 32(c-2)+31: consumer chunk q reads word q+2.
//  * Bands are handed out by an atomic ticket in order, so every wait is on a
//    band that a resident wave already holds.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "llcs_kernels.h"

namespace lcsl {

#define LCSL_GLOBAL __attribute__((address_space(1)))
typedef LCSL_GLOBAL unsigned long long gword;
typedef LCSL_GLOBAL const unsigned long long gcword;
typedef LCSL_GLOBAL const int32_t gsym;     // global, not flat: a flat load also counts against lgkmcnt,
                                            // so every LDS wait would wait for it too
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) volatile uint32_t lds_vu32;

constexpr int kStage = kChunk + kWave - 1;   // 95 symbols staged per chunk
constexpr int kStageWords = 96;              // one staging buffer (two are used)
constexpr int kZeroRowWords = kWave;         // row -1 of the table, all zero
constexpr unsigned kNoTag = 0xffffffffu;
constexpr int kSymRing = kLdsSymRing;        // LDS rings (powers of two): column symbols,
constexpr int kInRing = kLdsInRing;          // carry words of the band below
constexpr int kSymBatch = 1024;              // symbols per fetch round
// LDS control words of a workgroup
constexpr int kTask = 0, kGo = 1, kStop = 2, kInAvail = 3, kInUsed = 4, kSymAvail = 5, kSymUsed = 6;

// The workgroup's LDS: staging buffers, the zero row and the match-mask
// table (comb wave), the two rings and the control words (shared).
struct Lds {
    lds_u32 *stage, *tab, *sym, *in;
    lds_vu32 *ctl;
};

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

// Waits (bounded in time) until control word `w` of the workgroup exceeds
// `need`; `cache` keeps the last value seen.  False when the band must stop
// (the other wave gave up, or the time ran out).
__device__ __forceinline__ bool wait_ctl(const LlcsArgs &A, lds_vu32 *ctl, int w, uint32_t need, uint32_t &cache)
{
    if (cache > need) return true;
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        cache = __builtin_amdgcn_readfirstlane(ctl[w]);
        if (cache > need) return true;
        if (__builtin_amdgcn_readfirstlane(ctl[kStop])) return false;
        if (__builtin_amdgcn_s_memrealtime() - t0 > A.timeout_ticks) {
            __hip_atomic_store(A.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(0);
    }
}

// Sixteen steps of one lane's 32 rows.  Per step: the band's carry-in bit
// for lane 0 (bit B of w) goes to SCC; VCC, the previous step's carry-out
// mask, doubles into the carry-in mask (lane l-1's carry enters lane l) with
// SCC as bit 0, and its old bit 63 (lane 63's carry-out of the previous
// step) doubles into `o`; then U = V & M, (s, VCC) = V + U + VCC and
// V = s | (V & ~M) (v_bitop3 table 0xf4 over (s, V, M)).  Four SALU and
// three VALU issues.  VCC is written by VALU and read by SALU and the
// reverse, which needs no wait states on gfx950; the carry state enters and
// leaves in `c` so that no compiled code between blocks can disturb VCC.
#define LCSL_STEP(B, K)                                   \
    "s_bitcmp1_b32 %[w], " #B "\n\t"                    \
    "s_addc_u32 vcc_lo, vcc_lo, vcc_lo\n\t"              \
    "s_addc_u32 vcc_hi, vcc_hi, vcc_hi\n\t"              \
    "s_addc_u32 %[o], %[o], %[o]\n\t"                    \
    "v_and_b32 %[u], %[v], %[m" #K "]\n\t"               \
    "v_addc_co_u32_e32 %[s], vcc, %[v], %[u], vcc\n\t"   \
    "v_bitop3_b32 %[v], %[s], %[v], %[m" #K "] bitop3:0xf4\n\t"

#define LCSL_OPERANDS(M, OFF)                                                                          \
    : [v] "+v"(V), [o] "+s"(wout), [c] "+s"(carry), [u] "=&v"(u), [s] "=&v"(sum)                       \
    : [w] "s"(win), [m0] "v"(M[OFF + 0]), [m1] "v"(M[OFF + 1]), [m2] "v"(M[OFF + 2]),                  \
      [m3] "v"(M[OFF + 3]), [m4] "v"(M[OFF + 4]), [m5] "v"(M[OFF + 5]), [m6] "v"(M[OFF + 6]),           \
      [m7] "v"(M[OFF + 7])                                                                              \
    : "vcc", "scc"

// (readfirstlane returns int: each half goes through uint32_t so the low
// half is not sign-extended into the high one)
__device__ __forceinline__ uint64_t uniform64(uint64_t x)
{
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)x);
    return ((uint64_t)hi << 32) | lo;
}

// Eight steps (quarter Q of a chunk): carry-in bits come from w (bit 31 - t
// at step t of the chunk), carry-out bits of the previous step are
// collected into wout.  Volatile, so the LDS reads the chunk loop places
// between quarters stay there: at most a quarter's reads are outstanding at
// each block, well under lgkmcnt's limit of 15 (a wait for an older read
// then never waits for the newer ones).
template <int Q>
__device__ __forceinline__ void quarter_steps(uint32_t &V, const uint32_t (&M)[kChunk], uint32_t win, uint32_t &wout,
                                              unsigned long long &carry)
{
    uint32_t u, sum;
    if constexpr (Q == 0)
        asm volatile("s_mov_b64 vcc, %[c]\n\t"
                     LCSL_STEP(31, 0) LCSL_STEP(30, 1) LCSL_STEP(29, 2) LCSL_STEP(28, 3)
                     LCSL_STEP(27, 4) LCSL_STEP(26, 5) LCSL_STEP(25, 6) LCSL_STEP(24, 7)
                     "s_mov_b64 %[c], vcc"
                     LCSL_OPERANDS(M, 0));
    else if constexpr (Q == 1)
        asm volatile("s_mov_b64 vcc, %[c]\n\t"
                     LCSL_STEP(23, 0) LCSL_STEP(22, 1) LCSL_STEP(21, 2) LCSL_STEP(20, 3)
                     LCSL_STEP(19, 4) LCSL_STEP(18, 5) LCSL_STEP(17, 6) LCSL_STEP(16, 7)
                     "s_mov_b64 %[c], vcc"
                     LCSL_OPERANDS(M, 8));
    else if constexpr (Q == 2)
        asm volatile("s_mov_b64 vcc, %[c]\n\t"
                     LCSL_STEP(15, 0) LCSL_STEP(14, 1) LCSL_STEP(13, 2) LCSL_STEP(12, 3)
                     LCSL_STEP(11, 4) LCSL_STEP(10, 5) LCSL_STEP(9, 6) LCSL_STEP(8, 7)
                     "s_mov_b64 %[c], vcc"
                     LCSL_OPERANDS(M, 16));
    else
        asm volatile("s_mov_b64 vcc, %[c]\n\t"
                     LCSL_STEP(7, 0) LCSL_STEP(6, 1) LCSL_STEP(5, 2) LCSL_STEP(4, 3)
                     LCSL_STEP(3, 4) LCSL_STEP(2, 5) LCSL_STEP(1, 6) LCSL_STEP(0, 7)
                     "s_mov_b64 %[c], vcc"
                     LCSL_OPERANDS(M, 24));
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

// Comb wave of a band.  Column symbols come from the LDS symbol ring and the
// band below's carry words from the LDS in-ring, both filled by the fetch
// wave; its own carry words go straight to the global ring, one store by
// lane 0 per chunk.  The loop issues no global load, so nothing ever waits
// for those stores (on gfx950 vmcnt counts loads and stores alike); staging
// them through an LDS out-ring and the fetch wave instead cost 3 ms at
// m = n = 2^20, in the chunk's own LDS traffic and in the longer hand-off.
__device__ bool band_compute(const LlcsArgs &A, const LlcsDesc &D, int g, int lane, const Lds &L)
{
    const int sigma = A.sigma;
    lds_u32 *tab = L.tab;
    const uint32_t tab_addr = (uint32_t)(uintptr_t)tab;
    lds_vu32 *ctl = L.ctl;

    // match masks of this lane's 32 rows (each lane owns column `lane`)
    for (int c = -1; c < sigma; ++c) tab[c * kWave + lane] = 0u;
    const int64_t r0 = ((int64_t)g * kWave + lane) * kWordRows;
    bool bad = false;
    for (int i = 0; i < kWordRows; ++i) {
        const int64_t r = r0 + i;
        if (r < D.nrows) {
            const int32_t s = ((gsym *)D.rows)[r];
            if ((uint32_t)s < (uint32_t)sigma)
                tab[s * kWave + lane] |= 1u << i;
            else
                bad = true;
        }
    }

    const int32_t ncols = D.ncols, nchunks = D.nchunks;
    const bool has_in = g > 0, has_out = g + 1 < D.nbands;
    gword *rout = (gword *)(D.ring + (size_t)((g + 1) & 1) * D.nchunks);
    const unsigned long long tag_out = (unsigned long long)(g + 1) << 32;
    uint32_t in_avail = 0, sym_avail = 0;
    bool ok = true;

    // Chunk c covers steps [32c, 32c+32); lane l takes column symbol k - l at
    // step k, so the chunk needs symbols 32c-63 .. 32c+31: staged position
    // p = lane (symbol 32c-63+lane) and p = 64+lane (symbol 32c+1+lane).
    auto raw_syms = [&](int c, int32_t &sa, int32_t &sb) -> bool {
        const int64_t hi = min((int64_t)c * kChunk + kChunk, (int64_t)ncols);
        if (hi > 0 && !wait_ctl(A, ctl, kSymAvail, (uint32_t)(hi - 1), sym_avail)) return false;
        const int64_t ja = (int64_t)c * kChunk - (kWave - 1) + lane, jb = ja + kWave;
        // unconditional reads (every index is a ring slot); stage_chunk
        // ignores the values of positions outside the string
        sa = (int32_t)L.sym[ja & (kSymRing - 1)];
        sb = (int32_t)L.sym[jb & (kSymRing - 1)];
        return true;
    };
    auto stage_chunk = [&](int c, lds_u32 *buf, int32_t sa, int32_t sb) {
        const int64_t ja = (int64_t)c * kChunk - (kWave - 1) + lane, jb = ja + kWave;
        // table address of row code, lane l, minus 4t for t = p + l - 63:
        // lane l reads position t - l + 63 at step t and adds 4t
        const int32_t ca = code_of(sa, ja >= 0 && ja < ncols, sigma, bad);
        buf[lane] = (uint32_t)((int32_t)tab_addr + ca * (kWave * 4) + (kWave - 1) * 4 - 4 * lane);
        if (lane < kStage - kWave) {
            const int32_t cb = code_of(sb, jb >= 0 && jb < ncols, sigma, bad);
            buf[kWave + lane] = (uint32_t)((int32_t)tab_addr + cb * (kWave * 4) + (kWave - 1) * 4 - 4 * (kWave + lane));
        }
    };
    // staged addresses of a chunk (32 reads, 16 ds_read2), then its masks
    // in quarters of 8 dependent reads
    auto read_addrs = [&](const lds_u32 *buf, uint32_t (&M)[kChunk]) {
        const lds_u32 *sp = buf + (kWave - 1) - lane;
#pragma unroll
        for (int t = 0; t < kChunk; ++t) M[t] = sp[t];
    };
    auto read_quarter = [&](uint32_t (&M)[kChunk], int qtr) {
#pragma unroll
        for (int i = 0; i < kChunk / 4; ++i) {
            const int t = qtr * (kChunk / 4) + i;
            M[t] = *((const lds_u32 *)(uintptr_t)M[t] + t);
        }
    };
    auto read_masks = [&](const lds_u32 *buf, uint32_t (&M)[kChunk]) {
        read_addrs(buf, M);
        read_quarter(M, 0);
        read_quarter(M, 1);
        read_quarter(M, 2);
        read_quarter(M, 3);
    };

    uint32_t V = 0xffffffffu, wout = 0;
    unsigned long long carry = 0;
    uint32_t MA[kChunk], MB[kChunk];
    int32_t s0 = 0, s1 = 0;
    if (nchunks > 0) {
        if (!raw_syms(0, s0, s1)) return false;
        stage_chunk(0, L.stage, s0, s1);
        read_masks(L.stage, MA);
        if (nchunks > 1 && !raw_syms(1, s0, s1)) return false;
    }
    // chunk q: stage and read chunk q+1 (raw symbols of chunk q+2 ahead)
    // while chunk q combs; Mc and Mn swap roles, two chunks per iteration.
    // The band below's word q+2 holds the carry-out bits of symbols 32q ..
    // 32q+31 (it collects each bit one step late).
    // the band below's word for chunk q is read during chunk q-1 (before
    // that chunk's last quarter of mask reads, so waiting for it never waits
    // for them)
    // (the read itself is unconditional and masked only where the chunk
    // uses it, so that waiting for it never waits for later reads)
    auto ring_word = [&](int q, uint32_t &w, bool &use) -> bool {
        use = has_in && q + 2 < nchunks;
        if (use && !wait_ctl(A, ctl, kInAvail, (uint32_t)q + 2, in_avail)) return false;
        w = L.in[(q + 2) & (kInRing - 1)];
        if (use && (q & 15) == 15 && lane == 0) ctl[kInUsed] = (uint32_t)q + 3;
        return true;
    };
    uint32_t win_next = 0;
    bool use_next = false;
    if (!ring_word(0, win_next, use_next)) return false;
    auto do_chunk = [&](int q, uint32_t (&Mc)[kChunk], uint32_t (&Mn)[kChunk]) -> bool {
        uint32_t win = use_next ? win_next : 0u;
        // chunk q combs in four quarters; between them the next chunk's
        // masks are read a quarter at a time (addresses first).  After the
        // last chunk this stages chunk nchunks, all padding (zero-row
        // addresses), and reads nothing else.
        lds_u32 *nb = L.stage + ((q + 1) & 1) * kStageWords;
        stage_chunk(q + 1, nb, s0, s1);
        read_addrs(nb, Mn);
        // wave-uniform by construction; readfirstlane lets the compiler see
        // it across the early exits of the chunk loop
        win = __builtin_amdgcn_readfirstlane(win);
        wout = __builtin_amdgcn_readfirstlane(wout);
        carry = uniform64(carry);
        quarter_steps<0>(V, Mc, win, wout, carry);
        read_quarter(Mn, 0);
        quarter_steps<1>(V, Mc, win, wout, carry);
        read_quarter(Mn, 1);
        quarter_steps<2>(V, Mc, win, wout, carry);
        read_quarter(Mn, 2);
        if (!raw_syms(q + 2, s0, s1)) return false;
        if ((q & 15) == 15 && lane == 0) ctl[kSymUsed] = (uint32_t)max(0, (q + 3) * kChunk - (kWave - 1));
        if (!ring_word(q + 1, win_next, use_next)) return false;
        quarter_steps<3>(V, Mc, win, wout, carry);
        read_quarter(Mn, 3);
        if (has_out && lane == 0) st_relaxed(rout + q, tag_out | wout);
        return true;
    };
    int q = 0;
    for (; q + 1 < nchunks && ok; q += 2) {
        if (!do_chunk(q, MA, MB)) return false;
        if (!do_chunk(q + 1, MB, MA)) return false;
    }
    if (q < nchunks && !do_chunk(q, MA, MB)) return false;

    if (__any(bad) && lane == 0) *D.err = 1;
    const int64_t live = D.nrows - r0;
    const uint32_t valid = live >= kWordRows ? 0xffffffffu : live <= 0 ? 0u : ((1u << live) - 1u);
    int zeros = __popc(~V & valid);
    for (int off = 32; off > 0; off >>= 1) zeros += __shfl_xor(zeros, off);
    if (lane == 0) atomicAdd(D.out, zeros);
    return true;
}

// Fetch wave of a band: the band's global loads after its table is built.
// Each round it
//  * loads the next column symbols (up to kSymBatch, 16 per lane) into the
//    LDS symbol ring, as far as the comb wave's progress frees slots;
//  * loads the band below's next 64 carry words from the global ring and
//    copies the leading run that carries this band's tag to the LDS in-ring
//    (one load brings every word published so far).
// Both loads are issued before either result is awaited, so a round costs
// about one memory round trip.  Every wait is bounded in time.
__device__ void band_fetch(const LlcsArgs &A, const LlcsDesc &D, int g, int lane, const Lds &L)
{
    lds_vu32 *ctl = L.ctl;
    const int32_t ncols = D.ncols, nchunks = D.nchunks;
    const bool need_in = g > 0 && nchunks > 2;
    gcword *rin = (gcword *)(D.ring + (size_t)(g & 1) * nchunks);
    const uint32_t tag_in = (uint32_t)g;
    gsym *cols = (gsym *)D.cols;
    int32_t next_s = 0, next_w = 2;
    unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        const bool done_s = next_s >= ncols, done_w = !need_in || next_w >= nchunks;
        if (done_s && done_w) return;
        if (__builtin_amdgcn_readfirstlane(ctl[kStop])) return;
        bool progress = false;

        // issue: symbols, carry words in
        int32_t scnt = 0;
        int32_t sv[kSymBatch / kWave];
        if (!done_s) {
            const int32_t sym_used = (int32_t)__builtin_amdgcn_readfirstlane(ctl[kSymUsed]);
            scnt = min(min(sym_used + kSymRing - next_s, kSymBatch), ncols - next_s);
#pragma unroll
            for (int k = 0; k < kSymBatch / kWave; ++k) {
                const int32_t i = lane + k * kWave;
                sv[k] = i < scnt ? cols[next_s + i] : 0;
            }
        }
        int32_t want = 0;
        unsigned long long e = 0;
        if (!done_w) {
            const int32_t used = (int32_t)__builtin_amdgcn_readfirstlane(ctl[kInUsed]);
            want = min(min(used + kInRing - next_w, kWave), nchunks - next_w);
            if (want > 0 && lane < want) e = ld_relaxed(rin + next_w + lane);
        }
        // land the loads
        if (scnt > 0) {
#pragma unroll
            for (int k = 0; k < kSymBatch / kWave; ++k) {
                const int32_t i = lane + k * kWave;
                if (i < scnt) L.sym[(next_s + i) & (kSymRing - 1)] = (uint32_t)sv[k];
            }
            next_s += scnt;
            if (lane == 0) ctl[kSymAvail] = (uint32_t)next_s;   // after the symbols
            progress = true;
        }
        if (want > 0) {
            const unsigned long long fresh = __ballot(lane < want && (uint32_t)(e >> 32) == tag_in);
            const int k = ~fresh ? (int)__builtin_ctzll(~fresh) : kWave;   // leading run of fresh words
            if (k > 0) {
                if (lane < k) L.in[(next_w + lane) & (kInRing - 1)] = (uint32_t)e;
                next_w += k;
                if (lane == 0) ctl[kInAvail] = (uint32_t)next_w;   // after the words
                progress = true;
            }
        }
        if (progress) {
            t0 = __builtin_amdgcn_s_memrealtime();
        } else {
            if (__builtin_amdgcn_s_memrealtime() - t0 > A.timeout_ticks) {
                __hip_atomic_store(A.abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (lane == 0) ctl[kStop] = 1u;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// One workgroup of two waves per band: wave 0 combs, wave 1 fetches.  Thread
// 0 takes the ticket and publishes it (and whether to go on) in LDS; the
// barriers keep both waves on the same band.
__global__ __launch_bounds__(2 * kWave) void llcs_chain_kernel(LlcsArgs A)
{
    extern __shared__ uint32_t lds_dyn[];
    Lds L;
    L.stage = (lds_u32 *)lds_dyn;
    L.tab = L.stage + 2 * kStageWords + kZeroRowWords;
    L.sym = L.tab + A.sigma * kWave;
    L.in = L.sym + kSymRing;
    L.ctl = (lds_vu32 *)(L.in + kInRing);
    lds_vu32 *ctl = L.ctl;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    for (;;) {
        if (threadIdx.x == 0) {
            const unsigned t = atomicAdd(A.ticket, 1u);
            ctl[kTask] = t;
            ctl[kGo] = ((int)t < A.ntasks && !ld_flag(A.abort_flag)) ? 1u : 0u;
            ctl[kStop] = 0u;
            ctl[kInAvail] = 2u;
            ctl[kInUsed] = 2u;
            ctl[kSymAvail] = 0u;
            ctl[kSymUsed] = 0u;
        }
        __syncthreads();
        if (!__builtin_amdgcn_readfirstlane(ctl[kGo])) break;
        const unsigned t = __builtin_amdgcn_readfirstlane(ctl[kTask]);
        const int2 task = A.tasks[t];
        const LlcsDesc D = A.descs[task.x];
        if (wave == 0) {
            const bool traced = A.trace && (int)t < A.trace_cap;
            if (traced && lane == 0) A.trace[2 * (size_t)t] = __builtin_amdgcn_s_memrealtime();
            if (!band_compute(A, D, task.y, lane, L) && lane == 0) ctl[kStop] = 1u;
            if (traced && lane == 0) A.trace[2 * (size_t)t + 1] = __builtin_amdgcn_s_memrealtime();
        } else {
            band_fetch(A, D, task.y, lane, L);
        }
        __syncthreads();
    }
}

// grid (x: ring words, y: pair); block (0,0) clears the ticket and abort
// words, after folding the previous launch's abort flag into `sticky`
__global__ void llcs_prep_kernel(const LlcsDesc *descs, unsigned *ticket, unsigned *abort_flag, unsigned *sticky)
{
    const LlcsDesc D = descs[blockIdx.y];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * (int64_t)D.nchunks) D.ring[i] = (unsigned long long)kNoTag << 32;
    if (i == 0) {
        *D.out = 0;
        *D.err = 0;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (sticky && *abort_flag) *sticky = 1u;
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
    const unsigned px = (unsigned)(((long long)max_ring_words + 255) / 256);
    hipLaunchKernelGGL(llcs_prep_kernel, dim3(px > 0 ? px : 1, npairs), dim3(256), 0, stream, d_descs, args.ticket,
                       args.abort_flag, args.sticky);
    if (args.ntasks > 0 && grid_blocks > 0)
        hipLaunchKernelGGL(llcs_chain_kernel, dim3(grid_blocks), dim3(2 * kWave), llcs_lds_bytes(args.sigma), stream,
                           args);
    hipLaunchKernelGGL(llcs_finish_kernel, dim3((npairs + 255) / 256), dim3(256), 0, stream, d_descs, npairs);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
