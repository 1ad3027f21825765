// Launch interface of the combing kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One comb problem: a (m symbols, row r holds a[r]) against b (n symbols),
// cut into bands of 64*tile_m rows.  Band g reads its top-edge strands from
// ring slot g&1 of E and writes its bottom edge to slot (g+1)&1; the prep
// kernel seeds slot 0, the extract kernel copies the last slot to v_out.
struct PairDesc {
    const int32_t *a;            // m symbols, original order
    const int32_t *b;            // n symbols
    const int32_t *h_in;         // nullable: initial h strands (reference order), else i
    const int32_t *v_in;         // nullable: initial v strands, else m + j
    int32_t *h_out;              // m strands, reference order (index m-1-r)
    int32_t *v_out;              // n strands
    unsigned long long *E;       // 2*n ring words (tag << 32 | strand) + 64 trash words
    int32_t m, n, nbands, tile_m;
    // Nullable: b as byte symbols (lcsk_launch_sym8), read instead of b by
    // the register chain when *sym8_bad is 0 (every symbol in [0, 255]).
    // sym8_col0 is the column of b[0] within those bytes (a strip's offset).
    const uint32_t *sym8;
    const int32_t *sym8_bad;
    int32_t sym8_col0, sym8_stride;
};

// Byte symbols of a string b of n symbols: four copies (stride words each,
// lcsk_sym8_stride(n)), copy j holding at byte x the symbol of column
// x + j - kSym8Pad as a byte (0 outside [0, n)).  A lane whose C-step
// window starts at byte x0 reads copy x0 % 4 from a word boundary.
constexpr int32_t kSym8Pad = 128;
inline int32_t lcsk_sym8_stride(int32_t n) { return (int32_t)((((int64_t)n + 2 * kSym8Pad + 15) / 16) * 4); }
// Writes the four copies of b into `out` (4 * stride words) and sets *bad
// (zeroed first) to nonzero when a symbol of b is outside [0, 255].
int lcsk_launch_sym8(const int32_t *b, int32_t n, uint32_t *out, int32_t stride, int32_t *bad, hipStream_t stream);

// Two pairs combed together by the 16-bit packed chain
// (LCS_VARIANT_PACKED16): pair q (0 = X, low halves; 1 = Y, high halves)
// reads a[q], b[q] (m_[q], n_[q] int32 symbols in [0, 0xfffd]) and writes
// h_out[q], v_out[q]; each has m_[q] + n_[q] <= 65535 and m_[q], n_[q] >= 1.
// m, n are the larger extents; the shorter pair is padded with rows and
// columns that never match and never swap.
struct CoupleDesc {
    const int32_t *a[2];
    const int32_t *b[2];
    int32_t *h_out[2];
    int32_t *v_out[2];
    int32_t m_[2], n_[2];
    unsigned long long *E;       // 2*n ring words (tag << 32 | Y strand << 16 | X strand) + 64 trash words
    int32_t m, n, nbands, tile_m;
    int32_t onehot;              // both pairs' symbols in [0, 15]: one-hot cells (1 << symbol per half)
};

// Words of ring a pair of n columns needs.
inline size_t lcsk_ring_words(int32_t n) { return 2 * (size_t)n + 64; }

// One persistent chain launch: waves take tickets in order; ticket t runs
// band tasks[t].y of pair tasks[t].x (group_w == 1), or bands
// group_w*tasks[t].y .. +group_w-1 in one workgroup (group_w > 1).  Tasks must list every band after the
// band above it (same pair), which makes every wait one on a resident wave.
struct ChainArgs {
    const PairDesc *pairs;       // the register chain and LDS groups
    const CoupleDesc *couples;   // the packed chain (tasks[t].x indexes couples)
    const int2 *tasks;
    int32_t ntasks;
    int32_t group_w;             // 1: one wave per task (register chain); W > 1: LDS groups of W bands
    unsigned *ticket;            // zeroed by the prep kernel
    unsigned *abort_flag;        // zeroed by the prep kernel
    unsigned long long timeout_ticks;  // 100 MHz realtime ticks per wait
    unsigned long long *trace;   // nullable: [2t] start, [2t+1] end of task t (realtime ticks)
    int32_t trace_cap;           // tasks the trace buffer holds
    int32_t poll_sleep;          // 1: a wave waiting for its inputs sleeps between polls (launches that
                                 // share SIMDs with other launches' waves)
    unsigned *sticky;            // nullable: the prep kernel ORs the abort flag into it before clearing it
};

// Seeds the rings of pairs[0..npairs) and clears ticket/abort; then runs the
// chain for tile height tile_m; then extracts v_out.  All on `stream`.
int lcsk_launch_chain(const ChainArgs &args, const PairDesc *d_pairs, int npairs, int max_n, int tile_m,
                      int grid_blocks, int block_threads, hipStream_t stream);
int lcsk_launch_identity(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n, hipStream_t stream);
// A string whose symbols the packed chain needs in [0, 0xfffd]: bit 0 of
// bad[item] is set when one is outside, bit 1 when one is above 15 (the
// one-hot cells then cannot take the pair).
struct SymbolRange {
    const int32_t *x;
    int32_t len, item;
};
int lcsk_launch_check16(const SymbolRange *d_ranges, int32_t nranges, int32_t *d_bad, int32_t nitems,
                        hipStream_t stream);
// The same sequence (seed, chain, extract) for couples of pairs combed by
// the 16-bit packed chain.
int lcsk_launch_chain16(const ChainArgs &args, const CoupleDesc *d_couples, int ncouples, int max_n, int tile_m,
                        int grid_blocks, int block_threads, hipStream_t stream);
// row_to_col of the permutation of each pair from its strands (m+n entries
// each, written at perm + perm_off[p]).
int lcsk_launch_from_strands(const PairDesc *d_pairs, int npairs, int max_len, int32_t *perm,
                             const int64_t *d_perm_off, hipStream_t stream);
// row_to_col (size m+n) of one pair's permutation from its device strands.
int lcsk_launch_r2c_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n, int32_t *r2c,
                                 hipStream_t stream);
// dst[0..count) = src[0..count) (a kernel, so it can sit in a captured graph).
int lcsk_launch_copy(int32_t *dst, const int32_t *src, long long count, hipStream_t stream);
// Strands from a permutation (row_to_col, size m+n): h_out[l] = row of col
// n+l, v_out[j] = row of col j.  scratch holds m+n words.
int lcsk_launch_strands_from_perm(const int32_t *r2c, int32_t m, int32_t n, int32_t *scratch, int32_t *h_out,
                                  int32_t *v_out, hipStream_t stream);
