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
};

// Words of ring a pair of n columns needs.
inline size_t lcsk_ring_words(int32_t n) { return 2 * (size_t)n + 64; }

// One persistent chain launch: waves take tickets in order; ticket t runs
// band tasks[t].y of pair tasks[t].x (group_w == 1), or bands
// group_w*tasks[t].y .. +group_w-1 in one workgroup (group_w > 1).  Tasks must list every band after the
// band above it (same pair), which makes every wait one on a resident wave.
struct ChainArgs {
    const PairDesc *pairs;
    const int2 *tasks;
    int32_t ntasks;
    int32_t group_w;             // 1: one wave per task (register chain); W > 1: LDS groups of W bands
    unsigned *ticket;            // zeroed by the prep kernel
    unsigned *abort_flag;        // zeroed by the prep kernel
    unsigned long long timeout_ticks;  // 100 MHz realtime ticks per wait
    unsigned long long *trace;   // nullable: [2t] start, [2t+1] end of task t (realtime ticks)
    int32_t trace_cap;           // tasks the trace buffer holds
    int32_t pad1;
    unsigned *sticky;            // nullable: the prep kernel ORs the abort flag into it before clearing it
};

// Seeds the rings of pairs[0..npairs) and clears ticket/abort; then runs the
// chain for tile height tile_m; then extracts v_out.  All on `stream`.
int lcsk_launch_chain(const ChainArgs &args, const PairDesc *d_pairs, int npairs, int max_n, int tile_m,
                      int grid_blocks, int block_threads, hipStream_t stream);
int lcsk_launch_identity(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n, hipStream_t stream);
// row_to_col of the permutation of each pair from its strands (m+n entries
// each, written at perm + perm_off[p]).
int lcsk_launch_from_strands(const PairDesc *d_pairs, int npairs, int max_len, int32_t *perm,
                             const int64_t *d_perm_off, hipStream_t stream);
// row_to_col (size m+n) of one pair's permutation from its device strands.
int lcsk_launch_r2c_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n, int32_t *r2c,
                                 hipStream_t stream);
// Strands from a permutation (row_to_col, size m+n): h_out[l] = row of col
// n+l, v_out[j] = row of col j.  scratch holds m+n words.
int lcsk_launch_strands_from_perm(const int32_t *r2c, int32_t m, int32_t n, int32_t *scratch, int32_t *h_out,
                                  int32_t *v_out, hipStream_t stream);
