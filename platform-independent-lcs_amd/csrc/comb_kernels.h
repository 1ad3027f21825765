// Launch interface of the combing kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One single-pair band-chain launch.  The grid is a stack of bands of
// 64*tile_m rows; band g reads its top-edge strands from ring slot g&1 of E
// and writes its bottom edge to slot (g+1)&1 (the last band writes v_out).
struct ChainArgs {
    const int32_t *a;            // m symbols, original order (row r holds a[r])
    const int32_t *b;            // n symbols
    int32_t m, n, nbands, pad0;
    int32_t *h_out;              // m strands, reference order (index m-1-r)
    int32_t *v_out;              // n strands
    const int32_t *h_in;         // nullable: initial h strands (reference order), else i
    const int32_t *v_in;         // nullable: initial v strands, else m + j
    unsigned long long *E;       // 2*n ring words (tag << 32 | strand) + 64 trash words
    unsigned *ticket;            // ctrl[0]
    unsigned *abort_flag;        // ctrl[1]
    unsigned long long timeout_ticks;  // 100 MHz realtime ticks per wait
    unsigned long long *trace;   // nullable: [2g] start, [2g+1] end of band g (realtime ticks)
    int32_t trace_cap;           // bands the trace buffer holds
    int32_t pad1;
};

int lcsk_launch_chain(const ChainArgs &args, int tile_m, int grid_blocks, int block_threads, hipStream_t stream);
int lcsk_launch_identity(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n, hipStream_t stream);
