// Launch interface of the score-only bit-parallel LCS kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcsl {

constexpr int kWave = 64;
constexpr int kWordRows = 32;                    // rows of the row string per lane
constexpr int kBandRows = kWave * kWordRows;     // 2048 rows per band (one wave)
constexpr int kChunk = 32;                       // steps per chunk = bits per ring word
constexpr int kMaxSigma = 200;              // table, staging and rings stay under 64 KiB of LDS

// One LCS-length problem: `rows` (nrows symbols) is cut into bands of 2048
// rows, one 32-bit word per lane; `cols` (ncols symbols) streams through the
// band.  Band g reads the carry words of band g-1 from ring slot g&1 (tag g)
// and writes its own to slot (g+1)&1 (tag g+1).
struct LlcsDesc {
    const int32_t *rows;
    const int32_t *cols;
    int32_t nrows, ncols, nbands, nchunks;
    unsigned long long *ring;    // 2 * nchunks words (tag << 32 | 32 carry bits)
    int32_t *out;                // LCS length (zeroed by the prep kernel); -1 on a symbol outside [0, sigma)
    int32_t *err;                // set to 1 on a symbol outside [0, sigma)
};

inline int32_t llcs_nbands(int32_t nrows) { return (int32_t)(((int64_t)nrows + kBandRows - 1) / kBandRows); }
// Chunks a band runs: lane 63 takes the last symbol at step ncols+62 and its
// carry-out is collected one step later, in word (ncols-1)/32 + 2.
inline int32_t llcs_nchunks(int32_t ncols) { return ncols > 0 ? (ncols - 1) / kChunk + 3 : 0; }

struct LlcsArgs {
    const LlcsDesc *descs;
    const int2 *tasks;           // (pair, band), every band after the band below it
    int32_t ntasks;
    int32_t sigma;               // symbols in [0, sigma), sigma <= kMaxSigma
    unsigned *ticket;            // zeroed by the prep kernel
    unsigned *abort_flag;        // zeroed by the prep kernel
    unsigned long long timeout_ticks;
    unsigned long long *trace;   // nullable: [2t] start, [2t+1] end of task t (realtime ticks)
    int32_t trace_cap;
    int32_t pad;
    unsigned *sticky;            // nullable: the prep kernel ORs the abort flag into it before clearing it
};

constexpr int kLdsSymRing = 2048, kLdsInRing = 128;
// LDS: two buffers of 96 staged addresses, one all-zero row (padding
// symbols), sigma rows, the symbol and carry-in rings, 16 control words
inline size_t llcs_lds_bytes(int sigma)
{
    return (size_t)(2 * 96 + kWave + sigma * kWave + kLdsSymRing + kLdsInRing + 16) * 4;
}

}  // namespace lcsl

// Clears outputs, error words, ticket and abort flag and invalidates every
// ring word; runs the band chain; marks pairs with a bad symbol (-1).
int lcsl_launch(const lcsl::LlcsArgs &args, const lcsl::LlcsDesc *d_descs, int npairs, int max_ring_words,
                int grid_blocks, hipStream_t stream);
