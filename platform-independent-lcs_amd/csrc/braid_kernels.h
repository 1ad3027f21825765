// Launch interface of the steady-ant kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Largest padded core (2^L): cores up to 2^11 run in one workgroup's LDS,
// larger ones over the whole GPU (their grid of ant-walk boxes is sized
// for nodes of at most 2^22 points).
constexpr int kMaxCoreL = 22;
constexpr int kMaxCoreK = 1 << kMaxCoreL;

struct BraidCoreDesc {
    const int32_t *P;   // k: row -> shared column
    const int32_t *Q;   // k: shared row -> column
    int32_t *R;         // k: row -> column of the sticky product
    int32_t k, L;       // padded size 2^L >= k
    int32_t *scratch;   // lcsb_core_scratch_words(k) words (the big-core workspace; none for k <= 2^11)
};

// Core descriptors passed by value in the kernel arguments, one per block.
constexpr int kCoreBatch = 64;
struct CoreBatch {
    BraidCoreDesc d[kCoreBatch];
};

struct AssembleArgs {
    const int32_t *p, *q;
    int32_t *out;
    const int32_t *map_row, *map_col, *core;
    int32_t p_size, q_size, k, row_glue;
};

size_t lcsb_core_scratch_words(int32_t k);
// One-time kernel attribute setup (called by the launches; call it before
// capturing a launch sequence in a graph).
int lcsb_init(void);
int lcsb_core_L(int32_t k);
// One workgroup per descriptor (host array, copied into the launch arguments).
int lcsb_launch_core(const BraidCoreDesc *descs, int count, hipStream_t stream);
size_t lcsb_multiply_ws_words(int32_t p_size, int32_t q_size, int32_t k);
// staggered_multiply<row_glue>(p, q, k) on device pointers.
int lcsb_multiply_device(const int32_t *p, int32_t p_size, const int32_t *q, int32_t q_size, int32_t k,
                         int32_t row_glue, int32_t *out, int32_t *ws, hipStream_t stream);
// Several independent products; their cores share one launch.  ws of each
// holds lcsb_multiply_ws_words(p_size, q_size, k) words.
struct MulSpec {
    const int32_t *p;
    int32_t p_size;
    const int32_t *q;
    int32_t q_size, k, row_glue;
    int32_t *out, *ws;
};
int lcsb_multiply_many_device(const MulSpec *specs, int count, hipStream_t stream);
