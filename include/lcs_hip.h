/*
 * lcs_hip.h -- C ABI of the MI355X (gfx950) semi-local LCS engine.
 *
 * The engine computes the seaweed (strand) permutation of two int32 strings
 * with hand-written HIP kernels and merges permutations with a GPU steady
 * ant.  Every entry point takes plain pointers and sizes, returns an int
 * status (0 = OK, negative errno-style codes below) and never throws.
 *
 * Semantics follow egor-bv/platform-independent-lcs (paths relative to its
 * root):
 *   - a[0..m), b[0..n) int32 symbols, compared with ==;
 *   - strands are int32 ids: h strands 0..m-1 enter the rows (row i holds
 *     a[m-1-i]), v strands m..m+n-1 enter the columns (src/lcs_reference.hpp
 *     :5-69);
 *   - outputs h_out[i] (i < m) and v_out[j] (j < n) are the strands leaving
 *     row i on the right and column j at the bottom; the permutation of size
 *     m+n maps h_out[l] -> n+l and v_out[j] -> j (src/permutation_matrix.hpp
 *     :88-115).
 *
 * Threading: a context is used by one host thread at a time, except the
 * server entry points, which are internally synchronised.
 */
#ifndef LCS_HIP_H
#define LCS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LCS_OK 0
#define LCS_ENOENT (-2)     /* unknown algorithm / id                          */
#define LCS_EIO (-5)        /* HIP runtime error                              */
#define LCS_EAGAIN (-11)    /* result not ready yet / admission deferred      */
#define LCS_ENOMEM (-12)    /* device memory or HBM budget exhausted          */
#define LCS_EBUSY (-16)     /* object in use                                  */
#define LCS_EINVAL (-22)    /* bad argument (null pointer, negative length)   */
#define LCS_ERANGE (-34)    /* size beyond what the entry point supports      */
#define LCS_ETIMEDOUT (-110)/* a device-side wait gave up (kernel aborted)    */
#define LCS_ECANCELED (-125)/* request cancelled                              */

typedef struct lcs_ctx lcs_ctx;

/* A combing schedule.  Schedules only change speed: every schedule gives
 * the same bits.  A pair whose schedule is NULL or has tile_m == 0 gets one
 * chosen by the context (lcs_ctx_choose_schedule, in this order): the
 * context's autotune table, then its learned cost model, then the
 * exhaustive autotune of the shape when enabled, then the built-in
 * defaults; the caller's other non-zero fields are kept.  A field still 0
 * after that takes the built-in default. */
typedef struct lcs_schedule {
    int32_t tile_m;          /* rows per lane: 1, 2, 4, 8, 16 or 32          */
    int32_t waves_per_simd;  /* resident waves per SIMD the launch aims at   */
    int32_t variant;         /* LCS_VARIANT_*                                */
    int32_t strips;          /* column strips merged by the steady ant       */
} lcs_schedule;

#define LCS_VARIANT_AUTO 0
#define LCS_VARIANT_CHAIN 1      /* wave64 band chain, registers only         */
/* The register chain reads b as bytes when every symbol of b lies in
 * [0, 255] (checked on the device per call): each lane then holds its own
 * column symbols and none move between lanes.  The bytes are staged (two
 * small launches) for pairs of 2^22 cells or more, and at any size with
 * LCS_VARIANT_CHAIN_BYTES; the results are the same either way. */
#define LCS_VARIANT_CHAIN_LDS 2  /* band chain with b staged through LDS      */
/* Batches only: pairs of a batch with this variant are combed two per wave
 * with 16-bit strands (pair X in the low, pair Y in the high half of every
 * register, 5 VALU per 2 cells against 4 per cell; 4 per 2 cells when
 * every symbol of both pairs lies in [0, 15]).  A pair qualifies when
 * m, n >= 1, m + n <= 65535 and every symbol lies in [0, 65533]; the others
 * (and an odd one out) take the register chain.  Same bits as every other
 * variant.  A batch pair left to the library (tile_m 0, LCS_VARIANT_AUTO,
 * after the context's table and cost model) that qualifies takes it too.
 * Elsewhere it means LCS_VARIANT_CHAIN. */
#define LCS_VARIANT_PACKED16 3
/* LCS_VARIANT_CHAIN with the byte symbols staged whatever the pair's size
 * (single pairs; in a batch it means LCS_VARIANT_CHAIN). */
#define LCS_VARIANT_CHAIN_BYTES 4

/* ---- context ------------------------------------------------------------
 * Replaces the registry-owned sycl::queue (src/algorithm_registry.cpp:6-41,
 * 167-173): owns the HIP stream, device workspaces and graphs.
 * hbm_budget bounds the device memory the context may hold (0 = 3/4 of the
 * free memory at creation). */
int lcs_ctx_create(int device, size_t hbm_budget, lcs_ctx **out);
int lcs_ctx_destroy(lcs_ctx *ctx);
/* Name of the device the context runs on (the CSV `device` column of
 * src/test_setup.hpp:331-347). */
int lcs_ctx_device_name(lcs_ctx *ctx, char *buf, size_t len);
/* Number of compute units, and the kernel-measured time in milliseconds of
 * the last lcs_comb / lcs_comb_device call (HIP events around the kernels). */
int lcs_ctx_info(lcs_ctx *ctx, int32_t *num_cus, double *last_kernel_ms);
/* Time in milliseconds of the band-chain launch alone (its seed, chain and
 * extract kernels, HIP events) within the last single-pair call; for a pair
 * cut into column strips this leaves out the strip permutations and the
 * steady-ant merges.  Waits for that launch; LCS_ENOENT before any call. */
int lcs_ctx_chain_ms(lcs_ctx *ctx, double *chain_ms);
/* Waits for every kernel the context has enqueued (its own stream and the
 * streams given to the asynchronous entry points) and reports whether any
 * launch since the previous lcs_ctx_status gave up a device-side wait:
 * LCS_ETIMEDOUT (the outputs of that launch are undefined), else LCS_OK.
 * The asynchronous entry points (lcs_comb_device, lcs_llcs_device) return
 * after enqueueing, so this is how their callers learn of an abort.  The
 * report is cleared by the call.  No counterpart in the reference, whose
 * calls are synchronous (src/algorithm_registry.cpp:43-75). */
int lcs_ctx_status(lcs_ctx *ctx);
/* Bound of every device-side wait, in ticks of the 100 MHz device clock
 * (default 500000000 = 5 s).  A wait that exceeds it makes its launch give
 * up: every wave exits and the call (or lcs_ctx_status) reports
 * LCS_ETIMEDOUT.  0 makes every wait that finds its data not yet there give
 * up at once (a way to exercise the error path). */
int lcs_ctx_set_timeout(lcs_ctx *ctx, uint64_t ticks);
/* Device bytes the context holds now (workspaces grow on demand and stay
 * until lcs_ctx_destroy) and its budget. */
int lcs_ctx_memory(lcs_ctx *ctx, size_t *held, size_t *budget);

/* Diagnostics (no counterpart in the reference): record the start and end
 * time (100 MHz device clock) of each band of the following chain launches
 * (lcs_comb* and lcs_llcs*).
 * max_bands > 0 arms tracing, 0 disarms it; with host_out != NULL copies
 * 2*max_bands words of the last launch's trace out instead. */
int lcs_debug_band_trace(lcs_ctx *ctx, int32_t max_bands, uint64_t *host_out);

/* ---- single pair --------------------------------------------------------
 * Replaces an LcsProcedure registered in reg2 (src/lcs_interface_internal.hpp
 * :91, src/algorithm_registry.cpp:149-155) such as Lcs_Semi_Tiled_MT
 * (src/lcs_tiled.hpp:95-137) together with its residual fixup
 * (src/lcs_grid.hpp:451-493): host pointers, synchronous.  h_out has m
 * entries, v_out n entries; llcs_out (nullable) receives #{ j : v_out[j] < m }.
 * sched may be NULL.  Sizes: m + n <= 2^31 - 1 and n <= 2^31 - 4097 (the
 * widest b one band chain combs), else LCS_ERANGE; the same limits hold for
 * every comb entry point below. */
int lcs_comb(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
             const lcs_schedule *sched, int32_t *h_out, int32_t *v_out, int32_t *llcs_out);

/* Replaces SemiLocalLcsImpl::operator() (src/algorithm_registry.cpp:43-75):
 * the procedure followed by PermutationMatrix::FromStrands
 * (src/permutation_matrix.hpp:88-115), which the reference runs on the host
 * inside its timed region (:70-71).  Host pointers, synchronous.  Combs on
 * the GPU, builds the permutation there and copies out row_to_col and
 * col_to_row (m+n entries each, the reference's layout: strand h[l] ends in
 * column n+l, v[j] in column j); llcs_out (nullable) as for lcs_comb. */
int lcs_comb_permutation(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                         const lcs_schedule *sched, int32_t *row_to_col, int32_t *col_to_row, int32_t *llcs_out);

/* Same computation on device pointers, enqueued on `stream` (a hipStream_t;
 * NULL = the context's stream).  Returns after enqueueing; the caller
 * synchronises the stream and calls lcs_ctx_status to learn whether the
 * launch completed.  Workspace (ring, descriptors, control words) comes
 * from the context and is shared by all its calls: a call on a stream other
 * than the previous call's first waits, on the device, for the previous
 * call's kernels, so calls through one context never overlap.  Concurrent
 * combs take one context per stream, or lcs_batch_comb. */
int lcs_comb_device(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                    const lcs_schedule *sched, int32_t *d_h_out, int32_t *d_v_out, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same comb as lcs_comb_device (one strip, the
 * band-chain kernels) without a context: its ring, descriptor, task list and
 * control words live in a static device array of about 32.5 MiB per GPU,
 * held until the process ends, so the call allocates nothing and copies
 * nothing from host memory.  Limits: n <= 2^21 columns and at most 65536
 * band tasks (LCS_ERANGE), strips > 1 is LCS_EINVAL.  Asynchronous on
 * `stream` (a hipStream_t; NULL = the null stream) of the current device;
 * d_status (nullable, device int32) receives LCS_OK or LCS_ETIMEDOUT once
 * the launch's kernels are done.  Every call on one GPU shares that GPU's
 * array: calls whose kernels overlap on one GPU race, so the caller orders
 * them (one stream, or events). */
int lcs_comb_device_noalloc(const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n, const lcs_schedule *sched,
                            int32_t *d_h_out, int32_t *d_v_out, int32_t *d_status, void *stream);

/* The residual fixup of the reference as a GPU operation: comb the
 * sub-rectangle rows [i0, i0+isize) x columns [j0, j0+jsize) of the grid in
 * place, starting from the strands currently held in h/v (host pointers,
 * lengths m and n; a_prepared is a reversed, as in the reference's
 * LcsProblemContext::a_prepared).  Follows Lcs_Semi_Fixup
 * (src/lcs_residual_fixup.hpp:74-122). */
int lcs_comb_fixup(lcs_ctx *ctx, const int32_t *a_prepared, int32_t m, const int32_t *b, int32_t n,
                   int32_t *h, int32_t *v, int32_t i0, int32_t isize, int32_t j0, int32_t jsize);

/* ---- permutations ------------------------------------------------------ */
/* PermutationMatrix::FromStrands (src/permutation_matrix.hpp:88-115): host. */
int lcs_permutation_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n,
                                 int32_t *row_to_col, int32_t *col_to_row);
/* PermutationMatrix::hash (src/permutation_matrix.hpp:192-203) with the
 * reference's two's-complement wrap made well defined.  Host. */
int64_t lcs_permutation_hash(const int32_t *col_to_row, int32_t size);

/* Semi-local scores read from a permutation (row_to_col, host pointers) on
 * the GPU.  The reference reads no scores from its permutations (its only
 * score routine, Lcs_Prefix_Reference in src/lcs_reference.hpp:71-132, is
 * an unregistered DP); these follow the seaweed read-out:
 *   lcs_permutation_dominance:  out[q] = #{ r >= qr[q] : row_to_col[r] < qc[q] }
 *   lcs_semilocal_string_substring: out[q] = LCS(a, b[qi[q]:qj[q])) for
 *     0 <= qi <= qj <= n, from P(a, b) of size m+n. */
int lcs_permutation_dominance(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size, const int32_t *qr,
                              const int32_t *qc, int32_t nq, int32_t *out);
int lcs_semilocal_string_substring(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n, const int32_t *qi,
                                   const int32_t *qj, int32_t nq, int32_t *out);

/* The four semi-local quadrants of P(a, b) (size m+n), query q = (q0, q1):
 *   LCS_SEMI_STRING_SUBSTRING  LCS(a, b[q0:q1))        0 <= q0 <= q1 <= n
 *   LCS_SEMI_SUBSTRING_STRING  LCS(a[q0:q1), b)        0 <= q0 <= q1 <= m
 *   LCS_SEMI_PREFIX_SUFFIX     LCS(a[:q0), b[q1:))     0 <= q0 <= m, 0 <= q1 <= n
 *   LCS_SEMI_SUFFIX_PREFIX     LCS(a[q0:), b[:q1))     0 <= q0 <= m, 0 <= q1 <= n
 *   LCS_SEMI_DOMINANCE         #{ r >= q0 : row_to_col[r] < q1 }   (m, n: m+n = size)
 * All are H(i, j) = (j - i) - #{ r >= m + i : P[r] < j } of the seaweed
 * matrix, i in [-m, n], j in [0, m+n], read at fixed offsets. */
#define LCS_SEMI_STRING_SUBSTRING 0
#define LCS_SEMI_SUBSTRING_STRING 1
#define LCS_SEMI_PREFIX_SUFFIX 2
#define LCS_SEMI_SUFFIX_PREFIX 3
#define LCS_SEMI_DOMINANCE 4

/* One-shot quadrant queries from a host row_to_col (host pointers,
 * synchronous).  Large query sets go through a dominance index built for
 * the call, small ones through a direct count.  LCS_EINVAL for a query
 * outside its quadrant's range.  No counterpart in the reference. */
int lcs_semilocal_query(lcs_ctx *ctx, const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind,
                        const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out);

/* A device-resident dominance index (wavelet matrix over row_to_col + 1):
 * each query costs O(log size) dependent loads instead of a scan of the
 * permutation.  row_to_col must hold values in [-1, size) (-1 = unset);
 * anything else is LCS_EINVAL.  The index holds ceil(log2(size+1)) levels of
 * 16 bytes per 64 rows, charged to the context's HBM budget; destroy it
 * before the context.  No counterpart in the reference (it reads no scores
 * from its permutations).
 *   _build: host row_to_col; _build_device: device row_to_col (both
 *     synchronous, the input may be freed on return).
 *   _query: host q0/q1/out, synchronous, validated (LCS_EINVAL).
 *   _query_device: device q0/q1/out, asynchronous on the context's stream;
 *     a query outside its quadrant's range yields -1. */
typedef struct lcs_dominance_index lcs_dominance_index;
int lcs_dominance_index_build(lcs_ctx *ctx, const int32_t *row_to_col, int32_t size, lcs_dominance_index **out);
int lcs_dominance_index_build_device(lcs_ctx *ctx, const int32_t *d_row_to_col, int32_t size,
                                     lcs_dominance_index **out);
int lcs_dominance_index_query(lcs_dominance_index *idx, int32_t m, int32_t n, int32_t kind, const int32_t *q0,
                              const int32_t *q1, int32_t nq, int32_t *out);
int lcs_dominance_index_query_device(lcs_dominance_index *idx, int32_t m, int32_t n, int32_t kind,
                                     const int32_t *d_q0, const int32_t *d_q1, int32_t nq, int32_t *d_out);
int64_t lcs_dominance_index_bytes(const lcs_dominance_index *idx);
void lcs_dominance_index_destroy(lcs_dominance_index *idx);

/* staggered_multiply<RowGlue>(p, q, k) (src/braid_multiplication.hpp:273-406)
 * on the GPU: the sticky-braid product of p (size p_size) and q (size
 * q_size) glued on k strands; row_glue = 0 glues p's last k columns to q's
 * first k rows (b-split), 1 glues p's first k columns to q's last k rows
 * (a-split).  Host pointers (row_to_col form); out has p_size+q_size-k
 * entries. */
int lcs_braid_multiply(lcs_ctx *ctx, const int32_t *p_row_to_col, int32_t p_size,
                       const int32_t *q_row_to_col, int32_t q_size, int32_t k, int32_t row_glue,
                       int32_t *out_row_to_col);

/* ---- score-only LCS length (bit-parallel) --------------------------------
 * The length of the LCS of a and b alone, by the word-parallel LCS
 * recurrence on the GPU (one bit per cell instead of one strand); the
 * counterpart of the reference's disabled binary bit-parallel LLCS
 * (Lcs_Prefix_Binary_Reference, src/lcs_reference.hpp:135-211), for any
 * alphabet of at most 200 symbols.  Symbols of a and b must lie in
 * [0, sigma), 1 <= sigma <= 200; a symbol outside gives LCS_ERANGE (host
 * entry points) or a length of -1 (device entry point).  Equal to
 * #{v_out[j] < m} of lcs_comb on the same pair.  The string combed as
 * columns (the longer one, as a rule) may hold at most 2^31 - 4097
 * symbols, else LCS_ERANGE.
 *   lcs_llcs: host pointers, synchronous.
 *   lcs_llcs_device: device pointers, asynchronous on the context's stream;
 *     *d_llcs is a device int32.
 *   lcs_llcs_batch: npairs pairs in one launch; a_all/b_all host arrays
 *     with per-pair offsets and lengths (as lcs_batch_comb); out[p]. */
int lcs_llcs(lcs_ctx *ctx, const int32_t *a, int32_t m, const int32_t *b, int32_t n, int32_t sigma, int32_t *llcs);
int lcs_llcs_device(lcs_ctx *ctx, const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n, int32_t sigma,
                    int32_t *d_llcs);
int lcs_llcs_batch(lcs_ctx *ctx, int32_t npairs, const int32_t *a_all, const int64_t *a_off, const int32_t *m,
                   const int32_t *b_all, const int64_t *b_off, const int32_t *n, int32_t sigma, int32_t *out);

/* ---- batched pairs ------------------------------------------------------
 * No counterpart in the reference (it combs one pair per call); the batch
 * form the reference's harness would call once per line of its run script
 * (src/main.cpp:46-64).  Pair p is a[a_off[p] .. a_off[p]+m[p]) against
 * b[b_off[p] .. b_off[p]+n[p]) inside two concatenated host arrays; outputs
 * go to h_out[a_off[p]..] and v_out[b_off[p]..].  scheds (nullable) gives a
 * schedule per pair.  use_graph != 0 captures the launch in a hipGraph. */
int lcs_batch_comb(lcs_ctx *ctx, int32_t npairs,
                   const int32_t *a_all, const int64_t *a_off, const int32_t *m,
                   const int32_t *b_all, const int64_t *b_off, const int32_t *n,
                   const lcs_schedule *scheds, int32_t use_graph,
                   int32_t *h_out, int32_t *v_out);
/* The same with the symbols and strands in device memory (d_*: device
 * pointers; offsets, lengths and schedules stay host arrays).  Synchronous;
 * lcs_ctx_info reports the launch's kernel time. */
int lcs_batch_comb_device(lcs_ctx *ctx, int32_t npairs,
                          const int32_t *d_a_all, const int64_t *a_off, const int32_t *m,
                          const int32_t *d_b_all, const int64_t *b_off, const int32_t *n,
                          const lcs_schedule *scheds, int32_t use_graph,
                          int32_t *d_h_out, int32_t *d_v_out);

/* ---- autotuner ------------------------------------------------------------
 * No counterpart in the reference (its variants were swept offline by
 * notebooks, tests/test_prepare.ipynb).  Times each candidate schedule on
 * synthetic pairs of the shape (A uniform over `alphabet`, B = A mutated at
 * `mutation_rate`), best of `reps` runs after a warm-up: mode 0 times one
 * pair through lcs_comb_device (cost = kernel ms); mode 1 prices a pair in
 * a batch that keeps the GPU busy: K identical-shape pairs (enough for two
 * bands of 64*32 rows per SIMD, K in [8, 4096]) through lcs_batch_comb,
 * columns cut to about 2^30 / (K*m) (at least 1024), cost = kernel ms per
 * pair scaled by (n + 64) / (columns + 64).
 * cost_ms has ncand entries. */
int lcs_autotune_measure(lcs_ctx *ctx, int32_t m, int32_t n, int32_t alphabet, float mutation_rate, uint64_t seed,
                         const lcs_schedule *candidates, int32_t ncand, int32_t mode, int32_t reps, float *cost_ms);

/* ---- schedule selection ---------------------------------------------------
 * No counterpart in the reference.  Two modes, each with its own table,
 * model and on-miss autotune: LCS_TUNE_SINGLE prices one pair alone
 * (lcs_comb, lcs_comb_permutation, lcs_comb_device), LCS_TUNE_BATCH a pair
 * inside a batch (lcs_batch_comb*).  For a pair (m, n) left to the library:
 *   1. table: the autotuned shape nearest to (m, n) within 1/8 octave in
 *      both lengths gives its measured best schedule;
 *   2. model: the attached cost model (lcs_mlp, created on this context)
 *      predicts a cost per candidate from (log2 m, log2 n, log2 alphabet,
 *      mutation rate) and the cheapest candidate is taken.  The alphabet
 *      is 1 + the largest symbol for host strings, 4 for device strings;
 *      the mutation rate is taken as 0.1 (the comb's cost does not depend
 *      on the symbols).  Decisions are memoised per (m, n, alphabet);
 *   3. autotune on a miss (when enabled): lcs_autotune_measure times the
 *      on-miss candidates on a synthetic pair of the shape (seconds) and
 *      the best joins the table;
 *   4. the built-in defaults.
 * Table rows usually come from the autotuner's recorded tables (the
 * Python package loads them); the model is trained on the same timings. */
#define LCS_TUNE_SINGLE 0
#define LCS_TUNE_BATCH 1
#define LCS_SCHED_FROM_DEFAULT 0
#define LCS_SCHED_FROM_TABLE 1
#define LCS_SCHED_FROM_MODEL 2
#define LCS_SCHED_FROM_AUTOTUNE 3
#define LCS_SCHED_FROM_CALLER 4
/* Adds (or replaces) the table row of shape (m, n); best->tile_m != 0. */
int lcs_ctx_add_tuned(lcs_ctx *ctx, int32_t mode, int32_t m, int32_t n, const lcs_schedule *best);
int lcs_ctx_clear_tuned(lcs_ctx *ctx, int32_t mode);
/* Attaches a cost model whose n_out outputs price candidates[0..ncand)
 * (n_out == ncand, every tile_m != 0); NULL detaches.  The model must have
 * been created on this context and stays owned by the caller;
 * lcs_mlp_destroy detaches it. */
struct lcs_mlp;
int lcs_ctx_set_cost_model(lcs_ctx *ctx, int32_t mode, struct lcs_mlp *model, const lcs_schedule *candidates,
                           int32_t ncand);
/* Enables the exhaustive autotune of shapes that the table and the model
 * leave open (ncand = 0 disables it); reps timed runs per candidate. */
int lcs_ctx_set_autotune_on_miss(lcs_ctx *ctx, int32_t mode, const lcs_schedule *candidates, int32_t ncand,
                                 int32_t reps);
/* The schedule the context would use for a pair of this shape, and which
 * step chose it (LCS_SCHED_FROM_*). */
int lcs_ctx_choose_schedule(lcs_ctx *ctx, int32_t mode, int32_t m, int32_t n, int32_t alphabet, float rate,
                            lcs_schedule *out, int32_t *source);

/* ---- request server ------------------------------------------------------
 * No counterpart in the reference.  Continuous batching of pair comparisons:
 * a worker thread admits queued requests, oldest first, into a device arena
 * of hbm_budget bytes (0 = half the free memory) as memory and one of
 * max_inflight (1..8) micro-batch slots free up; micro-batches run on their
 * own streams and overlap.  Inputs are copied at submit; results stay until
 * lcs_server_release.  Each result equals lcs_comb's for the same pair.
 * features (nullable, 4 floats: the cost model's inputs) let an attached
 * model choose the schedule when sched is NULL or has tile_m == 0. */
typedef struct lcs_server lcs_server;
int lcs_server_create(lcs_ctx *ctx, size_t hbm_budget, int32_t max_inflight, lcs_server **out);
int lcs_server_destroy(lcs_server *server);
int lcs_server_submit(lcs_server *server, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                      const lcs_schedule *sched, const float *features, int64_t *request_id);
/* *done = 1 and the strands copied out when finished (0 = still queued or
 * running); returns the request's status once done */
int lcs_server_poll(lcs_server *server, int64_t request_id, int32_t *h_out, int32_t *v_out, int32_t *done);
/* blocks up to timeout_ms (< 0: forever); LCS_EAGAIN on timeout */
int lcs_server_wait(lcs_server *server, int64_t request_id, int32_t *h_out, int32_t *v_out, int32_t timeout_ms);
int lcs_server_cancel(lcs_server *server, int64_t request_id);
/* forget a finished request (its result and times) */
int lcs_server_release(lcs_server *server, int64_t request_id);
/* microseconds from submit to admission and from submit to result */
int lcs_server_request_times(lcs_server *server, int64_t request_id, double *queued_us, double *total_us);
/* queue length, completed requests, launched micro-batches, most slots in
 * flight at once, free arena bytes; returns the worker's last error */
int lcs_server_stats(lcs_server *server, int64_t *pending, int64_t *completed, int64_t *batches,
                     int32_t *max_inflight, size_t *free_bytes);

/* ---- learned schedule cost model ---------------------------------------
 * No counterpart in the reference (its variant choice is a compile-time
 * name, src/algorithm_registry.cpp:127-142).  An fp32 MLP
 * (n_in -> hidden -> hidden -> n_out, ReLU) trained on the GPU by the
 * library's own kernels on the timings the autotuner records; it predicts a
 * cost per candidate schedule from a pair's features and the cheapest
 * prediction is chosen.  Schedules only change speed, never results.
 * X is n x n_in, T and M (mask of measured entries) are n x n_out,
 * row-major floats, host pointers. */
typedef struct lcs_mlp lcs_mlp;
int lcs_mlp_create(lcs_ctx *ctx, int32_t n_in, int32_t hidden, int32_t n_out, uint64_t seed, lcs_mlp **out);
int lcs_mlp_destroy(lcs_mlp *mlp);
int lcs_mlp_param_count(lcs_mlp *mlp, int32_t *count);
int lcs_mlp_get_params(lcs_mlp *mlp, float *out);
int lcs_mlp_set_params(lcs_mlp *mlp, const float *in);
int lcs_mlp_predict(lcs_mlp *mlp, const float *X, int32_t n, float *Y);
/* masked mean squared error and its gradient (param_count floats) */
int lcs_mlp_loss_grad(lcs_mlp *mlp, const float *X, const float *T, const float *M, int32_t n, float *loss,
                      float *grad);
/* `steps` full-batch Adam steps (lr), loss after the last step */
int lcs_mlp_train(lcs_mlp *mlp, const float *X, const float *T, const float *M, int32_t n, int32_t steps, float lr,
                  float *final_loss);
/* index of the smallest predicted cost per row, computed on the GPU */
int lcs_mlp_choose(lcs_mlp *mlp, const float *X, int32_t n, int32_t *choice);
/* The server chooses schedules of requests submitted with features (and no
 * tile_m) among candidates[0..ncand) with the model (NULL detaches it).
 * Declared after lcs_mlp. */
int lcs_server_set_cost_model(lcs_server *server, lcs_mlp *model, const lcs_schedule *candidates, int32_t ncand);

#ifdef __cplusplus
}
#endif
#endif /* LCS_HIP_H */
