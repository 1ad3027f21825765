/*
 * lcs_workloads.h -- seeded synthetic inputs for the benchmarks and tests.
 *
 * All generators use std::mt19937_64 (its output sequence is fixed by the C++
 * standard) followed by explicit integer / IEEE-double transforms, so the same
 * seed gives the same strings on every machine with this library.  The
 * reference's own generator (src/test_utility.hpp:8-19: std::mt19937, output
 * & 1) is lcs_gen_binary_mt19937.
 */
#ifndef LCS_WORKLOADS_H
#define LCS_WORKLOADS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* generate_random_binary_sequence(size, seed) of src/test_utility.hpp:8-19. */
void lcs_gen_binary_mt19937(int32_t seed, int32_t len, int32_t *out);

/* i.i.d. uniform symbols in [0, alphabet). */
void lcs_gen_uniform(uint64_t seed, int32_t len, int32_t alphabet, int32_t *out);

/* i.i.d. Zipf(s) symbols in [0, alphabet): P(k) proportional to (k+1)^-s. */
void lcs_gen_zipf(uint64_t seed, int32_t len, int32_t alphabet, double s, int32_t *out);

/* B from A: walk A cyclically; per source symbol, with probability sub_rate
 * substitute a different symbol, with probability indel_rate/2 insert a random
 * symbol before it, with probability indel_rate/2 delete it; stop after
 * out_len output symbols. */
void lcs_gen_mutate(uint64_t seed, const int32_t *src, int32_t src_len, int32_t alphabet,
                    double sub_rate, double indel_rate, int32_t out_len, int32_t *out);

/* Zipf(s) background of length n with `copies` mutated copies of a (each
 * with sub_rate substitutions) planted at seeded, non-overlapping offsets
 * where room allows. */
void lcs_gen_planted(uint64_t seed, const int32_t *a, int32_t m, int32_t n, int32_t alphabet, double s,
                     int32_t copies, double sub_rate, int32_t *out);

/* The k-th draw of a seeded stream mapped to [0, 1): used for per-pair
 * lengths and rates of the batched and served workloads. */
double lcs_gen_unit(uint64_t seed, int64_t k);

#ifdef __cplusplus
}
#endif
#endif
