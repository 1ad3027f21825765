/* Test infrastructure: a self-checking driver for the CPU oracle, built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by `make -C oracle
 * sanitize` (tests/test_oracle_sanitizers.py runs it).  It drives every comb
 * order, the strand read-outs and the steady ant on small seeded inputs,
 * including empty and one-symbol strings, and checks each against an
 * independent answer:
 *   - the four comb orders give the same strands;
 *   - the LCS read from the strands equals the textbook DP and the
 *     bit-parallel recurrence;
 *   - the string-substring scores read from the permutation equal the DP on
 *     every substring b[i:j);
 *   - the steady-ant core equals the sticky product computed from its
 *     definition.
 * Exit status 0 and a final "ok" line mean every check passed; a sanitizer
 * report aborts the run. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lcs_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static uint32_t next_u32(void)
{
    /* xorshift64* */
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (uint32_t)((rng_state * 0x2545f4914f6cdd1dull) >> 32);
}

static int32_t *random_string(int32_t len, int32_t sigma)
{
    int32_t *s = (int32_t *)malloc(sizeof(int32_t) * (size_t)(len > 0 ? len : 1));
    for (int32_t i = 0; i < len; ++i) s[i] = (int32_t)(next_u32() % (uint32_t)sigma);
    return s;
}

static int fail(const char *what, int32_t m, int32_t n)
{
    fprintf(stderr, "mismatch: %s (m=%d n=%d)\n", what, m, n);
    return 1;
}

static int check_pair(int32_t m, int32_t n, int32_t sigma)
{
    int32_t *a = random_string(m, sigma), *b = random_string(n, sigma);
    const size_t hb = sizeof(int32_t) * (size_t)(m > 0 ? m : 1), vb = sizeof(int32_t) * (size_t)(n > 0 ? n : 1);
    int32_t *h0 = malloc(hb), *v0 = malloc(vb), *h1 = malloc(hb), *v1 = malloc(vb);
    int32_t *h2 = malloc(hb), *v2 = malloc(vb), *h3 = malloc(hb), *v3 = malloc(vb);
    int bad = 0;
    orc_comb_antidiag(a, m, b, n, h0, v0);
    orc_comb_rows(a, m, b, n, h1, v1);
    orc_comb_mt(a, m, b, n, h2, v2, 2);
    orc_comb_simd(a, m, b, n, h3, v3, 2, 0);
    if (memcmp(h0, h1, sizeof(int32_t) * (size_t)m) || memcmp(v0, v1, sizeof(int32_t) * (size_t)n))
        bad |= fail("rows vs antidiag", m, n);
    if (memcmp(h0, h2, sizeof(int32_t) * (size_t)m) || memcmp(v0, v2, sizeof(int32_t) * (size_t)n))
        bad |= fail("mt vs antidiag", m, n);
    if (memcmp(h0, h3, sizeof(int32_t) * (size_t)m) || memcmp(v0, v3, sizeof(int32_t) * (size_t)n))
        bad |= fail("simd vs antidiag", m, n);
    const int32_t llcs = orc_llcs_from_strands(v0, m, n);
    if (llcs != orc_llcs_dp(a, m, b, n)) bad |= fail("llcs from strands vs dp", m, n);
    if (m > 0 && n > 0 && llcs != orc_llcs_bitpar(a, m, b, n)) bad |= fail("llcs bit-parallel vs dp", m, n);
    if (m + n > 0 && n <= 48) {
        const int32_t size = m + n;
        int32_t *r2c = malloc(sizeof(int32_t) * (size_t)size), *c2r = malloc(sizeof(int32_t) * (size_t)size);
        int32_t *H = malloc(sizeof(int32_t) * (size_t)(n + 1) * (size_t)(n + 1));
        orc_from_strands(h0, m, v0, n, r2c, c2r);
        orc_string_substring(r2c, m, n, H);
        for (int32_t i = 0; i <= n && !bad; ++i)
            for (int32_t j = i; j <= n; ++j)
                if (H[(size_t)i * (size_t)(n + 1) + (size_t)j] != orc_llcs_dp(a, m, b + i, j - i)) {
                    bad |= fail("string-substring score vs dp", m, n);
                    break;
                }
        free(r2c); free(c2r); free(H);
    }
    free(a); free(b); free(h0); free(v0); free(h1); free(v1); free(h2); free(v2); free(h3); free(v3);
    return bad;
}

static void random_permutation(int32_t *p, int32_t k)
{
    for (int32_t i = 0; i < k; ++i) p[i] = i;
    for (int32_t i = k - 1; i > 0; --i) {
        const int32_t j = (int32_t)(next_u32() % (uint32_t)(i + 1));
        const int32_t t = p[i]; p[i] = p[j]; p[j] = t;
    }
}

static int check_steady_ant(int32_t k)
{
    int32_t p[16], q[16], r0[16], r1[16];
    random_permutation(p, k);
    random_permutation(q, k);
    orc_steady_ant(p, q, k, r0);
    orc_sticky_bruteforce(p, q, k, r1);
    return memcmp(r0, r1, sizeof(int32_t) * (size_t)k) ? fail("steady ant vs sticky product", k, k) : 0;
}

int main(void)
{
    static const int32_t sizes[][3] = {{0, 0, 4}, {0, 7, 4}, {5, 0, 4}, {1, 1, 2}, {1, 9, 2},   {9, 1, 3},
                                       {17, 23, 4}, {64, 64, 4}, {65, 31, 26}, {130, 47, 4}, {200, 333, 4},
                                       {333, 200, 20}, {1000, 777, 4}};
    int bad = 0;
    for (size_t c = 0; c < sizeof(sizes) / sizeof(sizes[0]); ++c)
        for (int rep = 0; rep < 3; ++rep) bad |= check_pair(sizes[c][0], sizes[c][1], sizes[c][2]);
    for (int32_t k = 1; k <= 12; ++k)
        for (int rep = 0; rep < 20; ++rep) bad |= check_steady_ant(k);
    if (bad) return 1;
    printf("ok\n");
    return 0;
}
