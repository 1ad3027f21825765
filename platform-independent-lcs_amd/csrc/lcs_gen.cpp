// Seeded synthetic workloads (declared in include/lcs_workloads.h).
#include <cmath>
#include <random>
#include <vector>

#include "lcs_workloads.h"

namespace {
inline uint32_t below(uint64_t x, uint32_t k) { return (uint32_t)(((x >> 32) * (uint64_t)k) >> 32); }
inline double unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

std::vector<double> zipf_cdf(int32_t alphabet, double s)
{
    std::vector<double> cdf((size_t)alphabet);
    double acc = 0.0;
    for (int32_t k = 0; k < alphabet; ++k) { acc += std::pow((double)(k + 1), -s); cdf[(size_t)k] = acc; }
    for (auto &c : cdf) c /= acc;
    cdf.back() = 1.0;
    return cdf;
}
int32_t zipf_draw(const std::vector<double> &cdf, double u)
{
    int32_t lo = 0, hi = (int32_t)cdf.size() - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi) / 2;
        if (u < cdf[(size_t)mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}
}  // namespace

extern "C" {

void lcs_gen_binary_mt19937(int32_t seed, int32_t len, int32_t *out)
{
    std::mt19937 rng((uint32_t)seed);
    for (int32_t i = 0; i < len; ++i) out[i] = (int32_t)(rng() & 1u);
}

void lcs_gen_uniform(uint64_t seed, int32_t len, int32_t alphabet, int32_t *out)
{
    std::mt19937_64 rng(seed);
    for (int32_t i = 0; i < len; ++i) out[i] = (int32_t)below(rng(), (uint32_t)alphabet);
}

void lcs_gen_zipf(uint64_t seed, int32_t len, int32_t alphabet, double s, int32_t *out)
{
    std::mt19937_64 rng(seed);
    const auto cdf = zipf_cdf(alphabet, s);
    for (int32_t i = 0; i < len; ++i) out[i] = zipf_draw(cdf, unit(rng()));
}

void lcs_gen_mutate(uint64_t seed, const int32_t *src, int32_t src_len, int32_t alphabet, double sub_rate,
                    double indel_rate, int32_t out_len, int32_t *out)
{
    std::mt19937_64 rng(seed);
    int32_t o = 0;
    int64_t i = 0;
    if (src_len <= 0) {
        for (; o < out_len; ++o) out[o] = (int32_t)below(rng(), (uint32_t)alphabet);
        return;
    }
    while (o < out_len) {
        const int32_t sym = src[i % src_len];
        ++i;
        const double u = unit(rng());
        if (u < sub_rate) {
            const uint32_t shift = alphabet > 1 ? 1u + below(rng(), (uint32_t)(alphabet - 1)) : 0u;
            out[o++] = (int32_t)(((uint32_t)sym + shift) % (uint32_t)alphabet);
        } else if (u < sub_rate + 0.5 * indel_rate) {
            out[o++] = (int32_t)below(rng(), (uint32_t)alphabet);
            if (o < out_len) out[o++] = sym;
        } else if (u < sub_rate + indel_rate) {
            // deletion: emit nothing
        } else {
            out[o++] = sym;
        }
    }
}

void lcs_gen_planted(uint64_t seed, const int32_t *a, int32_t m, int32_t n, int32_t alphabet, double s,
                     int32_t copies, double sub_rate, int32_t *out)
{
    lcs_gen_zipf(seed, n, alphabet, s, out);
    if (m <= 0 || m > n || copies <= 0) return;
    std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ull);
    const int64_t slots = n / m;  // non-overlapping slots of length m
    std::vector<int32_t> tmp((size_t)m);
    for (int32_t c = 0; c < copies; ++c) {
        int64_t off;
        if (slots >= copies) {
            const int64_t slot = (int64_t)c * slots / copies + below(rng(), (uint32_t)(slots / copies));
            off = slot * m;
        } else {
            off = below(rng(), (uint32_t)(n - m + 1));
        }
        lcs_gen_mutate(rng(), a, m, alphabet, sub_rate, 0.0, m, tmp.data());
        for (int32_t i = 0; i < m; ++i) out[off + i] = tmp[(size_t)i];
    }
}

double lcs_gen_unit(uint64_t seed, int64_t k)
{
    std::mt19937_64 rng(seed);
    if (k > 0) rng.discard((unsigned long long)k);
    return unit(rng());
}

}  // extern "C"
