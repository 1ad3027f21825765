// Schedule selection: which tile height, waves per SIMD and kernel variant
// a pair is combed with when its caller leaves the schedule to the library.
// In order: the context's autotune table (shapes the autotuner measured),
// the learned cost model attached to the context, the exhaustive autotune
// of the shape (when enabled; its result joins the table), and the
// built-in default.  Schedules change speed only, never the strands.
// No counterpart in the reference, whose kernel shape is a compile-time name
// (src/algorithm_registry.cpp:127-142) swept offline (tests/test_prepare.ipynb).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"

namespace lcs {

namespace {

// a table row answers for shapes within 1/8 octave of its own in m and n
constexpr double kTableRadius = 0.125;
// the cost model's inputs when the caller gives none (the comb's cost does
// not depend on the symbols: the cell rule is branch-free)
constexpr float kDefaultRate = 0.1f;

bool valid_mode(int mode) { return mode == LCS_TUNE_SINGLE || mode == LCS_TUNE_BATCH; }

// The cost model's four inputs, each scaled to about [0, 1]:
// log2 m / 18, log2 n / 18, log2 alphabet / 5, 4 * mutation rate.
void features(int32_t m, int32_t n, int32_t alphabet, float rate, float *f)
{
    f[0] = (float)(std::log2((double)std::max(m, 1)) / 18.0);
    f[1] = (float)(std::log2((double)std::max(n, 1)) / 18.0);
    f[2] = (float)(std::log2((double)std::max(alphabet, 2)) / 5.0);
    f[3] = rate * 4.0f;
}

const TunedShape *table_lookup(const Selector &S, int32_t m, int32_t n)
{
    const TunedShape *best = nullptr;
    double best_d = 1e30;
    const double lm = std::log2((double)std::max(m, 1)), ln = std::log2((double)std::max(n, 1));
    for (const auto &t : S.table) {
        const double dm = std::fabs(std::log2((double)std::max(t.m, 1)) - lm);
        const double dn = std::fabs(std::log2((double)std::max(t.n, 1)) - ln);
        if (dm > kTableRadius || dn > kTableRadius) continue;
        if (dm + dn < best_d) {
            best_d = dm + dn;
            best = &t;
        }
    }
    return best;
}

// zero fields of `given` take the chosen schedule's values
lcs_schedule merge(const lcs_schedule *given, const lcs_schedule &chosen)
{
    lcs_schedule r = chosen;
    if (given) {
        if (given->tile_m) r.tile_m = given->tile_m;
        if (given->waves_per_simd) r.waves_per_simd = given->waves_per_simd;
        if (given->variant) r.variant = given->variant;
        if (given->strips) r.strips = given->strips;
    }
    return r;
}

// The exhaustive autotune of one shape over the on-miss candidates.
int autotune_shape(lcs_ctx *ctx, int mode, int32_t m, int32_t n, int32_t alphabet, lcs_schedule &best)
{
    Selector &S = ctx->select[mode];
    std::vector<float> cost(S.miss_cands.size());
    const int rc = lcs_autotune_measure(ctx, m, n, std::max(alphabet, 1), kDefaultRate, 1, S.miss_cands.data(),
                                        (int32_t)S.miss_cands.size(), mode, S.miss_reps, cost.data());
    if (rc) return rc;
    size_t b = 0;
    for (size_t i = 1; i < cost.size(); ++i)
        if (cost[i] < cost[b]) b = i;
    best = S.miss_cands[b];
    S.table.push_back(TunedShape{m, n, best});
    return LCS_OK;
}

}  // namespace

int32_t alphabet_of(const int32_t *x, int32_t len)
{
    if (!x || len <= 0) return 4;
    int32_t mx = 0;
    for (int32_t i = 0; i < len; ++i) mx = std::max(mx, x[i]);
    return std::max<int32_t>(2, mx >= 0x3fffffff ? 0x3fffffff : mx + 1);
}

void detach_model(lcs_ctx *ctx, const lcs_mlp *M)
{
    for (auto &S : ctx->select)
        if (S.model == M) {
            S.model = nullptr;
            S.model_cands.clear();
            S.memo.clear();
        }
}

int choose_schedules(lcs_ctx *ctx, int mode, int32_t npairs, const int32_t *m, const int32_t *n,
                     const int32_t *alphabet, const lcs_schedule *given, std::vector<lcs_schedule> &out)
{
    if (!valid_mode(mode)) return LCS_EINVAL;
    Selector &S = ctx->select[mode];
    out.assign((size_t)npairs, lcs_schedule{});
    std::vector<int32_t> ask;   // pairs the model decides
    for (int32_t p = 0; p < npairs; ++p) {
        const lcs_schedule *g = given ? &given[p] : nullptr;
        if (g && g->tile_m) {
            out[(size_t)p] = *g;
            continue;
        }
        const int32_t al = alphabet ? alphabet[p] : 4;
        if (const TunedShape *t = table_lookup(S, m[p], n[p])) {
            out[(size_t)p] = merge(g, t->s);
            continue;
        }
        auto it = S.memo.find(std::make_tuple(m[p], n[p], al));
        if (it != S.memo.end()) {
            out[(size_t)p] = merge(g, it->second.first);
            continue;
        }
        if (S.model) {
            ask.push_back(p);
            continue;
        }
        if (!S.miss_cands.empty()) {
            lcs_schedule best{};
            const int rc = autotune_shape(ctx, mode, m[p], n[p], al, best);
            if (rc) return rc;
            out[(size_t)p] = merge(g, best);
            continue;
        }
        out[(size_t)p] = merge(g, lcs_schedule{});   // built-in default
    }
    if (!ask.empty()) {
        std::vector<float> X(ask.size() * 4);
        for (size_t i = 0; i < ask.size(); ++i) {
            const int32_t p = ask[i];
            features(m[p], n[p], alphabet ? alphabet[p] : 4, kDefaultRate, &X[4 * i]);
        }
        std::vector<int32_t> choice(ask.size());
        const int rc = lcs_mlp_choose(S.model, X.data(), (int32_t)ask.size(), choice.data());
        if (rc) return rc;
        for (size_t i = 0; i < ask.size(); ++i) {
            const int32_t p = ask[i];
            const int32_t c = choice[i];
            const lcs_schedule pick =
                (c >= 0 && (size_t)c < S.model_cands.size()) ? S.model_cands[(size_t)c] : lcs_schedule{};
            S.memo[std::make_tuple(m[p], n[p], alphabet ? alphabet[p] : 4)] = {pick, LCS_SCHED_FROM_MODEL};
            out[(size_t)p] = merge(given ? &given[p] : nullptr, pick);
        }
    }
    return LCS_OK;
}

int choose_schedule(lcs_ctx *ctx, int mode, int32_t m, int32_t n, int32_t alphabet, float rate,
                    const lcs_schedule *given, lcs_schedule *out, int32_t *source)
{
    if (!valid_mode(mode)) return LCS_EINVAL;
    Selector &S = ctx->select[mode];
    if (given && given->tile_m) {
        *out = *given;
        if (source) *source = LCS_SCHED_FROM_CALLER;
        return LCS_OK;
    }
    if (const TunedShape *t = table_lookup(S, m, n)) {
        *out = merge(given, t->s);
        if (source) *source = LCS_SCHED_FROM_TABLE;
        return LCS_OK;
    }
    if (S.model) {
        const auto key = std::make_tuple(m, n, alphabet);
        auto it = S.memo.find(key);
        if (it == S.memo.end()) {
            float X[4];
            features(m, n, alphabet, rate, X);
            int32_t c = -1;
            const int rc = lcs_mlp_choose(S.model, X, 1, &c);
            if (rc) return rc;
            const lcs_schedule pick =
                (c >= 0 && (size_t)c < S.model_cands.size()) ? S.model_cands[(size_t)c] : lcs_schedule{};
            it = S.memo.emplace(key, std::make_pair(pick, (int)LCS_SCHED_FROM_MODEL)).first;
        }
        *out = merge(given, it->second.first);
        if (source) *source = LCS_SCHED_FROM_MODEL;
        return LCS_OK;
    }
    if (!S.miss_cands.empty()) {
        lcs_schedule best{};
        const int rc = autotune_shape(ctx, mode, m, n, alphabet, best);
        if (rc) return rc;
        *out = merge(given, best);
        if (source) *source = LCS_SCHED_FROM_AUTOTUNE;
        return LCS_OK;
    }
    *out = merge(given, lcs_schedule{});
    if (source) *source = LCS_SCHED_FROM_DEFAULT;
    return LCS_OK;
}

}  // namespace lcs

using namespace lcs;

namespace {
bool valid_schedule(const lcs_schedule &s)
{
    return (s.tile_m == 0 || valid_tile_m(s.tile_m)) && s.waves_per_simd >= 0 && s.waves_per_simd <= 8 &&
           group_width(s.variant) != 0 && s.strips >= 0;
}
}  // namespace

extern "C" {

int lcs_ctx_add_tuned(lcs_ctx *ctx, int32_t mode, int32_t m, int32_t n, const lcs_schedule *best)
{
    if (!ctx || !valid_mode(mode) || m < 1 || n < 1 || !best || !best->tile_m || !valid_schedule(*best))
        return LCS_EINVAL;
    auto &T = ctx->select[mode].table;
    for (auto &t : T)
        if (t.m == m && t.n == n) {
            t.s = *best;
            return LCS_OK;
        }
    T.push_back(TunedShape{m, n, *best});
    return LCS_OK;
}

int lcs_ctx_clear_tuned(lcs_ctx *ctx, int32_t mode)
{
    if (!ctx || !valid_mode(mode)) return LCS_EINVAL;
    ctx->select[mode].table.clear();
    ctx->select[mode].memo.clear();
    return LCS_OK;
}

int lcs_ctx_set_cost_model(lcs_ctx *ctx, int32_t mode, lcs_mlp *model, const lcs_schedule *candidates,
                           int32_t ncand)
{
    if (!ctx || !valid_mode(mode) || ncand < 0 || (ncand && !candidates)) return LCS_EINVAL;
    if (model && (ncand < 1 || mlp_ctx(model) != ctx)) return LCS_EINVAL;
    if (model) {
        int n_out = 0;
        if (mlp_shape(model, nullptr, nullptr, &n_out) || n_out != ncand) return LCS_EINVAL;
        for (int32_t i = 0; i < ncand; ++i)
            if (!candidates[i].tile_m || !valid_schedule(candidates[i])) return LCS_EINVAL;
    }
    Selector &S = ctx->select[mode];
    S.model = model;
    S.model_cands.assign(candidates, candidates + (model ? ncand : 0));
    S.memo.clear();
    return LCS_OK;
}

int lcs_ctx_set_autotune_on_miss(lcs_ctx *ctx, int32_t mode, const lcs_schedule *candidates, int32_t ncand,
                                 int32_t reps)
{
    if (!ctx || !valid_mode(mode) || ncand < 0 || (ncand && !candidates) || reps < 1) return LCS_EINVAL;
    for (int32_t i = 0; i < ncand; ++i)
        if (!candidates[i].tile_m || !valid_schedule(candidates[i])) return LCS_EINVAL;
    Selector &S = ctx->select[mode];
    S.miss_cands.assign(candidates, candidates + ncand);
    S.miss_reps = reps;
    return LCS_OK;
}

int lcs_ctx_choose_schedule(lcs_ctx *ctx, int32_t mode, int32_t m, int32_t n, int32_t alphabet, float rate,
                            lcs_schedule *out, int32_t *source)
{
    if (!ctx || !out || m < 0 || n < 0) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    return choose_schedule(ctx, mode, m, n, alphabet, rate, nullptr, out, source);
}

}  // extern "C"
