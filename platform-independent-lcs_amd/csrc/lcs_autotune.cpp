// The autotuner: times candidate schedules on synthetic pairs of a shape.
// Its measurements are the training data of the cost model and the
// exhaustive baseline that scores the model's choices.  Schedules only
// change speed, so every candidate computes the same strands.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"
#include "lcs_workloads.h"

using namespace lcs;

extern "C" int lcs_autotune_measure(lcs_ctx *ctx, int32_t m, int32_t n, int32_t alphabet, float mutation_rate,
                                    uint64_t seed, const lcs_schedule *candidates, int32_t ncand, int32_t mode,
                                    int32_t reps, float *cost_ms)
{
    if (!ctx || m < 1 || n < 1 || alphabet < 1 || ncand < 1 || !candidates || !cost_ms || reps < 1) return LCS_EINVAL;
    if (mode != 0 && mode != 1) return LCS_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return LCS_EIO;
    // pairs: A uniform, B = A mutated (substitutions and indels 4:1).
    // Mode 1 prices a pair inside a batch that keeps every SIMD busy, as a
    // large batch of uneven pairs does: K pairs give even the coarsest
    // schedule (64*32 rows per band) two bands per SIMD, and the columns are
    // cut to n_s so that a timed launch stays near 2^30 cells; a band's time
    // grows linearly in its columns, so the cost is scaled back by
    // (n + 64) / (n_s + 64).
    const int32_t n_full = n;
    int K = 1;
    if (mode == 1) {
        const long long simds = 4LL * ctx->num_cus;
        const long long coarse = std::max<long long>(1, bands_of(m, n, 32));
        K = (int)std::min<long long>(4096, std::max<long long>(8, (2 * simds + coarse - 1) / coarse));
        const long long cap = std::max<long long>(1024, (1LL << 30) / ((long long)K * m));
        n = (int32_t)std::min<long long>(n, cap);
    }
    const double scale = (double)(n_full + 64) / (double)(n + 64);
    std::vector<int32_t> a_all((size_t)K * m), b_all((size_t)K * n);
    std::vector<int64_t> a_off((size_t)K), b_off((size_t)K);
    std::vector<int32_t> ms((size_t)K, m), ns((size_t)K, n);
    for (int p = 0; p < K; ++p) {
        a_off[(size_t)p] = (int64_t)p * m;
        b_off[(size_t)p] = (int64_t)p * n;
        lcs_gen_uniform(seed * 7919 + 2 * (uint64_t)p, m, alphabet, a_all.data() + a_off[(size_t)p]);
        lcs_gen_mutate(seed * 7919 + 2 * (uint64_t)p + 1, a_all.data() + a_off[(size_t)p], m, alphabet,
                       0.8 * mutation_rate, 0.2 * mutation_rate, n, b_all.data() + b_off[(size_t)p]);
    }
    std::vector<int32_t> h_all((size_t)K * m), v_all((size_t)K * n);
    int rc;
    if (mode == 0) {
        DevBuf buf;
        if ((rc = buf.reserve(ctx, (size_t)2 * (m + n) * 4))) return rc;
        int32_t *da = (int32_t *)buf.ptr, *db = da + m, *dh = db + n, *dv = dh + m;
        if (hipMemcpy(da, a_all.data(), (size_t)m * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(db, b_all.data(), (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess) {
            buf.release(ctx);
            return LCS_EIO;
        }
        for (int c = 0; c < ncand; ++c) {
            float best = 1e30f;
            for (int r = 0; r <= reps; ++r) {  // r == 0 warms up
                if ((rc = lcs_comb_device(ctx, da, m, db, n, &candidates[c], dh, dv, nullptr))) { buf.release(ctx); return rc; }
                if (hipStreamSynchronize(ctx->stream) != hipSuccess) { buf.release(ctx); return LCS_EIO; }
                double ms_k = 0;
                lcs_ctx_info(ctx, nullptr, &ms_k);
                if ((rc = check_abort(ctx))) { buf.release(ctx); return rc; }
                if (r) best = std::min(best, (float)ms_k);
            }
            cost_ms[c] = best;
        }
        buf.release(ctx);
        return LCS_OK;
    }
    std::vector<lcs_schedule> sched((size_t)K);
    for (int c = 0; c < ncand; ++c) {
        std::fill(sched.begin(), sched.end(), candidates[c]);
        float best = 1e30f;
        for (int r = 0; r <= reps; ++r) {
            rc = lcs_batch_comb(ctx, K, a_all.data(), a_off.data(), ms.data(), b_all.data(), b_off.data(), ns.data(),
                                sched.data(), 1, h_all.data(), v_all.data());
            if (rc) return rc;
            double ms_k = 0;
            lcs_ctx_info(ctx, nullptr, &ms_k);
            if (r) best = std::min(best, (float)(ms_k / K * scale));
        }
        cost_ms[c] = best;
    }
    return LCS_OK;
}
