// The reference's registry boundary (src/algorithm_registry.{hpp,cpp}) over
// the engine's C ABI.  See include/lcs_registry.hpp.
#include "lcs_registry.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

namespace lcsreg {

PermutationMatrix::PermutationMatrix(int size) : size_(size)
{
    if (size > 0) {
        row_to_col_.assign((size_t)size, -1);
        col_to_row_.assign((size_t)size, -1);
    } else {
        size_ = 0;
    }
}

PermutationMatrix::PermutationMatrix(PermutationMatrix &&o) noexcept
    : row_to_col_(std::move(o.row_to_col_)), col_to_row_(std::move(o.col_to_row_)), size_(o.size_)
{
    o.size_ = 0;
}

PermutationMatrix &PermutationMatrix::operator=(PermutationMatrix &&o) noexcept
{
    row_to_col_ = std::move(o.row_to_col_);
    col_to_row_ = std::move(o.col_to_row_);
    size_ = o.size_;
    o.size_ = 0;
    return *this;
}

PermutationMatrix PermutationMatrix::FromStrands(const int *h_strands, int m, const int *v_strands, int n)
{
    PermutationMatrix p(m + n);
    if (m + n > 0) lcs_permutation_from_strands(h_strands, m, v_strands, n, p.row_to_col_.data(), p.col_to_row_.data());
    return p;
}

PermutationMatrix PermutationMatrix::FromRowToCol(const int *row_to_col, int size)
{
    PermutationMatrix p(size);
    for (int r = 0; r < size; ++r)
        if (row_to_col[r] >= 0 && row_to_col[r] < size) p.set_point(r, row_to_col[r]);
    return p;
}

void PermutationMatrix::set_point(int row, int col)
{
    row_to_col_[(size_t)row] = col;
    col_to_row_[(size_t)col] = row;
}

void PermutationMatrix::unset_all()
{
    std::fill(row_to_col_.begin(), row_to_col_.end(), -1);
    std::fill(col_to_row_.begin(), col_to_row_.end(), -1);
}

bool PermutationMatrix::operator==(const PermutationMatrix &o) const
{
    return size_ == o.size_ && row_to_col_ == o.row_to_col_ && col_to_row_ == o.col_to_row_;
}

double PermutationMatrix::similarity(const PermutationMatrix &o) const
{
    if (size_ != o.size_ || size_ == 0) return 0.0;
    long long good = 0;
    for (int i = 0; i < size_; ++i) {
        good += row_to_col_[(size_t)i] == o.row_to_col_[(size_t)i];
        good += col_to_row_[(size_t)i] == o.col_to_row_[(size_t)i];
    }
    return (double)good / (2.0 * size_);
}

int64 PermutationMatrix::hash() const { return lcs_permutation_hash(col_to_row_.data(), size_); }

// ------------------------------------------------------------ procedures

void Lcs_Semi_Hip(LcsProblemContext &ctx)
{
    const int m = ctx.a_length, n = ctx.b_length;
    ctx.status = LCS_EINVAL;
    if (!ctx.queue || m < 0 || n < 0) return;
    int *h = new (std::nothrow) int[(size_t)(m > 0 ? m : 1)];
    int *v = new (std::nothrow) int[(size_t)(n > 0 ? n : 1)];
    if (!h || !v) {
        delete[] h;
        delete[] v;
        ctx.status = LCS_ENOMEM;
        return;
    }
    int32_t llcs = -1;
    const int rc = lcs_comb(ctx.queue, (const int32_t *)ctx.a_given, m, (const int32_t *)ctx.b_given, n,
                            &ctx.schedule, h, v, &llcs);
    ctx.status = rc;
    if (rc != LCS_OK) {
        delete[] h;
        delete[] v;
        return;
    }
    delete[] ctx.h_strands;
    delete[] ctx.v_strands;
    ctx.h_strands = reinterpret_cast<Word *>(h);
    ctx.v_strands = reinterpret_cast<Word *>(v);
    ctx.h_strands_length = m;
    ctx.v_strands_length = n;
    ctx.length_of_lcs = llcs;
}

void Lcs_Semi_Hip_Function(const LcsInput &input, LcsContext &ctx)
{
    LcsProblemContext p;
    p.a_given = reinterpret_cast<const Word *>(input.a_data);
    p.b_given = reinterpret_cast<const Word *>(input.b_data);
    p.a_length = input.a_size;
    p.b_length = input.b_size;
    p.queue = ctx.queue;
    Lcs_Semi_Hip(p);
    if (!p.h_strands) return;
    ctx.h_strands = reinterpret_cast<int *>(p.h_strands);
    ctx.v_strands = reinterpret_cast<int *>(p.v_strands);
    ctx.h_strands_size = p.h_strands_length;
    ctx.v_strands_size = p.v_strands_length;
    ctx.llcs = p.length_of_lcs;
    p.h_strands = nullptr;
    p.v_strands = nullptr;
}

void Lcs_Llcs_Hip(LcsProblemContext &ctx)
{
    lcs_ctx *gpu = ctx.queue;
    if (!gpu || ctx.a_length < 0 || ctx.b_length < 0) return;
    const int32_t *a = reinterpret_cast<const int32_t *>(ctx.a_given);
    const int32_t *b = reinterpret_cast<const int32_t *>(ctx.b_given);
    const int32_t m = ctx.a_length, n = ctx.b_length;
    // the bit-parallel kernel takes symbols in [0, sigma), sigma <= 200
    int32_t lo = 0, hi = 0;
    for (int32_t i = 0; i < m; ++i) lo = std::min(lo, a[i]), hi = std::max(hi, a[i]);
    for (int32_t j = 0; j < n; ++j) lo = std::min(lo, b[j]), hi = std::max(hi, b[j]);
    int32_t llcs = -1;
    if (lo >= 0 && hi < 200) {
        if (lcs_llcs(gpu, a, m, b, n, hi + 1, &llcs) == LCS_OK) ctx.length_of_lcs = llcs;
        return;
    }
    std::vector<int32_t> h((size_t)std::max(m, 1)), v((size_t)std::max(n, 1));
    if (lcs_comb(gpu, a, m, b, n, nullptr, h.data(), v.data(), &llcs) == LCS_OK) ctx.length_of_lcs = llcs;
}

// ------------------------------------------------------------ registry

int SemiLocalLcsImpl::length(const LcsInput &input)
{
    if (func) {
        LcsContext ctx(queue);
        func(input, ctx);
        return ctx.llcs;
    }
    if (proc) {
        LcsProblemContext p;
        p.a_given = reinterpret_cast<const Word *>(input.a_data);
        p.a_length = input.a_size;
        p.b_given = reinterpret_cast<const Word *>(input.b_data);
        p.b_length = input.b_size;
        p.queue = queue;
        p.schedule = schedule;
        proc(p);
        return p.length_of_lcs;
    }
    return -1;
}

PermutationMatrix SemiLocalLcsImpl::operator()(const LcsInput &input)
{
    if (func) {
        LcsContext ctx(queue);
        func(input, ctx);
        if (ctx.h_strands && ctx.v_strands)
            return PermutationMatrix::FromStrands(ctx.h_strands, ctx.h_strands_size, ctx.v_strands,
                                                  ctx.v_strands_size);
        return std::move(ctx.matrix);
    }
    if (proc && device_permutation && queue) {
        // the same result as the procedure + FromStrands below, with the
        // permutation built on the GPU
        const int m = input.a_size, n = input.b_size;
        if (m < 0 || n < 0) return PermutationMatrix();
        PermutationMatrix pm(m + n);
        if (lcs_comb_permutation(queue, input.a_data, m, input.b_data, n, &schedule, pm.row_to_col_.data(),
                                 pm.col_to_row_.data(), nullptr) != LCS_OK)
            return PermutationMatrix();
        return pm;
    }
    if (proc) {
        LcsProblemContext p;
        p.a_given = reinterpret_cast<const Word *>(input.a_data);
        p.a_length = input.a_size;
        p.b_given = reinterpret_cast<const Word *>(input.b_data);
        p.b_length = input.b_size;
        p.queue = queue;
        p.schedule = schedule;
        proc(p);
        if (p.h_strands && p.v_strands)
            return PermutationMatrix::FromStrands(reinterpret_cast<int *>(p.h_strands), p.h_strands_length,
                                                  reinterpret_cast<int *>(p.v_strands), p.v_strands_length);
    }
    return PermutationMatrix();
}

LcsAlgorithmRegistry::LcsAlgorithmRegistry()
{
    reg2["hip_auto"] = Lcs_Semi_Hip;
    schedules["hip_auto"] = lcs_schedule{0, 0, 0, 0};
    // "tiled": the name the reference's own run scripts give its tiled comb
    // (tests/test_prepare.ipynb), so those scripts run with device_type=gpu
    reg2["tiled"] = Lcs_Semi_Hip;
    schedules["tiled"] = lcs_schedule{0, 0, 0, 0};
    reg["hip_auto_fn"] = Lcs_Semi_Hip_Function;
    reg2["hip_llcs"] = Lcs_Llcs_Hip;
    score_only["hip_llcs"] = true;
    const int tiles[] = {1, 2, 4, 8, 16, 32};
    const int strips[] = {1, 2, 4, 8, 16};
    for (int tm : tiles)
        for (int wps = 1; wps <= 2; ++wps)
            for (int s : strips) {
                char name[64];
                std::snprintf(name, sizeof(name), "hip_tiled_%d_%d_%d", tm, wps, s);
                reg2[name] = Lcs_Semi_Hip;
                schedules[name] = lcs_schedule{tm, wps, LCS_VARIANT_CHAIN, s};
            }
}

LcsAlgorithmRegistry::~LcsAlgorithmRegistry()
{
    if (QUEUE_GPU) lcs_ctx_destroy(QUEUE_GPU);
}

lcs_ctx *LcsAlgorithmRegistry::GetQueueForDeviceType(const std::string &device_type)
{
    if (device_type == "gpu") {
        if (!QUEUE_GPU && lcs_ctx_create(0, 0, &QUEUE_GPU) != LCS_OK) QUEUE_GPU = nullptr;
        return QUEUE_GPU;
    }
    return nullptr;  // the reference returns a null queue for unknown types too
}

SemiLocalLcsImpl LcsAlgorithmRegistry::Get(const std::string &name, const std::string &device_type)
{
    SemiLocalLcsImpl result{};
    if (reg.count(name)) {
        result.queue = GetQueueForDeviceType(device_type);
        result.func = reg[name];
        result.ok = true;  // as in the reference: a known name is ok, the queue may still be null
        return result;
    }
    if (reg2.count(name)) {
        result.queue = GetQueueForDeviceType(device_type);
        result.proc = reg2[name];
        result.schedule = schedules[name];
        result.score_only = score_only.count(name) > 0;
        result.device_permutation = result.proc == Lcs_Semi_Hip;
        result.ok = true;  // as in the reference: a known name is ok, the queue may still be null
        return result;
    }
    return result;
}

}  // namespace lcsreg
