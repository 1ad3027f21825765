// lcs_registry.hpp -- C++ host interface mirroring the reference's plugin
// boundary (egor-bv/platform-independent-lcs, paths relative to its root),
// implemented over the C ABI of include/lcs_hip.h.
//
//   Word, LcsProblemContext, LcsProcedure   src/lcs_interface_internal.hpp:16-92
//   LcsInput, LcsContext, LcsFunction       src/lcs_types.hpp:10-55
//   PermutationMatrix                       src/permutation_matrix.hpp:5-204
//   SemiLocalLcsImpl, LcsAlgorithmRegistry  src/algorithm_registry.hpp:15-42,
//                                           src/algorithm_registry.cpp:43-173
//
// Same names, argument meaning and error behaviour: Get() of an unknown
// name returns an impl with ok == false; a failed procedure leaves the
// strands unset and operator() returns an empty PermutationMatrix.  The
// registry owns one engine context per device type, as the reference owns
// one sycl::queue per device type; the only device type is "gpu".
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "lcs_hip.h"

namespace lcsreg {

using int64 = long long;

// A 32-bit symbol or strand id compared with == and signed >.
struct Word {
    union {
        int value;
        uint32_t uvalue;
    };
    Word() : value(0x0f0f0f0f) {}
    Word(int x) : value(x) {}
    bool operator==(const Word &o) const { return value == o.value; }
    bool operator>(const Word &o) const { return value > o.value; }
};
static_assert(sizeof(Word) == 4, "Word must be 32 bits");

class PermutationMatrix {
public:
    PermutationMatrix() = default;
    explicit PermutationMatrix(int size);
    PermutationMatrix(PermutationMatrix &&o) noexcept;
    PermutationMatrix &operator=(PermutationMatrix &&o) noexcept;
    PermutationMatrix(const PermutationMatrix &) = delete;
    PermutationMatrix &operator=(const PermutationMatrix &) = delete;

    // row = strand id, col = end position: h[l] -> n+l, v[j] -> j
    static PermutationMatrix FromStrands(const int *h_strands, int m, const int *v_strands, int n);
    static PermutationMatrix FromRowToCol(const int *row_to_col, int size);

    int row_count() const { return size_; }
    int col_count() const { return size_; }
    int get_col(int row) const { return row_to_col_[(size_t)row]; }
    int get_row(int col) const { return col_to_row_[(size_t)col]; }
    void set_point(int row, int col);
    void unset_all();
    bool operator==(const PermutationMatrix &o) const;
    double similarity(const PermutationMatrix &o) const;
    // rolling hash of the reference's CSV (two's-complement wrap made explicit)
    int64 hash() const;
    const int *row_to_col() const { return row_to_col_.data(); }
    const int *col_to_row() const { return col_to_row_.data(); }
    bool empty() const { return size_ == 0; }

private:
    friend struct SemiLocalLcsImpl;   // fills both arrays from the GPU (lcs_comb_permutation)
    std::vector<int> row_to_col_, col_to_row_;
    int size_ = 0;
};

struct LcsInput {
    const int *a_data;
    const int *b_data;
    int a_size;
    int b_size;
    LcsInput(const int *a, int a_sz, const int *b, int b_sz) : a_data(a), b_data(b), a_size(a_sz), b_size(b_sz) {}
};

// The engine context stands where the reference keeps a sycl::queue*.
struct LcsContext {
    lcs_ctx *queue;
    int *h_strands = nullptr;
    int *v_strands = nullptr;
    int h_strands_size = -1;
    int v_strands_size = -1;
    int llcs = -1;
    PermutationMatrix matrix;
    explicit LcsContext(lcs_ctx *q = nullptr) : queue(q) {}
    ~LcsContext()
    {
        delete[] h_strands;
        delete[] v_strands;
    }
};

typedef void LcsFunction(const LcsInput &input, LcsContext &ctx);

struct RuntimeParameters {
    int num_sections;
};

struct LcsProblemContext {
    const Word *a_given = nullptr;
    const Word *b_given = nullptr;
    int a_length = 0;
    int b_length = 0;
    lcs_ctx *queue = nullptr;
    RuntimeParameters params{};
    Word *a_prepared = nullptr;  // reversed a, as in the reference
    Word *b_prepared = nullptr;
    Word *h_strands = nullptr;
    Word *v_strands = nullptr;
    int h_strands_length = 0;
    int v_strands_length = 0;
    int length_of_lcs = -1;
    PermutationMatrix permutation_result;
    lcs_schedule schedule{};     // the variant's schedule (set by the registry)
    int status = LCS_OK;         // last engine status of the procedure
    ~LcsProblemContext()
    {
        delete[] h_strands;
        delete[] v_strands;
        delete[] a_prepared;
        delete[] b_prepared;
    }
};

typedef void LcsProcedure(LcsProblemContext &ctx);

struct SemiLocalLcsImpl {
    bool ok = false;
    bool score_only = false;   // the procedure fills length_of_lcs only (no strands)
    // the procedure is the GPU comb: operator() builds the permutation on the
    // GPU (lcs_comb_permutation) instead of FromStrands on the host
    bool device_permutation = false;
    lcs_ctx *queue = nullptr;
    LcsFunction *func = nullptr;
    LcsProcedure *proc = nullptr;
    lcs_schedule schedule{};
    PermutationMatrix operator()(const LcsInput &input);
    // Extension (the reference returns permutations only): the procedure's
    // length_of_lcs, -1 on failure.  Works for every name; the score-only
    // ones compute nothing else.
    int length(const LcsInput &input);
};

// Names: "hip_auto" (schedule chosen by the engine) and
// "hip_tiled_<tile_m>_<waves_per_simd>_<strips>" for tile_m in
// {1,2,4,8,16,32}, waves_per_simd in {1,2}, strips in {1,2,4,8,16};
// "hip_llcs": score only (length_of_lcs), the bit-parallel kernel when every
// symbol lies in [0, 200), else the comb's read-out.
struct LcsAlgorithmRegistry {
    std::map<std::string, LcsFunction *> reg;
    std::map<std::string, LcsProcedure *> reg2;
    std::map<std::string, lcs_schedule> schedules;
    std::map<std::string, bool> score_only;
    lcs_ctx *QUEUE_GPU = nullptr;

    lcs_ctx *GetQueueForDeviceType(const std::string &device_type);
    SemiLocalLcsImpl Get(const std::string &name, const std::string &device_type);

    LcsAlgorithmRegistry();
    ~LcsAlgorithmRegistry();
    LcsAlgorithmRegistry(const LcsAlgorithmRegistry &) = delete;
    LcsAlgorithmRegistry &operator=(const LcsAlgorithmRegistry &) = delete;
};

// The procedure every hip_* name maps to: combs on the GPU with
// ctx.schedule and fills h_strands / v_strands (new Word[]).
void Lcs_Semi_Hip(LcsProblemContext &ctx);
// LcsFunction form (legacy signature of src/lcs_types.hpp:55).
void Lcs_Semi_Hip_Function(const LcsInput &input, LcsContext &ctx);
// Score-only procedure ("hip_llcs"): sets length_of_lcs, leaves the strands
// unset, as the reference's unregistered score routines do
// (Lcs_Prefix_Binary_Reference, src/lcs_reference.hpp:135-211).
void Lcs_Llcs_Hip(LcsProblemContext &ctx);

}  // namespace lcsreg
