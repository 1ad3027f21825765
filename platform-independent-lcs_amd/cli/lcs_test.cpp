// lcs_test: the reference's script-driven harness (src/main.cpp:16-99,
// src/test_setup.hpp:105-381) over the HIP engine's registry.
//
//   lcs_test [script.txt] [out.csv]
//
// Each script line holds key=value words (algorithm, device_type, size_a,
// size_b, seed_a, seed_b, iterations, file_a, file_b, rng_a, rng_b; '#'
// starts a comment).  Inputs are generate_random_binary_sequence
// (std::mt19937 & 1, src/test_utility.hpp:8-19) with seeds 1000, 1001, ...
// when not given, or whitespace-separated integers read from file_a/file_b
// when given (the reference parses those keys but never reads the files).
// Output: the reference's CSV columns, speed in cells/us (MCUPS), hash of
// the permutation; for score-only algorithms (hip_llcs), which produce no
// permutation, the hash column holds the LCS length.  Blank lines are
// skipped (the reference's parser does not advance past them).
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "lcs_registry.hpp"

namespace {

struct Options {
    std::string algorithm, device_type, file_a, file_b, rng_a, rng_b;
    int size_a = 256, size_b = 256, seed_a = -1, seed_b = -1, iterations = 1;
};

bool parse_int(const std::string &s, int *out)
{
    if (s.empty()) return false;
    size_t i = 0;
    int sign = 1;
    if (s[0] == '-') { sign = -1; i = 1; }
    if (i == s.size()) return false;
    long long v = 0;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        v = v * 10 + (s[i] - '0');
    }
    *out = (int)(sign * v);
    return true;
}

// One script line: words "key=value" separated by blanks; '#' to end of line
// is a comment; words without '=' are ignored.
std::map<std::string, std::string> parse_line(const std::string &line)
{
    std::map<std::string, std::string> kv;
    std::string body = line.substr(0, line.find('#'));
    size_t i = 0;
    while (i < body.size()) {
        while (i < body.size() && (body[i] == ' ' || body[i] == '\t' || body[i] == '\r')) ++i;
        size_t k0 = i;
        while (i < body.size() && body[i] != ' ' && body[i] != '\t' && body[i] != '=' && body[i] != '\r') ++i;
        std::string key = body.substr(k0, i - k0);
        while (i < body.size() && (body[i] == ' ' || body[i] == '\t')) ++i;
        if (i < body.size() && body[i] == '=') {
            ++i;
            while (i < body.size() && (body[i] == ' ' || body[i] == '\t')) ++i;
            size_t v0 = i;
            while (i < body.size() && body[i] != ' ' && body[i] != '\t' && body[i] != '\r') ++i;
            kv[key] = body.substr(v0, i - v0);
        }
    }
    return kv;
}

Options to_options(const std::map<std::string, std::string> &kv)
{
    Options o;
    auto get = [&](const char *k) { auto it = kv.find(k); return it == kv.end() ? std::string() : it->second; };
    o.algorithm = get("algorithm");
    o.device_type = get("device_type");
    o.file_a = get("file_a");
    o.file_b = get("file_b");
    o.rng_a = get("rng_a");
    o.rng_b = get("rng_b");
    int v;
    if (parse_int(get("size_a"), &v)) o.size_a = v;
    if (parse_int(get("size_b"), &v)) o.size_b = v;
    if (parse_int(get("seed_a"), &v)) o.seed_a = v;
    if (parse_int(get("seed_b"), &v)) o.seed_b = v;
    if (parse_int(get("iterations"), &v)) o.iterations = v;
    return o;
}

std::vector<int> binary_sequence(int size, int seed)
{
    std::vector<int> r((size_t)(size > 0 ? size : 0));
    std::mt19937 rng((uint32_t)seed);
    for (auto &x : r) x = (int)(rng() & 1u);
    return r;
}

bool read_ints(const std::string &path, std::vector<int> &out)
{
    std::ifstream f(path);
    if (!f) return false;
    out.clear();
    long long x;
    while (f >> x) out.push_back((int)x);
    return true;
}

}  // namespace

int main(int argc, char **argv)
{
    const char *in_name = argc >= 2 ? argv[1] : "test_default.txt";
    FILE *out = argc >= 3 ? std::fopen(argv[2], "w") : stdout;
    if (!out) { std::fprintf(stderr, "Error: cannot open %s\n", argv[2]); return 1; }
    std::ifstream in(in_name);
    if (!in) { std::fprintf(stderr, "Error: cannot read %s\n", in_name); return 1; }
    lcsreg::LcsAlgorithmRegistry reg;
    std::fprintf(out, "algorithm,device_type,device,size_a,size_b,source_a,source_b,elapsed_ms,hash,speed\n");
    int seed_counter = 1000;
    std::string line;
    int failures = 0;
    while (std::getline(in, line)) {
        const auto kv = parse_line(line);
        if (kv.empty()) continue;
        Options o = to_options(kv);
        if (o.seed_a == -1) o.seed_a = seed_counter++;
        if (o.seed_b == -1) o.seed_b = seed_counter++;
        std::vector<int> a, b;
        if (!o.file_a.empty()) {
            if (!read_ints(o.file_a, a)) { std::fprintf(stderr, "Error: cannot read %s\n", o.file_a.c_str()); ++failures; continue; }
        } else a = binary_sequence(o.size_a, o.seed_a);
        if (!o.file_b.empty()) {
            if (!read_ints(o.file_b, b)) { std::fprintf(stderr, "Error: cannot read %s\n", o.file_b.c_str()); ++failures; continue; }
        } else b = binary_sequence(o.size_b, o.seed_b);
        auto impl = reg.Get(o.algorithm, o.device_type);
        if (!impl.ok) {
            std::fprintf(stderr, "Error: Unable to find algorihm '%s'\n", o.algorithm.c_str());
            ++failures;
            continue;
        }
        char device[256] = "unknown";
        if (impl.queue) lcs_ctx_device_name(impl.queue, device, sizeof(device));
        lcsreg::LcsInput input(a.data(), (int)a.size(), b.data(), (int)b.size());
        for (int it = 0; it < o.iterations; ++it) {
            const auto t0 = std::chrono::steady_clock::now();
            lcsreg::PermutationMatrix perm;
            int64_t result = 0;
            if (impl.score_only) {
                result = impl.length(input);   // score-only names: the hash column holds the LCS length
            } else {
                perm = impl(input);
                result = perm.hash();
            }
            const auto t1 = std::chrono::steady_clock::now();
            if (impl.score_only && result < 0) {
                std::fprintf(stderr, "Error: algorithm '%s' on '%s' produced no length\n", o.algorithm.c_str(),
                             o.device_type.c_str());
                ++failures;
                continue;
            }
            if (!impl.score_only && perm.empty() && a.size() + b.size() > 0) {
                std::fprintf(stderr, "Error: algorithm '%s' on '%s' produced no permutation\n", o.algorithm.c_str(),
                             o.device_type.c_str());
                ++failures;
                continue;
            }
            const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            const double speed = (double)a.size() * (double)b.size() / (ms * 1000.0);
            std::fprintf(out, "%s,%s,%s,%d,%d,", o.algorithm.c_str(), o.device_type.c_str(), device, (int)a.size(),
                         (int)b.size());
            if (!o.file_a.empty()) std::fprintf(out, "%s,", o.file_a.c_str());
            else std::fprintf(out, "%s(%d),", o.rng_a.c_str(), o.seed_a);
            if (!o.file_b.empty()) std::fprintf(out, "%s,", o.file_b.c_str());
            else std::fprintf(out, "%s(%d),", o.rng_b.c_str(), o.seed_b);
            std::fprintf(out, "%f,%" PRId64 ",%f\n", ms, result, speed);
        }
        std::fflush(out);
    }
    if (out != stdout) std::fclose(out);
    return failures ? 2 : 0;
}
