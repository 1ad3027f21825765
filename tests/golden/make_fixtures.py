#!/usr/bin/env python3
"""Writes tests/golden/comb_fixtures.json.

Inputs come from seeded generators; expected strands from the CPU oracle
(pinned by DP, semi-local windows and the composition law, see
tests/test_oracle.py); the permutation hash from the reference's own
PermutationMatrix (src/permutation_matrix.hpp:192-203, compiled from source
into oracle/_ref) when that build exists, else from the oracle's
restatement (they are tested equal).  The reference itself stores no
expected outputs (its CSVs are git-ignored), so these fixtures are the
project's own regression vectors.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "platform-independent-lcs_amd")]

import lcs_amd  # noqa: E402  (generators only)
import oracle  # noqa: E402


def perm_hash(h, v):
    if oracle.ref_available():
        return oracle.ref_from_strands_hash(h, v)[2]
    return oracle.perm_hash(oracle.from_strands(h, v)[1])


def digest(h, v):
    return hashlib.sha256(np.concatenate([h, v]).astype("<i4").tobytes()).hexdigest()


def main():
    cases = [{"a": [1, 2], "b": [1]}, {"a": [5], "b": [5]}, {"a": [], "b": [3, 4]}, {"a": [7, 7], "b": []}]
    rng = np.random.default_rng(2024)
    for t in range(24):
        m, n, s = int(rng.integers(1, 48)), int(rng.integers(1, 48)), int(rng.integers(2, 6))
        cases.append({"a": rng.integers(0, s, m).tolist(), "b": rng.integers(0, s, n).tolist()})
    for c in cases:
        h, v = oracle.comb(c["a"], c["b"])
        c["h"], c["v"] = h.tolist(), v.tolist()
        c["hash"] = perm_hash(h, v)
    big = []
    specs = [("binary", 7330, 9944, 888, 999), ("binary", 9771, 11302, 888, 999), ("uniform4096", 4096, 4096, 1, 2)]
    for kind, m, n, sa, sb in specs:
        if kind == "binary":   # generate_random_binary_sequence (src/test_utility.hpp:8-19)
            a, b = lcs_amd.gen_binary(sa, m), lcs_amd.gen_binary(sb, n)
        else:
            a, b = lcs_amd.gen_uniform(sa, m, 4), lcs_amd.gen_uniform(sb, n, 4)
        h, v = oracle.comb(a, b)
        big.append({"kind": kind, "m": m, "n": n, "seed_a": sa, "seed_b": sb, "llcs": int((v < m).sum()),
                    "hash": perm_hash(h, v), "sha256": digest(h, v)})
    with open(os.path.join(HERE, "comb_fixtures.json"), "w") as f:
        json.dump({"cases": cases, "large": big}, f, separators=(",", ":"))
    print(f"wrote {len(cases)} small and {len(big)} large fixtures")


if __name__ == "__main__":
    main()
