"""The reference-style boundary: registry names, the LcsProcedure path and
the lcs_test harness (CSV hash) against the committed golden fixtures."""
import csv
import io
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "platform-independent-lcs_amd", "bin", "lcs_test")

pytestmark = pytest.mark.gpu


def golden():
    with open(os.path.join(ROOT, "tests", "golden", "comb_fixtures.json")) as f:
        return {(c["m"], c["n"], c["seed_a"], c["seed_b"]): c for c in json.load(f)["large"] if c["kind"] == "binary"}


def test_cli_script_matches_golden_hashes(tmp_path):
    script = tmp_path / "script.txt"
    script.write_text(
        "algorithm=hip_auto device_type=gpu size_a=7330 size_b=9944 seed_a=888 seed_b=999 iterations=2\n"
        "\n"
        "# a comment line\n"
        "algorithm=hip_tiled_4_1_2 device_type=gpu size_a=9771 size_b=11302 seed_a=888 seed_b=999\n"
        "algorithm=no_such_algorithm device_type=gpu size_a=10 size_b=10\n"
        "algorithm=hip_tiled_16_1_1 device_type=gpu size_a=9771 size_b=11302 seed_a=888 seed_b=999\n"
        "algorithm=tiled device_type=gpu size_a=7330 size_b=9944 seed_a=888 seed_b=999\n")
    out = tmp_path / "out.csv"
    r = subprocess.run([CLI, str(script), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r.stderr          # one unknown algorithm
    assert "no_such_algorithm" in r.stderr
    rows = list(csv.DictReader(io.StringIO(out.read_text())))
    assert [x["algorithm"] for x in rows] == ["hip_auto", "hip_auto", "hip_tiled_4_1_2", "hip_tiled_16_1_1", "tiled"]
    g = golden()
    for row in rows:
        key = (int(row["size_a"]), int(row["size_b"]), 888, 999)
        assert int(row["hash"]) == g[key]["hash"], row
        assert row["source_a"] == "(888)" and float(row["speed"]) > 0


def test_cli_unknown_device_type_fails_cleanly(tmp_path):
    script = tmp_path / "s.txt"
    script.write_text("algorithm=hip_auto device_type=fpga size_a=100 size_b=100\n")
    r = subprocess.run([CLI, str(script), str(tmp_path / "o.csv")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "produced no permutation" in r.stderr


def test_cli_reads_input_files(tmp_path):
    import numpy as np
    import oracle
    rng = np.random.default_rng(0)
    a = rng.integers(0, 26, 3000)
    b = rng.integers(0, 26, 5000)
    (tmp_path / "a.txt").write_text(" ".join(map(str, a)))
    (tmp_path / "b.txt").write_text("\n".join(map(str, b)))
    script = tmp_path / "s.txt"
    script.write_text(f"algorithm=hip_auto device_type=gpu file_a={tmp_path / 'a.txt'} file_b={tmp_path / 'b.txt'}\n")
    out = tmp_path / "o.csv"
    subprocess.run([CLI, str(script), str(out)], check=True, timeout=60)
    row = next(csv.DictReader(io.StringIO(out.read_text())))
    h, v = oracle.comb(a, b)
    assert int(row["hash"]) == oracle.perm_hash(oracle.from_strands(h, v)[1])


def test_cli_score_only_llcs(tmp_path):
    """hip_llcs: the hash column holds the LCS length; binary script inputs
    against the golden fixtures, a 26-symbol file pair (bit-parallel path)
    and a pair with symbols outside [0, 200) (the comb read-out) against the
    DP."""
    import numpy as np
    import oracle
    rng = np.random.default_rng(3)
    a26, b26 = rng.integers(0, 26, 2500), rng.integers(0, 26, 3100)
    aw, bw = rng.integers(-5, 1000, 1200), rng.integers(-5, 1000, 900)
    for name, x in (("a26", a26), ("b26", b26), ("aw", aw), ("bw", bw)):
        (tmp_path / f"{name}.txt").write_text(" ".join(map(str, x)))
    script = tmp_path / "s.txt"
    script.write_text(
        "algorithm=hip_llcs device_type=gpu size_a=7330 size_b=9944 seed_a=888 seed_b=999\n"
        "algorithm=hip_llcs device_type=gpu size_a=9771 size_b=11302 seed_a=888 seed_b=999\n"
        f"algorithm=hip_llcs device_type=gpu file_a={tmp_path / 'a26.txt'} file_b={tmp_path / 'b26.txt'}\n"
        f"algorithm=hip_llcs device_type=gpu file_a={tmp_path / 'aw.txt'} file_b={tmp_path / 'bw.txt'}\n")
    out = tmp_path / "o.csv"
    r = subprocess.run([CLI, str(script), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = list(csv.DictReader(io.StringIO(out.read_text())))
    g = golden()
    assert int(rows[0]["hash"]) == g[(7330, 9944, 888, 999)]["llcs"]
    assert int(rows[1]["hash"]) == g[(9771, 11302, 888, 999)]["llcs"]
    assert int(rows[2]["hash"]) == oracle.llcs_dp(a26, b26)
    assert int(rows[3]["hash"]) == oracle.llcs_dp(aw, bw)
