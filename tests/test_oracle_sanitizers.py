"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds oracle/sanitize_check.c with lcs_oracle.c
(gcc, -fsanitize=address,undefined, host code only) and the driver checks
every comb order, the strand read-outs and the steady ant on small seeded
inputs against independent answers (textbook DP, bit-parallel LCS, the
sticky product from its definition).  A sanitizer report or a mismatch
fails the test."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_clean_under_asan_and_ubsan():
    build = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "sanitize" in build.stderr and "cannot find" in build.stderr:
        pytest.skip("the compiler's sanitizer runtimes are not installed")
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    run = subprocess.run([os.path.join(ORACLE, "_san", "sanitize_check")], capture_output=True, text=True,
                         timeout=300, env=env)
    assert run.returncode == 0 and run.stdout.strip().endswith("ok"), (run.stdout[-2000:], run.stderr[-4000:])
