import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "platform-independent-lcs_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def cpu_threads() -> int:
    """Host threads the CPU checker uses: this process's CPU share, at most
    16 (the host share of one GPU on the test machines)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="session")
def ctx():
    import lcs_amd
    c = lcs_amd.Context(0)
    yield c
    c.close()
