import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu():
    """Skip only where no GPU device node exists; on a GPU machine any failure
    to load or use the library fails the test."""
    if not gpu_present():
        pytest.skip("no /dev/kfd: not a GPU machine")
    import linac_amd
    from linac_amd import _lib
    _lib.load()
    assert linac_amd.GPU_FOUND, "GPU device node present but liblinac_amd reports no usable GPU: " + _lib.last_error()
    return linac_amd


def load_small_cases():
    z = numpy.load(os.path.join(GOLDEN, "gf_small_cases.npz"), allow_pickle=False)
    out = []
    for k in range(int(z["count"][0])):
        out.append((int(z[f"p_{k}"][0]), z[f"A_{k}"], z[f"R_{k}"], z[f"piv_{k}"].tolist()))
    return out
