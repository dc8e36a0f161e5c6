import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


def pytest_runtest_setup(item):
    # name the case on stdout before any GPU work (helps when a box dies)
    if item.get_closest_marker("gpu") is not None:
        print(f"\n[gpu-test] {item.nodeid}", flush=True)


@pytest.fixture(scope="session")
def gpu():
    from gnss_gpu_amd import _lib

    n = _lib.device_count()
    if n < 1:
        raise RuntimeError("GPU test run without a visible GPU")
    return n
