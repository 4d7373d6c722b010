import os
import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
if str(REPO) not in sys.path:
    sys.path.insert(0, str(REPO))
if str(REPO / "tests") not in sys.path:
    sys.path.insert(0, str(REPO / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def rglib():
    from raytracedgl1_amd import ffi
    return ffi.lib()


@pytest.fixture(scope="session")
def gpu(rglib):
    from raytracedgl1_amd import ffi
    n = ffi.device_count()
    assert n >= 1, "no HIP device visible to librtgl1_amd.so"
    return n


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.Oracle()
