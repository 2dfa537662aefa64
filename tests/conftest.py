import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblassie.so")


@pytest.fixture(scope="session")
def gpu_lib():
    """liblassie with a visible HIP device; a GPU test must not pass without one."""
    from lassie_amd import _lib
    n = _lib.device_count()
    assert n >= 1, "gpu-marked test run without a visible HIP device"
    return _lib
