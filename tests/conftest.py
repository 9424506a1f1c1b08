import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU-side check")


@pytest.fixture(scope="session")
def gpu():
    """The HIP library, loaded; fails (does not skip) when no device is present,
    so a GPU run can never pass on a silent fallback."""
    from novaphy_amd import capi
    lib = capi.load()
    n = capi.device_count()
    assert n > 0, "no HIP device visible to libnpx_hip"
    return lib
