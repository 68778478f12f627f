import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdeft_amd.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")


@pytest.fixture(scope="session")
def gpu():
    """The product library on a HIP device; fails (never skips) when either is missing."""
    import deft_amd
    n = deft_amd.device_count()
    assert n > 0, "no HIP device visible"
    return deft_amd
