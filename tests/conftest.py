import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libvmx.so")


@pytest.fixture(scope="session")
def gpu():
    """Skips unless a HIP device is visible; the library itself must load (no fallback)."""
    from vampire_amd import context
    if context.device_count() < 1:
        pytest.skip("no HIP device visible")
    return True
