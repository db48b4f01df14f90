import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and libcraw.so")


@pytest.fixture(scope="session")
def gpu_available():
    from libcraw_amd import hip
    if hip.device_count() < 1:
        pytest.skip("no GPU in this environment")
    return True
