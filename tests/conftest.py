import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")


@pytest.fixture(autouse=True)
def _precision():
    """Each test starts in complex64 unless it switches to complex128 itself."""
    import jaqsi_amd

    jaqsi_amd.enable_x64(False)
    yield
    jaqsi_amd.enable_x64(False)
