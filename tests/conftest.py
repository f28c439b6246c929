import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libehip's C ABI)")


@pytest.fixture(scope="session")
def ehip():
    """libehip initialised on device 0.  GPU tests fail (not skip) when the
    library or the device is missing: they run only where a GPU is."""
    from elmerfem_amd import _lib
    return _lib.init(0)
