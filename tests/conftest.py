import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when no device is visible: a GPU test that
    silently skipped would hide a missing native path."""
    from ducc_amd import _lib
    _lib.require_gpu()
    return True
