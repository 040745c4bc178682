import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


@pytest.fixture(scope="session")
def hip():
    """The loaded HIP library bound to device 0; skips when no GPU is visible."""
    from pymoresane_amd import _lib
    if not _lib.gpu_available():
        pytest.skip("no GPU visible")
    return _lib.lib()
