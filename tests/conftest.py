import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmoons_amd on cuda:0)")


@pytest.fixture(scope="session")
def gpu():
    """Initialises libmoons_amd on device 0 (GPU tests only)."""
    import moons_amd

    moons_amd.init(0)
    return moons_amd
