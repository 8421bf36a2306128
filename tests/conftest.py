import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.backend(0)


@pytest.fixture(scope="session")
def gpu():
    from horsepower_amd.backend import GpuBackend
    return GpuBackend()
