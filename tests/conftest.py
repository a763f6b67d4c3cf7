import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "openmp-offloading-intranode-artifact_amd")
for d in (PKG, os.path.join(REPO, "tests")):
    if d not in sys.path:
        sys.path.insert(0, d)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libh3d.so)")


@pytest.fixture(scope="session")
def h3d():
    import h3d_py
    h3d_py.lib()
    return h3d_py


@pytest.fixture(scope="session")
def oracle():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def gpu(h3d):
    """The device library with at least one GPU present (gpu tests only)."""
    if h3d.device_count() < 1:
        pytest.fail("no HIP device visible: gpu tests need an MI355X")
    return h3d
