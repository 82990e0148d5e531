import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-load-balance_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def orb():
    import orb_amd
    orb_amd.lib()
    if orb_amd.lib().orb_device_count() < 1:
        pytest.fail("liborb_hip.so loaded but no GPU is visible")
    return orb_amd


@pytest.fixture(scope="session")
def orb_host():
    """The HIP library for its host-only entry points (no GPU needed)."""
    import orb_amd
    orb_amd.lib()
    return orb_amd
