import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "proxy-geos-hc_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def phc():
    import phc as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def gpu(phc):
    if phc.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return phc


@pytest.fixture(scope="session")
def oracle():
    import oracle as m
    m.lib()
    return m
