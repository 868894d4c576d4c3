import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ora():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def ooc():
    import oocdag_amd
    oocdag_amd.lib()
    return oocdag_amd


@pytest.fixture(scope="session")
def gpu(ooc):
    if ooc.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return ooc
