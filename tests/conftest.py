import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ffi

    return oracle_ffi.Oracle()


@pytest.fixture(scope="session")
def dev():
    import rstsr_amd as rt

    if rt.device_count() < 1:
        pytest.fail("no GPU visible: gpu-marked tests must run on the MI355X box")
    return rt.DeviceHIP(0)
