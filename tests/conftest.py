import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def lib():
    from medyan_amd import mechanics
    return mechanics.load_library()


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.Oracle()


@pytest.fixture(scope="session")
def gpu_ctx():
    from medyan_amd import mechanics
    if mechanics.load_library().mdy_device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on an MI355X")
    return mechanics.Context(0)
