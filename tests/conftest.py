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
def orc():
    import _oracle
    _oracle.lib()
    return _oracle


@pytest.fixture(scope="session")
def bvh():
    """The product package (HIP library); GPU tests only."""
    import bvhtest_amd
    L = bvhtest_amd.lib()
    if L.bvht_deviceCount() < 1:
        pytest.fail("no HIP device visible for a test marked gpu")
    return bvhtest_amd
