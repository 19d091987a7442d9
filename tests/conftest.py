import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    """The CPU restatement (test infrastructure only)."""
    from ice_amd import native
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(path):
        pytest.fail("oracle/liboracle.so is missing: run __graft_entry__.build()")
    return native.bind(ctypes.CDLL(path), "orc_")


@pytest.fixture(scope="session")
def gpu():
    from ice_amd import native
    return native.load()
