import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP path)")


@pytest.fixture(scope="session")
def M():
    """The product front end (map_amd.map over libmapx.so)."""
    from map_amd import map as M
    return M


@pytest.fixture(scope="session")
def G(M):
    """The product front end with the GPU set up."""
    M.setupDevices("", M.DEV_GPU, "")
    return M


@pytest.fixture()
def O():
    """The oracle evaluator (FP32-semantics mode unless a test switches it)."""
    from oracle import ref as O
    O.set_mode("f32")
    yield O
    O.set_mode("f32")
