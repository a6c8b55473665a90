import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    from ysim_amd.hip import HipContext
    c = HipContext(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx_formal():
    import ctypes as C
    from ysim_amd import _lib
    from ysim_amd.hip import HipContext
    c = HipContext.__new__(HipContext)
    h = C.c_void_p()
    _lib.call("ysim_hip_ctx_create", 0, 1, C.byref(h))
    c.h = h
    yield c
    c.close()


@pytest.fixture(scope="session")
def oracle_threads():
    from oracle import oracle as O
    return O.set_threads(min(16, os.cpu_count() or 1))
