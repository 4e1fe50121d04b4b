import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "ref: needs the reference CPU build under oracle/_ref")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(autouse=True)
def _release_device_buffers():
    yield
    from tests import helpers
    helpers.release_all()


@pytest.fixture(scope="session")
def lib():
    from dualsphysics_amd import _native
    return _native.load()


@pytest.fixture(scope="session")
def ctx(lib):
    import ctypes as C
    h = C.c_void_p()
    rc = lib.dsph_ctx_create(0, C.byref(h))
    if rc:
        pytest.fail(f"dsph_ctx_create failed: {rc}")
    rc = lib.dsph_ctx_reserve(h, 1 << 21, 1 << 20)
    assert rc == 0, lib.dsph_last_error(h)
    yield h
    lib.dsph_ctx_destroy(h)
