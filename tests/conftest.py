import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X / gfx950) and the built HIP library")


def _gpu_available():
    try:
        from cumat_amd import _capi
        import ctypes
        n = ctypes.c_int(0)
        _capi.lib().cm_device_count(ctypes.byref(n))
        return n.value > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly (no silent skip) when a gpu-marked test runs without a GPU
    or without the HIP library."""
    if not _gpu_available():
        pytest.fail("gpu test requested but no HIP device / library is available")
    import cumat_amd
    return cumat_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build(with_ref=False)
    return oracle
