import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gpu-spect-pet-sysmat_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsysmat_hip.so")


def gpu_available() -> bool:
    try:
        from sysmat import _lib
        return os.path.exists(_lib.LIB_PATH) and _lib.load().sm_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def lib():
    from sysmat import _lib
    if not gpu_available():
        pytest.fail("GPU test run without a visible HIP device or without libsysmat_hip.so")
    return _lib.load()
