import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


_GPU = None


def gpu_available():
    """True when libvoxphoton.so sees at least one GPU."""
    global _GPU
    if _GPU is None:
        import ctypes
        from voxphoton_amd import _lib
        c = ctypes.c_int(0)
        _GPU = _lib.lib.vp_device_count(ctypes.byref(c)) == 0 and c.value > 0
    return _GPU


@pytest.fixture(scope='session')
def vp_ctx():
    if not gpu_available():
        pytest.skip('no GPU')
    from voxphoton_amd import vie
    return vie.default_context()
