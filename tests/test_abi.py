"""CPU: the C-ABI library loads and exports every symbol include/voxphoton.h declares."""
import os
import re

from tests.conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, 'include', 'voxphoton.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(vp_[a-z0-9_]+)\s*\(', txt)))


def test_header_symbols_exported():
    from voxphoton_amd import _lib
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(_lib.lib, s)]
    assert not missing, f'not exported: {missing}'


def test_binding_covers_header():
    from voxphoton_amd import _lib
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_no_device_calls_fail_cleanly():
    import ctypes
    from voxphoton_amd import _lib
    c = ctypes.c_int(-1)
    assert _lib.lib.vp_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0
    h = ctypes.c_void_p()
    if c.value == 0:
        rc = _lib.lib.vp_ctx_create(0, ctypes.byref(h))
        assert rc != 0 and _lib.lib.vp_last_error()
    assert _lib.lib.vp_op_apply(None, 0, None, None, 0, 1) == _lib.VP_ERR_INVALID


def test_recycle_width_rejects_a_missing_context():
    from voxphoton_amd import _lib
    assert _lib.lib.vp_ctx_set_recycle_width(None, 30) == _lib.VP_ERR_INVALID
    assert b'ctx' in _lib.lib.vp_last_error()
