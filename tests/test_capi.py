"""The C-ABI library (CPU): it loads, exports every entry point of include/bbgpu.h, and rejects invalid
arguments before any device call."""
import ctypes as C

import pytest

from bbarolo2_amd import _lib


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.declared_symbols()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib._SIGS) <= set(names)


def test_argument_errors_before_device_calls():
    L = _lib.lib()
    # NULL output handle
    rc = L.bb_model_create(None, None, None, None, 0, None, 0, 0, 1.0, None)
    assert rc != 0 and b"NULL" in L.bb_last_error()
    # invalid grid
    g = _lib.bb_grid()
    out = C.c_void_p()
    rc = L.bb_model_create(C.byref(g), None, None, None, 0, None, 0, 0, 1.0, C.byref(out))
    assert rc != 0 and b"nz" in L.bb_last_error() and not out.value
    # null model handles
    assert L.bb_galmod_analytic(None, None, _lib.bb_box(0, 0, 4, 4), None, 1, 1, None, 0, None, 0, None) != 0
    assert L.bb_azplan_create(None, None, 3, None) != 0
    assert L.bb_analytic_work_size(None, 4) < 0
    assert L.bb_conv2d_same(None, None, None, 1, 4, 4, None) != 0
    # residual: too small a work buffer is rejected
    assert L.bb_residual_work_size(3, 10000) > 0
    rc = L.bb_residual_gauss(C.c_void_p(8), C.c_void_p(8), 3, 10000, 1.0, 0.0, C.c_void_p(8), 1, C.c_void_p(8),
                             None, None)
    assert rc != 0 and b"work" in L.bb_last_error()
    with pytest.raises(_lib.BBError):
        _lib.check(rc)
