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
    rc = L.bb_model_create(None, None, None, None, 0, 0, 1.0, None)
    assert rc != 0 and b"NULL" in L.bb_last_error()
    # invalid grid
    g = _lib.bb_grid()
    out = C.c_void_p()
    rc = L.bb_model_create(C.byref(g), None, None, None, 0, 0, 1.0, C.byref(out))
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


def test_argument_errors_of_the_profile_mc_and_ring_entry_points():
    L = _lib.lib()
    box = _lib.bb_box(0, 0, 8, 8)
    v = C.c_void_p(8)
    # ring maps and the profile
    assert L.bb_ring_maps_deposit(None, None, box, None, 1, 1, None, None, 2, 0, 2, 1.0, None, 0, None, None) != 0
    assert L.bb_ring_maps_finish(None, None, 1, 4, 4, 1.0, 1, None, None) != 0
    d = _lib.bb_profile_desc(0, 10, 10, 0, 0.3, 1.0, 0.05, 1.0, 1.0)
    out = C.c_void_p()
    rc = L.bb_profile_create(C.byref(d), v, v, None, None, None, C.byref(out))
    assert rc != 0 and b"nr" in L.bb_last_error() and not out.value
    d = _lib.bb_profile_desc(4, 10, 11, 0, 0.3, 1.0, 0.05, 1.0, 1.0)
    assert L.bb_profile_create(C.byref(d), v, v, None, None, None, C.byref(out)) != 0
    d = _lib.bb_profile_desc(4, 10, 10, 0, 0.3, 1.0, 0.05, 1.0, 1.0)
    rc = L.bb_profile_create(C.byref(d), v, v, None, None, None, C.byref(out))
    assert rc != 0 and b"prior" in L.bb_last_error()
    assert L.bb_profile_work_size(None, 1) < 0
    assert L.bb_profile_solve(None, None, 1, None, 0, None, None) != 0
    assert L.bb_profile_set_dens(None, 1, 3, 0, None, None, 5, None, None, None) != 0
    # Monte-Carlo GALMOD
    assert L.bb_mc_work_size(None, box, 3, 1, 100) < 0
    assert L.bb_galmod_mc(None, box, None, 3, 1, 1, 7, 100, None, 0, None, None, None, 0, None) != 0
    # the 3DFIT ring residual
    assert L.bb_norm_work_size(1, 0, 4, 4, 0) < 0
    assert L.bb_norm_work_size(2, 4, 4, 3, 1) > L.bb_norm_work_size(2, 4, 4, 3, 0) > 0
    args = [v, 1, 4, 4, 3, v, v, 10, 10, 0, 0, v, v]
    tail = [None, None, None, 0, None, v, 100, v, None, None, None]
    assert L.bb_norm3dfit(*args, None, *tail) != 0 and b"par" in L.bb_last_error()
    p = _lib.bb_norm_par(0, 5, 0, 0, 2.0, 1.0, 1.0, 0.0)
    assert L.bb_norm3dfit(*args, C.byref(p), *tail) != 0 and b"ftype" in L.bb_last_error()
    p = _lib.bb_norm_par(0, 2, 0, 0, 2.0, 1.0, 1.0, 0.0)
    bad = list(args)
    bad[9] = 8                                   # box [8, 12) outside a 10-pixel cube
    assert L.bb_norm3dfit(*bad, C.byref(p), *tail) != 0 and b"outside" in L.bb_last_error()
    assert L.bb_norm3dfit(*args, C.byref(p), *tail[:-5], 1, *tail[-4:]) != 0 and b"work" in L.bb_last_error()
    # deposit events: both or neither
    assert L.bb_trace_deposit(v, None) != 0
    assert L.bb_trace_deposit(None, None) == 0


def test_argument_errors_of_the_mask_dilation():
    L = _lib.lib()
    p = C.c_void_p(64)
    assert L.bb_dilate_rmax() >= 32
    assert L.bb_dilate_mask(p, 70000, 8, 8, 2.0, 1, p, C.c_void_p(128), None) != 0 and b"sizes" in L.bb_last_error()
    assert L.bb_dilate_mask(p, 4, 8, 8, float("nan"), 1, p, C.c_void_p(128), None) != 0
    assert L.bb_dilate_mask(p, 4, 8, 8, 2.0, -1, p, C.c_void_p(128), None) != 0
    rc = L.bb_dilate_mask(p, 4, 8, 8, L.bb_dilate_rmax() + 1.5, 1, C.c_void_p(96), C.c_void_p(128), None)
    assert rc != 0 and b"exceeds" in L.bb_last_error()
    assert L.bb_dilate_mask(p, 4, 8, 8, 2.0, 1, None, C.c_void_p(128), None) != 0 and b"tmp" in L.bb_last_error()
    assert L.bb_dilate_mask(p, 4, 8, 8, 2.0, 1, C.c_void_p(96), p, None) != 0 and b"alias" in L.bb_last_error()
    assert L.bb_dilate_mask(None, 4, 8, 8, 2.0, 1, C.c_void_p(96), C.c_void_p(128), None) != 0
    assert L.bb_dilate_mask(None, 0, 8, 8, 2.0, 1, None, None, None) == 0          # (an empty cube: nothing to do)


def test_argument_errors_of_the_likelihood_context():
    L = _lib.lib()
    out = C.c_void_p()
    d = _lib.bb_like_desc()
    # NULL handles and arguments are rejected before any device call
    assert L.bb_ctx_create(None, C.byref(d), None, None, None, None, None, None, None, None, None, None) != 0
    assert L.bb_ctx_create(None, C.byref(d), None, None, None, None, None, None, None, None, None,
                           C.byref(out)) != 0 and b"NULL" in L.bb_last_error() and not out.value
    assert L.bb_loglike_launch(None, None, 1, None) != 0
    assert L.bb_loglike_collect(None, None, None) != 0
    assert L.bb_loglike_batch(None, None, 1, None, None, None) != 0
    assert L.bb_ctx_tables(None, None, 1, None, None) != 0
    assert L.bb_ctx_info(None, None, None, None, None, None) != 0
    assert L.bb_ctx_set_halves(None, 0) != 0 and L.bb_ctx_set_graphs(None, 1) != 0
    assert L.bb_ctx_graph_replays(None) < 0
    assert L.bb_ctx_trace(None, None) != 0
    assert L.bb_ctx_destroy(None) == 0
    assert L.bb_model_set_bin(None, 4) != 0
