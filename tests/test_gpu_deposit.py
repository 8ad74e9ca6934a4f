"""GPU (MI355X): the two kernels of the ANALYTIC deposit give the same model, bit for bit.

The lane-per-sample deposit (bb_model_set_deposit 0, the default) and the wavefront deposit (variant 1: one lane per
sample walks the vertical column, then the wavefront's lanes along the channels add each sample's line profile to
its runs) add the same FP32 values as integers, so their models must be equal to the last bit. Cases: every
vertical layer, a height-dependent rotation (the lag path), a warp with radial and vertical motions, face-on, a
box away from the origin with half edges, a frequency axis with two lines, long vertical columns (more than 16
samples, fine vertical steps), line windows of more than 64 channels, per-vector model boxes, and tile shapes. Both
are compared with the FP64 oracle elsewhere (test_gpu_parity.py, test_gpu_fullsize.py). One exception to the bit
for bit rule: a long column with more than 16 runs in one tile, whose runs past the 16th the lane-per-sample kernel
adds through its FP64-argument line routine (acc_line): the lane-per-sample models are made on 4x4-pixel tiles,
where no column has more than 16 runs.
"""
import numpy as np
import pytest

from bbarolo2_amd.galmod import GalmodPar, model_rings, nzq_for
from bbarolo2_amd.header import emptyheader
from conftest import small_rings

pytestmark = pytest.mark.gpu

CASES = {
    "thin": dict(z0=0.),
    "gauss": dict(z0=10.),
    "sech2": dict(z0=10., ltype=2),
    "exp": dict(z0=10., ltype=3),
    "lorentz": dict(z0=10., ltype=4),
    "box": dict(z0=10., ltype=5),
    "lag": dict(z0=15., dvdz=0.5, zcyl=5.),
    "warp": dict(z0=5., inc=np.linspace(50., 75., 8), phi=np.linspace(30., 60., 8), vrad=20., vvert=5.),
    "faceon": dict(z0=5., inc=0.),
}


def _both(dm, sr, blo, bhi, vbox=None):
    """The lane-per-sample model on 4x4-pixel tiles (at most 16 runs of a column in a tile: every run through its
    register path) and the wavefront model on the model's own tile."""
    import ctypes as C
    import torch
    from bbarolo2_amd import _lib
    h = dm.handle(nzq_for(sr))
    tx, ty = C.c_int(), C.c_int()
    _lib.check(_lib.lib().bb_model_get_tile(h, C.byref(tx), C.byref(ty)))
    out = []
    for variant, tile in ((0, (4, 4)), (1, (tx.value, ty.value))):
        _lib.check(_lib.lib().bb_model_set_tile(h, *tile))
        dm.set_deposit(variant)
        m = dm.array_batch(sr, blo, bhi, convolve=False, vbox=vbox)
        torch.cuda.synchronize()
        out.append(m.cpu().numpy())
    return out


def _sr(dm, rings):
    sr, _ = dm.setup.subrings(model_rings(rings, dm.ringmode, dm.setup.radsep))
    return sr


@pytest.mark.parametrize("name", list(CASES))
def test_wave_deposit_equals_lane_deposit(small_header, name):
    from bbarolo2_amd.engine import DiskModel
    kw = dict(CASES[name])
    ltype = kw.pop("ltype", 1)
    dm = DiskModel(small_header, par=GalmodPar(LTYPE=ltype), galmodtype="ANALYTIC")
    r = small_rings(dens=np.linspace(2., 0.5, 8), vrot=np.linspace(80., 160., 8), **kw)
    sr = _sr(dm, r)
    a, b = _both(dm, np.stack([sr, sr * 1.0]), (0, 0), (64, 64))
    print(f"[deposit {name}] nsub={sr.shape[1]} max {a.max():.4e} max|diff| {np.abs(a - b).max():.3e}", flush=True)
    assert a.max() > 0
    assert np.array_equal(a, b)


def test_wave_deposit_frequency_axis_two_lines():
    from bbarolo2_amd.engine import DiskModel
    h = emptyheader([80, 72, 48], [-8 / 3600., 8 / 3600., 1.0e5], beam=30 / 3600.,
                    ctypes=["RA---SIN", "DEC--SIN", "FREQ"], cunits=["deg", "deg", "Hz"],
                    crvals=[60, 0, 1.4204e9], FREQ0=1.4204057518e9)
    par = GalmodPar(RADMAXA=90., RADMAXR=140., RESTFREQ=[1.4204057518e9, 1.42e9], RELINT=[1.0, 0.4])
    dm = DiskModel(h, par=par, galmodtype="ANALYTIC")
    sr = _sr(dm, small_rings(xpos=39., ypos=35., z0=8., vsys=10.))
    a, b = _both(dm, sr[None], (7, 5), (70, 66))
    assert a.max() > 0
    assert np.array_equal(a, b)


def test_wave_deposit_long_columns_and_tiles():
    """Fine vertical steps (no beam: 0.25-pixel steps) on a thick layer, and an edge-on thick disk, give columns of
    more than 16 samples; every tile shape gives the same wavefront model, bit for bit."""
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.engine import DiskModel
    h = emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=0., obj="NOBEAM")
    for kw in (dict(z0=20., inc=70.), dict(z0=30., inc=88.)):
        dm = DiskModel(h, par=GalmodPar(LTYPE=2, SM=False), galmodtype="ANALYTIC")
        sr = _sr(dm, small_rings(dens=np.linspace(2., 0.5, 8), **kw))
        a, b = _both(dm, sr[None], (0, 0), (64, 64))
        print(f"[long columns {kw}] max {a.max():.4e} max|diff| {np.abs(a - b).max():.3e}", flush=True)
        assert a.max() > 0
        assert np.array_equal(a, b)
        hnd = dm.handle(nzq_for(sr))
        for tx, ty in ((8, 8), (16, 4), (4, 16), (8, 2)):
            _lib.check(_lib.lib().bb_model_set_tile(hnd, tx, ty))
            dm.set_deposit(1)
            m = dm.array_batch(sr[None], (0, 0), (64, 64), convolve=False)
            torch.cuda.synchronize()
            assert np.array_equal(m.cpu().numpy(), a), (kw, tx, ty)


def test_wave_deposit_wide_lines():
    """Line windows of up to 185 channels (three blocks of 64 edges) on a 128-channel axis."""
    from bbarolo2_amd.engine import DiskModel
    h = emptyheader([64, 64, 128], [-10. / 3600., 10. / 3600., 5.], beam=30. / 3600., obj="WIDE")
    dm = DiskModel(h, galmodtype="ANALYTIC")
    sr = _sr(dm, small_rings(z0=10., vdisp=np.linspace(20., 90., 8), vrot=np.linspace(50., 120., 8)))
    a, b = _both(dm, sr[None], (0, 0), (64, 64))
    assert a.max() > 0
    assert np.array_equal(a, b)


def test_wave_deposit_vector_boxes(small_header):
    """Per-vector model boxes (free centre): each vector's deposit clipped to its own box."""
    from bbarolo2_amd.engine import DiskModel
    dm = DiskModel(small_header, galmodtype="ANALYTIC")
    srs = np.stack([_sr(dm, small_rings(z0=8., xpos=x, ypos=y)) for x, y in ((31.5, 31.5), (28., 34.), (35., 30.))])
    vbox = np.array([[0, 0, 64, 64], [4, 2, 60, 58], [10, 6, 64, 50]], np.int32)
    a, b = _both(dm, srs, (0, 0), (64, 64), vbox=vbox)
    assert a.max() > 0
    assert np.array_equal(a, b)
