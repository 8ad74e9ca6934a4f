"""GPU parity (MI355X): the HIP kernels through the C-ABI against the FP64 oracle on the same inputs.

Tolerances (FP32 arithmetic on the device, FP64 in the oracle):
  models (before and after the beam)   max |gpu - oracle| <= 1e-5 x max |oracle|
  residual sums and log-likelihoods    |gpu - oracle| <= 1e-5 x |oracle|
"""
import math

import numpy as np
import pytest

import oracle
from bbarolo2_amd.galmod import GalmodSetup, GalmodPar, model_rings, nzq_for, PHI_TAB, model_box
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.rings import Rings
from conftest import small_rings

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _dm(h, **kw):
    from bbarolo2_amd.engine import DiskModel
    return DiskModel(h, **kw)


def _oracle_model(dm, rings, blo, bhi, final=True):
    st = dm.setup
    if final:
        rings = model_rings(rings, dm.ringmode, st.radsep)
    sr, _ = st.subrings(rings)
    return oracle.galmod_analytic(st, sr, blo, bhi, nzq_for(sr), PHI_TAB), sr


def _gpu_model(dm, sr, blo, bhi, convolve=False):
    return dm.array_batch(sr[None], blo, bhi, convolve=convolve)[0].double().cpu().numpy()


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


@pytest.mark.parametrize("name", list(CASES))
def test_analytic_deposit(small_header, name):
    kw = dict(CASES[name])
    ltype = kw.pop("ltype", 1)
    dm = _dm(small_header, par=GalmodPar(LTYPE=ltype), galmodtype="ANALYTIC")
    r = small_rings(dens=np.linspace(2., 0.5, 8), vrot=np.linspace(80., 160., 8), **kw)
    blo, bhi = (0, 0), (64, 64)
    ref, sr = _oracle_model(dm, r, blo, bhi)
    got = _gpu_model(dm, sr, blo, bhi)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"[analytic {name}] nsub={sr.shape[1]} max rel err {err:.2e}", flush=True)
    assert err <= TOL


def test_analytic_box_and_edges():
    """A box away from the origin, RADMAXA/RADMAXR half edges, a frequency axis with two lines."""
    h = emptyheader([80, 72, 48], [-8 / 3600., 8 / 3600., 1.0e5], beam=30 / 3600.,
                    ctypes=["RA---SIN", "DEC--SIN", "FREQ"], cunits=["deg", "deg", "Hz"],
                    crvals=[60, 0, 1.4204e9], FREQ0=1.4204057518e9)
    par = GalmodPar(RADMAXA=90., RADMAXR=140., RESTFREQ=[1.4204057518e9, 1.42e9], RELINT=[1.0, 0.4])
    dm = _dm(h, par=par, galmodtype="ANALYTIC")
    r = small_rings(xpos=39., ypos=35., z0=8., vsys=10.)
    blo, bhi = (7, 5), (70, 66)
    ref, sr = _oracle_model(dm, r, blo, bhi)
    got = _gpu_model(dm, sr, blo, bhi)
    assert np.abs(got - ref).max() <= TOL * np.abs(ref).max()


@pytest.mark.parametrize("beam", [(30., 30., 0.), (40., 30., 30.)])
def test_conv2d_same(beam):
    h = emptyheader([64, 64, 40], [-10 / 3600., 10 / 3600., 10.], beam=tuple(b / 3600. if i < 2 else b
                                                                              for i, b in enumerate(beam)))
    dm = _dm(h, galmodtype="ANALYTIC")
    r = small_rings(z0=10.)
    blo, bhi = (2, 3), (61, 60)
    ref, sr = _oracle_model(dm, r, blo, bhi)
    refc = oracle.conv_same(ref, dm.cfield, dm.scalefactor)
    got = _gpu_model(dm, sr, blo, bhi, convolve=True)
    err = np.abs(got - refc).max() / refc.max()
    print(f"[conv {beam}] max rel err {err:.2e}", flush=True)
    assert err <= TOL


def test_residual_sums():
    import torch
    from bbarolo2_amd import _lib
    rng = np.random.default_rng(5)
    B, ny, nx, nz = 3, 37, 41, 23
    model = rng.random((B, ny, nx, nz)).astype(np.float32)
    data = (rng.random((nz, ny, nx)) + 0.2).astype(np.float32)
    data[rng.random(data.shape) < 0.05] = 0.
    data[rng.random(data.shape) < 0.02] = np.nan
    mask = rng.random(data.shape) < 0.7
    region = rng.random((ny, nx)) < 0.9
    for maskmode in (0, 1, 2):
        from bbarolo2_amd.likelihood import GpuLikelihood  # noqa: F401 (same data preparation)
        blank = ~np.isfinite(data) | (data == 0)
        if maskmode == 2:
            use = mask & ~blank & region[None]
            dv = np.where(use, data, np.nan)
        elif maskmode == 1:
            use = ~blank & region[None]
            dv = np.where(use, np.where(mask, data, 0.), np.nan)
        else:
            use = ~blank & region[None]
            dv = np.where(use, data, np.nan)
        dvalid = torch.from_numpy(np.ascontiguousarray(dv.transpose(1, 2, 0), np.float32)).cuda()
        dmod = torch.from_numpy(model).cuda()
        nvox = ny * nx * nz
        for eps2 in (0.0, 0.04):
            wl = int(_lib.lib().bb_residual_work_size(B, nvox))
            work = torch.empty(wl, dtype=torch.float64, device="cuda")
            ss = torch.empty(B, dtype=torch.float64, device="cuda")
            lg = torch.empty(B, dtype=torch.float64, device="cuda")
            _lib.check(_lib.lib().bb_residual_gauss(dmod.data_ptr(), dvalid.data_ptr(), B, nvox, 0.3, eps2,
                                                    work.data_ptr(), wl, ss.data_ptr(), lg.data_ptr(), None))
            torch.cuda.synchronize()
            for b in range(B):
                mz = model[b].transpose(2, 0, 1).astype(np.float64)
                if eps2 == 0:
                    ref = oracle.sumsq_residuals(mz, data, mask, maskmode, region)
                    assert abs(ss[b].item() - ref) <= TOL * ref
                else:
                    ref, rlg = oracle.residuals_fracerr(mz, data, mask, maskmode, region, 0.3, eps2)
                    assert abs(ss[b].item() - ref) <= TOL * ref and abs(lg[b].item() - rlg) <= TOL * abs(rlg)


def test_batch_rows_independent(small_header):
    """Row b of a batch equals the model of table b alone (no cross-talk between parameter vectors)."""
    dm = _dm(small_header, galmodtype="ANALYTIC")
    tabs = []
    for v in (100., 150., 200.):
        sr, _ = dm.subrings(small_rings(vrot=v, z0=8.))
        tabs.append(sr)
    sr = np.stack(tabs)
    blo, bhi = (0, 0), (64, 64)
    got = dm.array_batch(sr, blo, bhi, convolve=True).cpu().numpy()
    for b in range(3):
        one = dm.array_batch(sr[b:b + 1], blo, bhi, convolve=True)[0].cpu().numpy()
        assert np.abs(got[b] - one).max() <= 1e-6 * np.abs(one).max()


def test_conv_residual_fused(small_header):
    """bb_conv_residual_gauss (the convolution and the residual in one pass) against bb_conv2d_same followed by
    bb_residual_gauss (each pinned to the oracle above): the same FP32 terms summed in FP64 in another order, so
    within 1e-12 relative; both variants of the residual, blanks, and voxels under the 1e-12 cut."""
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.engine import DiskModel
    L = _lib.lib()
    dm = DiskModel(small_header, galmodtype="ANALYTIC")
    h = dm.handle()
    rng = np.random.default_rng(11)
    B, by, bx, nz = 3, 37, 41, small_header.DimAx(2)
    dep = rng.random((B, by, bx, nz)).astype(np.float32)
    dep[rng.random(dep.shape) < 0.5] = 0.            # (zeros around which the convolution falls under 1e-12)
    dv = (rng.random((by, bx, nz)) + 0.1).astype(np.float32)
    dv[rng.random(dv.shape) < 0.1] = np.nan
    ddep, ddv = torch.from_numpy(dep).cuda(), torch.from_numpy(dv).cuda()
    mod = torch.empty_like(ddep)
    _lib.check(L.bb_conv2d_same(h, ddep.data_ptr(), mod.data_ptr(), B, by, bx, None))
    nvox = by * bx * nz
    for eps2 in (0.0, 0.04):
        s2 = 0.3
        wl = int(L.bb_residual_work_size(B, nvox))
        work = torch.empty(wl, dtype=torch.float64, device="cuda")
        ss0, lg0 = (torch.empty(B, dtype=torch.float64, device="cuda") for _ in range(2))
        _lib.check(L.bb_residual_gauss(mod.data_ptr(), ddv.data_ptr(), B, nvox, s2, eps2, work.data_ptr(), wl,
                                       ss0.data_ptr(), lg0.data_ptr(), None))
        cwl = int(L.bb_conv_residual_work_size(h, B, by, bx))
        cwork = torch.empty(cwl, dtype=torch.float64, device="cuda")
        ss1, lg1 = (torch.empty(B, dtype=torch.float64, device="cuda") for _ in range(2))
        _lib.check(L.bb_conv_residual_gauss(h, ddep.data_ptr(), ddv.data_ptr(), B, by, bx, s2, eps2,
                                            cwork.data_ptr(), cwl, ss1.data_ptr(), lg1.data_ptr(), None))
        a, b = ss0.cpu().numpy(), ss1.cpu().numpy()
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a)), (eps2, a, b)
        if eps2 > 0:
            a, b = lg0.cpu().numpy(), lg1.cpu().numpy()
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a)), (a, b)
    # too small a work buffer is rejected before any launch
    assert L.bb_conv_residual_gauss(h, ddep.data_ptr(), ddv.data_ptr(), B, by, bx, 0.3, 0.0, cwork.data_ptr(), 1,
                                    ss1.data_ptr(), None, None) != 0
