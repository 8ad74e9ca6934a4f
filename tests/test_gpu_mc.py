"""GPU parity (MI355X) of the counter-based Monte-Carlo GALMOD: the integer count cube of the device kernels is
bit-exact to the CPU restatement (oracle/mc.c) for the same draws; the float model equals count x q x cd2i; a
box is the crop of the full model; batch rows are independent."""
import numpy as np
import pytest

import oracle
from bbarolo2_amd.galmod import GalmodPar, model_rings
from bbarolo2_amd.header import emptyheader
from conftest import small_rings

pytestmark = pytest.mark.gpu


def _dm(h, **kw):
    from bbarolo2_amd.engine import DiskModel
    return DiskModel(h, **kw)


def _gpu(dm, sr, blo, bhi, seed):
    import torch
    B = sr.shape[0]
    by, bx = bhi[1] - blo[1], bhi[0] - blo[0]
    counts = torch.empty((B, by, bx, dm.setup.nsubs), dtype=torch.int64, device="cuda")
    m = dm.array_batch_mc(sr, blo, bhi, convolve=False, seed=seed, counts=counts)
    return counts.cpu().numpy().astype(np.uint64), m.double().cpu().numpy()


CASES = {
    "thin": dict(z0=0.),
    "gauss": dict(z0=10.),
    "sech2": dict(z0=10., ltype=2),
    "exp": dict(z0=10., ltype=3),
    "lorentz": dict(z0=10., ltype=4),
    "box": dict(z0=10., ltype=5),
    "lag": dict(z0=15., dvdz=0.5, zcyl=5.),
    "warp": dict(z0=5., inc=np.linspace(50., 75., 8), phi=np.linspace(30., 60., 8), vrad=20., vvert=5.),
}


@pytest.mark.parametrize("name", list(CASES))
def test_mc_counts_bit_exact(small_header, name):
    kw = dict(CASES[name])
    ltype = kw.pop("ltype", 1)
    dm = _dm(small_header, par=GalmodPar(LTYPE=ltype, GALMODTYPE="MC"))
    r = small_rings(dens=np.linspace(2., 0.5, 8), vrot=np.linspace(80., 160., 8), **kw)
    sr, nv = dm.setup.subrings(model_rings(r, dm.ringmode, dm.setup.radsep))
    blo, bhi = (0, 0), (64, 64)
    cref, mref, q = oracle.galmod_mc(dm.setup, sr, nv, blo, bhi, 4321)
    cgot, mgot = _gpu(dm, sr[None], blo, bhi, 4321)
    ndiff = int(np.count_nonzero(cgot[0] != cref))
    print(f"[mc {name}] sub-cloud units {int(cref.sum())}, voxels differing {ndiff}", flush=True)
    assert ndiff == 0
    assert np.abs(mgot[0] - mref).max() <= 1e-6 * mref.max()


def test_mc_box_freq_two_lines():
    """A box away from the origin, half edges, a frequency axis with two lines (the non-linear channel path)."""
    h = emptyheader([80, 72, 48], [-8 / 3600., 8 / 3600., 1.0e5], beam=30 / 3600.,
                    ctypes=["RA---SIN", "DEC--SIN", "FREQ"], cunits=["deg", "deg", "Hz"],
                    crvals=[60, 0, 1.4204e9], FREQ0=1.4204057518e9)
    par = GalmodPar(RADMAXA=90., RADMAXR=140., RESTFREQ=[1.4204057518e9, 1.42e9], RELINT=[1.0, 0.4],
                    GALMODTYPE="MC", NV=20)
    dm = _dm(h, par=par)
    r = small_rings(xpos=39., ypos=35., z0=8., vsys=10.)
    sr, nv = dm.setup.subrings(model_rings(r, dm.ringmode, dm.setup.radsep))
    blo, bhi = (7, 5), (70, 66)
    cref, mref, q = oracle.galmod_mc(dm.setup, sr, nv, blo, bhi, 99)
    cgot, mgot = _gpu(dm, sr[None], blo, bhi, 99)
    assert cref.sum() > 0 and np.array_equal(cgot[0], cref)


def test_mc_batch_and_crop(small_header):
    """Rows of a batch are independent and a box model is the crop of the full model, on the device."""
    dm = _dm(small_header, par=GalmodPar(GALMODTYPE="MC"))
    tabs = [dm.setup.subrings(model_rings(small_rings(vrot=v, z0=8.), dm.ringmode, dm.setup.radsep))[0]
            for v in (100., 150., 200.)]
    sr = np.stack(tabs)
    cfull, _ = _gpu(dm, sr, (0, 0), (64, 64), 7)
    for b in range(3):
        one, _ = _gpu(dm, sr[b:b + 1], (0, 0), (64, 64), 7)
        assert np.array_equal(one[0], cfull[b])
    cbox, _ = _gpu(dm, sr, (5, 8), (60, 58), 7)
    assert np.array_equal(cbox, cfull[:, 8:58, 5:60])


def test_likelihood_mc_with_profile():
    """GpuLikelihood with the MC model, INC and PA free and DENS not fitted: the log-likelihood of every vector
    agrees with the oracle pipeline (ANALYTIC ring maps -> profile -> float32 densities -> MC model -> beam ->
    residual)."""
    import math
    from oracle import profile as op
    from bbarolo2_amd.galmod import nzq_for, PHI_TAB
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars, float32_rings
    h = emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., bunit="JY/BEAM", obj="MOCK")
    dm = _dm(h, par=GalmodPar(GALMODTYPE="MC", ISEED=11))
    st, kern, sf = dm.setup, dm.cfield, dm.scalefactor
    r = small_rings(8, dens=2. * np.exp(-(np.arange(8) * 20. + 10.) / 50.), z0=5.)
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    m = oracle.conv_same(oracle.galmod_analytic(st, sr, (0, 0), (64, 64), nzq_for(sr), PHI_TAB), kern, sf)
    sig = m.max() / 100.
    data = (m + np.random.default_rng(3).normal(0., sig, m.shape)).transpose(2, 0, 1).astype(np.float32)
    free = FreePars([("vrot", "single"), ("inc", "single"), ("phi", "single")], r.nr)
    like = GpuLikelihood(dm, data, r, free, sigma=sig, likefactor=1., densprof_smooth=0.3)
    theta = np.array([[150., 60., 45.], [140., 55., 50.]])
    ll = like.loglike(theta)
    nz, ny, nx = data.shape
    ymap = data.astype(np.float64).sum(axis=0)
    smap = np.full((ny, nx), sig * math.sqrt(nz))
    blo, bhi = like.blo, like.bhi
    for i, th in enumerate(theta):
        rr = float32_rings(r)
        for k, v in free.values(th[None]).items():
            setattr(rr, k, v[0])
        d = op.deconvolved(st, kern, sf, rr, ((0, 0), (nx, ny)), ymap, smap, np.ones((ny, nx), bool), smooth=0.3,
                           beam_arcsec=30.)
        rr.dens = d.astype(np.float32).astype(np.float64)
        sr, nv = st.subrings(model_rings(rr, "CENTRED", st.radsep))
        _, mm, _ = oracle.galmod_mc(st, sr, nv, blo, bhi, 11)
        mm = oracle.conv_same(mm, kern, sf)
        dbox = data[:, blo[1]:bhi[1], blo[0]:bhi[0]]
        S = oracle.sumsq_residuals(np.ascontiguousarray(mm.transpose(2, 0, 1)), dbox, np.ones(dbox.shape, bool), 2,
                                   np.ones(dbox.shape[1:], bool)) + like.ss_out
        ref = -(0.5 * S / sig ** 2 + 0.5 * like.nvox * math.log(2 * math.pi * sig ** 2))
        print(f"MC loglike {i}: gpu {ll[i]:.6f} oracle {ref:.6f}", flush=True)
        assert abs(ll[i] - ref) <= 1e-5 * 0.5 * S / sig ** 2
