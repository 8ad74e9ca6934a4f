"""GPU (MI355X): the likelihood context (bb_ctx_create / bb_loglike_batch, called through GpuLikelihood).

  * the sub-ring tables it builds on the device from theta equal the host's (SubringPlan.tables, the reference's
    GalmodSetup.subrings per batch) bit for bit: a thin disk with the profile of every vector (and the tables of its
    exterior annulus), a thick disk with an edge of the disk, edges of its halves and NODES rings;
  * with XPOS, YPOS, Z0, VSYS, INC and PA free (each vector in its own model box, the profile of every vector) the
    log-likelihoods agree with the FP64 oracle pipeline, which builds each vector's model in its own box and adds the
    data outside it from a summed-area table; the same with SIGMA and FRACERR free; the same with the MC model;
  * a vector with a negative positive-only value or SIGMA <= 0 gets -inf without disturbing the others.
Tolerance: 1e-5 of the residual term (FP32 models against FP64)."""
import math

import numpy as np
import pytest

import oracle
from oracle import pipeline
from bbarolo2_amd.galmod import GalmodPar, model_rings, nzq_for, PHI_TAB
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.rings import Rings

pytestmark = pytest.mark.gpu


def _disk():
    h = emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., bunit="JY/BEAM", obj="MOCK")
    r = Rings().setRings(radii=np.arange(8) * 20. + 10., xpos=31.5, ypos=32.25, vsys=5., vrot=150., vdisp=10., z0=5.,
                         dens=2. * np.exp(-(np.arange(8) * 20. + 10.) / 50.), inc=60., phi=45.)
    return h, r


def _data(h, r, par, noise=100., seed=3):
    """Noisy data of the rings (oracle model, ANALYTIC, beam) and the oracle's set-up pieces."""
    from bbarolo2_amd.galmod import GalmodSetup, zstep_for
    from bbarolo2_amd.smooth import beam_kernel
    st = GalmodSetup(h, par)
    kern, _ = beam_kernel(h)
    st.zstep = zstep_for(h, st, True)
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    m = oracle.conv_same(oracle.galmod_analytic(st, sr, (0, 0), (h.DimAx(0), h.DimAx(1)), nzq_for(sr), PHI_TAB),
                         kern, 1.0)
    sig = m.max() / noise
    data = (m + np.random.default_rng(seed).normal(0., sig, m.shape)).transpose(2, 0, 1).astype(np.float32)
    return data, sig


def _compare(like, orc, theta, tag):
    ll = like.loglike(theta)
    for i, th in enumerate(theta):
        ref = orc.loglike(th)
        if not np.isfinite(ref):
            assert ll[i] == ref, (tag, i)
            continue
        tol = 1e-5 * like.likefactor * orc.last_resid
        print(f"[{tag} {i}] box {orc.last_box} gpu {ll[i]:.8f} oracle {ref:.8f} |d| {abs(ll[i] - ref):.3e}",
              flush=True)
        assert abs(ll[i] - ref) <= tol, (tag, i)
    return ll


def test_device_tables_profile_per_vector():
    from bbarolo2_amd.bayes import synthetic_fit
    fit = synthetic_fit(5, chunk=16)
    like = fit["like"]
    assert like.update_prof
    lo, hi = fit["lo"], fit["hi"]
    theta = lo + np.random.default_rng(1).random((16, len(lo))) * (hi - lo)
    sr_d, ext_d = like.device_tables(theta)
    sr_h, vals = like.tables(theta)
    rows = [r for r in range(16) if r != 5]          # (the density row comes from the profile)
    assert np.array_equal(sr_d[:, rows], sr_h[:, rows])
    P = like.profile
    ext_h = P.plan_ext.tables({k: np.asarray(v)[:, -1:] for k, v in vals.items() if k in
                               ("xpos", "ypos", "vsys", "vrot", "vdisp", "z0", "inc", "phi")})[0]
    ext_h = np.broadcast_to(ext_h, ext_d.shape)
    assert np.array_equal(ext_d, ext_h)


def test_device_tables_thick_edges_nodes():
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    r.fthick = np.full(r.nr, 0.3)
    r.z0thick = np.full(r.nr, 12.)
    r.vdispthick = np.full(r.nr, 25.)
    r.edge, r.edgewidth = 130., 0.
    r.edgeapp, r.edgerec = 120., 140.
    dm = DiskModel(h, par=GalmodPar(GALMODTYPE="ANALYTIC", RINGMODE="NODES"))
    data, sig = _data(h, Rings().setRings(radii=r.radii, xpos=31.5, ypos=32.25, vrot=150., vdisp=10., dens=1.,
                                          inc=60., phi=45.), GalmodPar())
    free = FreePars([("vrot", "ring"), ("vdisp", "single"), ("dens", "ring"), ("z0", "single"), ("fthick", "single"),
                     ("vsys", "single"), ("z0thick", "single"), ("phi", "ring")], r.nr)
    like = GpuLikelihood(dm, data, r, free, sigma=sig, chunk=8)
    assert like.plan.thick and not like.useNorm
    rng = np.random.default_rng(2)
    theta = np.column_stack([rng.uniform(50, 250, (8, r.nr)), rng.uniform(1, 30, (8, 1)),
                             rng.uniform(0.1, 3, (8, r.nr)), rng.uniform(0, 15, (8, 1)), rng.uniform(0, 1, (8, 1)),
                             rng.uniform(-20, 20, (8, 1)), rng.uniform(-1, 30, (8, 1)), rng.uniform(30, 60, (8, r.nr))])
    sr_d, _ = like.device_tables(theta)
    sr_h, _ = like.tables(theta)
    assert sr_d.shape == sr_h.shape
    bad = np.argwhere(sr_d != sr_h)
    assert bad.size == 0, bad[:10]


def test_free_centre_height_vs_oracle():
    """XPOS, YPOS, Z0, VSYS, INC, PA free: each vector in its own model box within the whole-cube likelihood box."""
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    data, sig = _data(h, r, par)
    spec = [("vrot", "single"), ("vdisp", "single"), ("xpos", "single"), ("ypos", "single"), ("z0", "single"),
            ("vsys", "single"), ("inc", "single"), ("phi", "single")]
    region = np.ones((64, 64), bool)
    box = ((0, 0), (64, 64))
    dm = DiskModel(h, par=par)
    like = GpuLikelihood(dm, data, r, FreePars(spec, r.nr), region=region, likebox=box, sigma=sig, likefactor=0.5,
                         chunk=16)
    assert like.box_free and like.update_prof
    orc = pipeline.Likelihood(h, par, r, spec, data, None, region, box, sig, 0.5)
    theta = np.array([[150., 10., 31.5, 32.25, 5., 5., 60., 45.],
                      [140., 12., 27.2, 35.7, 0., 0., 55., 50.],
                      [160., 8., 36.9, 29.1, 9.5, -10., 65., 40.],
                      [150., 10., 20.4, 41.8, 3., 15., 62., 47.],
                      [150., 10., 31.5, 32.25, 5., 5., 60., 45.],
                      [150., -1., 31.5, 32.25, 5., 5., 60., 45.]])          # (negative VDISP: -inf)
    ll = _compare(like, orc, theta, "free centre")
    assert np.isneginf(ll[-1]) and np.all(np.isfinite(ll[:-1])) and ll[0] == ll[4]


def test_free_sigma_fracerr_vs_oracle():
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    data, sig = _data(h, r, par)
    spec = [("vrot", "single"), ("vdisp", "single"), ("inc", "single"), ("sigma", "single"), ("fracerr", "single")]
    region = np.ones((64, 64), bool)
    box = ((0, 0), (64, 64))
    like = GpuLikelihood(DiskModel(h, par=par), data, r, FreePars(spec, r.nr), region=region, likebox=box, sigma=sig,
                         chunk=8)
    orc = pipeline.Likelihood(h, par, r, spec, data, None, region, box, sig, 1.0)
    theta = np.array([[150., 10., 60., sig, 0.05], [140., 12., 55., 2 * sig, 0.], [150., 10., 60., 0.5 * sig, 0.2],
                      [150., 10., 60., -sig, 0.1]])
    ll = _compare(like, orc, theta, "sigma fracerr")
    assert np.isneginf(ll[-1])


def test_mc_free_centre_vs_oracle():
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    par = GalmodPar(GALMODTYPE="MC", ISEED=7)
    data, sig = _data(h, r, GalmodPar(GALMODTYPE="ANALYTIC"), noise=50.)
    spec = [("vrot", "single"), ("xpos", "single"), ("ypos", "single"), ("inc", "single")]
    region = np.ones((64, 64), bool)
    box = ((0, 0), (64, 64))
    like = GpuLikelihood(DiskModel(h, par=par), data, r, FreePars(spec, r.nr), region=region, likebox=box, sigma=sig,
                         chunk=8)
    orc = pipeline.Likelihood(h, par, r, spec, data, None, region, box, sig, 1.0)
    theta = np.array([[150., 31.5, 32.25, 60.], [140., 28.4, 34.9, 55.], [155., 35.1, 30.3, 64.]])
    _compare(like, orc, theta, "MC free centre")
    assert like.info()["mc_retries"] == 0
