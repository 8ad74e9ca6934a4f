"""GPU (MI355X) parity at the full sizes of the benchmark configurations, against the FP64 oracle pipeline
(oracle/pipeline.py: theta -> float32 rings -> deconvolved profile -> sub-rings per vector -> ANALYTIC model ->
'same' beam convolution -> masked residual + the squared data outside the box):

  * 160x160x96 GLOBAL (40 rings, 82-D, DENS from the profile of every vector): the truth, both prior corners, five
    prior draws and a nearly edge-on vector (long vertical columns: runs beyond the register slots, many chunks per
    tile); the same log-likelihoods with the tiles' chunk lists off and at one entry per chunk (every vector's
    lists overflow: its tiles scan all the chunk boxes), bit for bit; the deposit alone against orc_galmod_analytic;
  * 128x128x64 (8 rings, 16-D, the profile of the initial rings);
  * one ring of the 128x128x64 CLASSIC configuration (MC model, FTYPE 2, NORM LOCAL) against orc_norm.

Tolerances: log-likelihood within 1e-5 of its residual term; profiles within 1e-5 of their largest density (the
normal equations against the oracle's QR); models within 1e-5 of their peak (FP32 line profiles against FP64)."""
import math

import numpy as np
import pytest

import oracle
from oracle import pipeline
from bbarolo2_amd.galmod import model_rings, nzq_for, PHI_TAB

pytestmark = pytest.mark.gpu


def _thetas(fit, n_draws, seed=1):
    lo, hi = fit["lo"], fit["hi"]
    u = np.random.default_rng(seed).random((n_draws, fit["free"].ndim))
    return np.vstack([fit["theta_true"][None], lo[None], hi[None], lo + u * (hi - lo)])


def _check_loglike(fit, orc, theta, tag):
    like = fit["like"]
    ll = like.loglike(theta)
    dens = getattr(like, "last_dens", None)
    for i, th in enumerate(theta):
        ref = orc.loglike(th)
        tol = 1e-5 * fit["likefactor"] * orc.last_resid
        print(f"[{tag} {i}] gpu {ll[i]:.8f} oracle {ref:.8f} |d| {abs(ll[i] - ref):.3e} tol {tol:.3e}", flush=True)
        assert abs(ll[i] - ref) <= tol, (tag, i)
        if like.update_prof:
            d = orc.last_dens
            assert np.abs(dens[i] - d).max() <= 1e-5 * np.abs(d).max(), (tag, i)
    return ll


@pytest.fixture(scope="module")
def cfg2():
    from bbarolo2_amd.bayes import synthetic_fit
    fit = synthetic_fit(2, chunk=64)
    return fit, pipeline.of_fit(fit)


def _edge_on(fit):
    th = fit["theta_true"].copy()
    th[fit["free"].index["inc"]] = 86.
    return th


def test_config2_loglike_vs_oracle(cfg2):
    fit, orc = cfg2
    theta = np.vstack([_thetas(fit, 5), _edge_on(fit)[None]])
    _check_loglike(fit, orc, theta, "160x160x96")


def test_config2_chunk_lists_do_not_change_the_results(cfg2):
    fit, _ = cfg2
    theta = np.vstack([_thetas(fit, 3, seed=4), _edge_on(fit)[None]])
    like, dm = fit["like"], fit["dm"]
    ref = like.loglike(theta)
    try:
        for per in (0, 1):
            dm.set_bin(per)
            got = like.loglike(theta)
            print(f"[chunk lists {per} per chunk] max |d| {np.abs(got - ref).max():.3e}", flush=True)
            assert np.array_equal(got, ref)
    finally:
        dm.set_bin(8)


def test_config2_deposit_vs_oracle(cfg2):
    fit, orc = cfg2
    dm = fit["dm"]
    st = dm.setup
    tabs = []
    for th in (fit["theta_true"], _edge_on(fit)):
        rings, _ = orc.rings_of(th)
        rings.dens = np.asarray(orc.density(rings), np.float32).astype(np.float64)
        tabs.append(st.subrings(model_rings(rings, dm.ringmode, st.radsep))[0])
    sr = np.stack(tabs)
    blo, bhi = (0, 0), (160, 160)
    got = dm.array_batch(sr, blo, bhi, convolve=False).double().cpu().numpy()
    for b in range(sr.shape[0]):
        ref = oracle.galmod_analytic(st, sr[b], blo, bhi, nzq_for(sr[b]), PHI_TAB)
        err = np.abs(got[b] - ref).max() / ref.max()
        print(f"[deposit {b}] max |gpu - oracle| / max = {err:.3e}", flush=True)
        assert err < 1e-5


def test_config3_loglike_vs_oracle():
    from bbarolo2_amd.bayes import synthetic_fit
    fit = synthetic_fit(3, chunk=64)
    orc = pipeline.of_fit(fit)
    # (DENS is the profile of the initial rings, computed once)
    assert not fit["like"].update_prof and not orc.update_prof
    assert np.abs(fit["like"].dens0 - orc.dens0).max() <= 1e-5 * np.abs(orc.dens0).max()
    _check_loglike(fit, orc, _thetas(fit, 5), "128x128x64")


def test_config1_ring_residual_vs_oracle():
    """One ring of the CLASSIC configuration: 128x128x64, 10" pixels, 5 km/s, beam 30", MC model of 12 rings,
    FTYPE 2 (|m - o|), NORM LOCAL, WFUNC 2. The ring values of RingFit against orc_norm on the same models."""
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import RingFit, Fit3DPar
    from bbarolo2_amd.synth import config, add_noise
    h, truth, par, noise, seed, _, _ = config(1)
    dm = DiskModel(h, par=par, densflux=True)
    sr, _ = dm.subrings(truth)
    mod = dm.array_batch_mc(sr[None], (0, 0), (128, 128))[0].double().cpu().numpy()
    torch.cuda.synchronize()
    mz = mod.transpose(2, 0, 1)
    data = add_noise(mz, noise * mz.max(), dm.cfield, seed)
    mask = mz > 2 * noise * mz.max()
    inr = truth.copy()
    inr.vrot = np.full(truth.nr, 120.)
    inr.vdisp = np.full(truth.nr, 20.)
    fp = Fit3DPar(FREE="VROT VDISP", FTYPE=2, NORM="LOCAL", WFUNC=2)
    fit = RingFit(dm, data, mask, inr, par=fp, data_noise=float(noise * mz.max()))
    ps = fit.pixscale
    for ir in (2, 11):
        d = fit.make_dring(ir)
        trials = []
        for vrot, vdisp in ((140., 10.), (100., 25.)):
            r = d.copy()
            r.vrot[:], r.vdisp[:] = vrot, vdisp
            trials.append(r)
        vals = fit.values(trials)
        for r, v in zip(trials, vals):
            blo, bhi = dm.box(r, final=False)
            srr, _ = dm.setup.subrings(r)
            m = dm.array_batch_mc(srr[None], blo, bhi)[0].double().cpu().numpy()
            ref = oracle.norm3dfit(m, data.transpose(1, 2, 0), mask.transpose(1, 2, 0).astype(np.uint8), blo,
                                   r.xpos[-1] - blo[0], r.ypos[-1] - blo[1], r.phi[-1], r.inc[-1], r.radii[0] / ps,
                                   r.radii[-1] / ps, fp.WFUNC, 0, fp.FTYPE, np.ones(data.shape[0]), fit.data_noise,
                                   fp.BWEIGHT, 0)[0]
            print(f"[128x128x64 ring {ir}] vrot {r.vrot[0]} vdisp {r.vdisp[0]}: gpu {v:.12e} oracle {ref:.12e}",
                  flush=True)
            assert math.isfinite(ref) and abs(v - ref) <= 1e-9 * abs(ref)
