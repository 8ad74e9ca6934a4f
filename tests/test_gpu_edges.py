"""GPU (MI355X): the edges of the disk as free parameters of the likelihood context (RADMAX, RADMAXA, RADMAXR: the
reference's default FREE holds RADMAX, keywords.py:700; _update_rings, bayesian/fitter.py:1112-1120).

  * the model's sub-ring tables with a free RADMAX (the density edge weights of each vector, edge_weights,
    galmod.py:1122-1131) and with free RADMAXA / RADMAXR (the side weights) equal the host's tables of the same
    rings with that edge fixed, bit for bit; the exterior annulus of each vector (its radius from the vector's
    edge, profile.py:45-64; float32-accumulated sub-ring radii) equals the host's plan of the annulus of that edge;
  * the log-likelihoods with RADMAX free (sharp and Gaussian edges, the edge beyond the rings, in the last ring,
    inside the disk so that the annulus falls between the rings and the prior's rows change order, near the centre,
    negative: -inf) and with RADMAXA and RADMAXR free (one half without an edge) agree with the FP64 oracle pipeline,
    which sets rings.edge / edgeapp / edgerec from theta and builds the profile, annulus and model of those rings;
    the same with DENS fitted (no profile: the edge weights go into the model's densities only);
  * a free RADMAX at a value gives the log-likelihood of the same context with that edge fixed (to 1e-7 of the
    residual term: the two differ in the order of the products of the ring weights and in the azimuths of the
    annulus, taken from cos and sin per vector rather than from the plan's recurrence).
Tolerance against the oracle: 1e-5 of the residual term (FP32 models against FP64)."""
import numpy as np
import pytest

from oracle import pipeline
from bbarolo2_amd.galmod import GalmodPar, SubringPlan
from test_gpu_ctx import _disk, _data, _compare

pytestmark = pytest.mark.gpu

SPEC = [("vrot", "single"), ("vdisp", "single"), ("inc", "single"), ("radmax", "single")]


def _setup(spec, radmax="edge", width=None, chunk=16, rings=None):
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    if rings is not None:
        r = rings
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    data, sig = _data(h, r, par)
    region = np.ones((64, 64), bool)
    box = ((0, 0), (64, 64))
    like = GpuLikelihood(DiskModel(h, par=par), data, r, FreePars(spec, r.nr), region=region, likebox=box, sigma=sig,
                         likefactor=0.5, chunk=chunk, radmax=radmax, radmax_width=width)
    return h, r, par, data, sig, region, box, like


def test_edge_tables_vs_host():
    from bbarolo2_amd.profile import exterior_annulus, ANNULUS_KEYS
    spec = [("vrot", "single"), ("inc", "single"), ("radmax", "single"), ("radmaxa", "single"), ("radmaxr", "single")]
    h, r, par, data, sig, region, box, like = _setup(spec, chunk=8)
    assert like.update_prof and like.ann_cap
    st = like.dm.setup
    theta = np.array([[150., 60., 200., -1., -1.], [140., 55., 155., 120., -1.], [150., 62., 130., -1., 90.],
                      [150., 60., 75., 100., 140.], [160., 58., 12.5, 60., 30.], [150., 60., 63.7, -1., 200.]])
    sr_d, ext_d = like.device_tables(theta)
    beam = abs(h.Bmaj()) * 3600.
    for i, th in enumerate(theta):
        rr = like.rings0.copy()
        rr.vrot = np.full(rr.nr, float(np.float32(th[0])))
        rr.inc = np.full(rr.nr, float(np.float32(th[1])))
        rr.edge, rr.edgewidth = float(th[2]), 0.
        rr.edgeapp, rr.edgerec = float(th[3]), float(th[4])
        plan = SubringPlan(st, rr, like.dm.ringmode, final=True)
        sr_h = plan.tables({})[0][0]
        rows = [k for k in range(16) if k != 5]
        assert np.array_equal(sr_d[i][rows], sr_h[rows]), i
        # the density row before the profile: the ring table's densities times the edge weights
        assert np.array_equal(sr_d[i][5], sr_h[5]), i
        ext, rext = exterior_annulus(st, rr, like.dm.ringmode, beam)
        pe = SubringPlan(st, ext, "CENTRED", final=True)
        e_h = pe.tables({k: np.asarray(getattr(rr, k))[-1:] for k in ANNULUS_KEYS})[0][0]
        n = e_h.shape[1]
        print(f"[edges {th[2:]}] annulus at {rext:.3f}\": {n} sub-rings of {like.ann_cap}", flush=True)
        assert 0 < n <= like.ann_cap
        assert np.array_equal(ext_d[i][:, :n], e_h), i
        assert np.all(ext_d[i][14:, n:] == 0.), i


@pytest.mark.parametrize("mode", ["edge", "taper"])
def test_radmax_free_vs_oracle(mode):
    h, r, par, data, sig, region, box, like = _setup(SPEC, radmax=mode)
    width = like.edgewidth
    assert (width > 0) == (mode == "taper")
    orc = pipeline.Likelihood(h, par, r, SPEC, data, None, region, box, sig, 0.5, radmax_width=width)
    theta = np.array([[150., 10., 60., 250.],       # beyond the rings: the annulus at their outer edge
                      [150., 10., 60., 155.],       # in the outer half of the last ring
                      [145., 11., 58., 131.5],      # between the last two radii
                      [150., 10., 60., 75.],        # inside the disk: the annulus before the outer rings
                      [155., 9., 62., 22.],         # near the centre
                      [150., 10., 60., -3.]])       # negative: zero likelihood
    ll = _compare(like, orc, theta, f"RADMAX {mode}")
    assert np.isneginf(ll[-1]) and np.all(np.isfinite(ll[:-1]))
    assert like.info()["nnls_fail"] == 0


def test_radmax_halves_vs_oracle():
    spec = [("vrot", "single"), ("inc", "single"), ("radmaxa", "single"), ("radmaxr", "single")]
    h, r, par, data, sig, region, box, like = _setup(spec)
    assert like.update_prof
    orc = pipeline.Likelihood(h, par, r, spec, data, None, region, box, sig, 0.5)
    theta = np.array([[150., 60., 100., 140.], [150., 60., -1., 120.], [145., 57., 130., -1.], [150., 60., 60., 90.],
                      [150., 61., -1., -1.]])
    _compare(like, orc, theta, "RADMAXA RADMAXR")


def test_free_edge_equals_fixed_edge():
    spec_free = [("vrot", "single"), ("inc", "single"), ("radmax", "single")]
    spec_fix = spec_free[:2]
    h, r, par, data, sig, region, box, like = _setup(spec_free, chunk=8)
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars
    for E in (200., 131.5, 75.):
        rf = r.copy()
        rf.edge, rf.edgewidth = E, 0.
        fixed = GpuLikelihood(DiskModel(h, par=par), data, rf, FreePars(spec_fix, r.nr), region=region, likebox=box,
                              sigma=sig, likefactor=0.5, chunk=8)
        th = np.array([[150., 60.], [140., 56.]])
        a = like.loglike(np.column_stack([th, [E, E]]))
        b = fixed.loglike(th)
        resid = np.abs(b)
        print(f"[edge {E}] free {a} fixed {b} |d| {np.abs(a - b)}", flush=True)
        assert np.all(np.abs(a - b) <= 1e-7 * resid)


def test_radmax_free_with_dens_fitted_vs_oracle():
    """DENS fitted (no profile): the edge weights go into the model's densities only."""
    spec = [("vrot", "single"), ("dens", "ring"), ("radmax", "single")]
    h, r, par, data, sig, region, box, like = _setup(spec, chunk=8)
    assert not like.useNorm and like.ann_cap is None
    orc = pipeline.Likelihood(h, par, r, spec, data, None, region, box, sig, 0.5)
    d = np.asarray(r.dens, float)
    theta = np.array([np.r_[150., d, 250.], np.r_[145., 0.8 * d, 131.5], np.r_[150., d, 75.], np.r_[150., d, -1.]])
    ll = _compare(like, orc, theta, "RADMAX, DENS fitted")
    assert np.isneginf(ll[-1]) and np.all(np.isfinite(ll[:-1]))
