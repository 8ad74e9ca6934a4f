"""Pins of the oracle's whole likelihood (oracle/pipeline.py, CPU): on noiseless data the truth leaves only the
normalisation term; with a moved centre (a box of its own) the residual plus the outside sum equals a brute-force
sum over the whole region of the model embedded in an empty cube; a negative free parameter gives -inf; the
fracerr form reduces to the Gaussian one as eps -> 0."""
import math

import numpy as np

import oracle
from oracle import pipeline
from bbarolo2_amd.galmod import GalmodPar, GalmodSetup, model_rings, model_box, nzq_for, zstep_for, PHI_TAB
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.rings import Rings
from bbarolo2_amd.smooth import beam_kernel


def _setup():
    h = emptyheader([48, 48, 32], [-10 / 3600., 10 / 3600., 10.], beam=30 / 3600.)
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    truth = Rings().setRings(radii=np.arange(15., 140., 30.), xpos=23.5, ypos=24.25, vsys=0., vrot=120., vdisp=10.,
                             z0=5., dens=np.linspace(2., 1., 5), inc=55., phi=40.)
    st = GalmodSetup(h, par)
    kern, _ = beam_kernel(h)
    st.zstep = zstep_for(h, st, True)
    return h, par, truth, st, kern, 1.0          # (DENSFLUX models: no flux scale after the convolution)


def _model_full(st, kern, sf, rings, blo, bhi, shape):
    r32 = rings.copy()
    for k, v in rings.as_dict().items():
        setattr(r32, k, np.asarray(v, np.float32).astype(np.float64))
    sr, _ = st.subrings(model_rings(r32, "CENTRED", st.radsep))
    m = oracle.conv_same(oracle.galmod_analytic(st, sr, blo, bhi, nzq_for(sr), PHI_TAB), kern, sf)
    full = np.zeros(shape)
    full[:, blo[1]:bhi[1], blo[0]:bhi[0]] = m.transpose(2, 0, 1)
    return full


def test_pipeline_truth_and_moved_centre():
    h, par, truth, st, kern, sf = _setup()
    shape = (32, 48, 48)
    blo, bhi = model_box(h, truth, "CENTRED", final=True)
    data = _model_full(st, kern, sf, truth, blo, bhi, shape).astype(np.float32)
    sigma, lf = 0.01, 0.4
    region = np.ones((48, 48), bool)
    region[:, :3] = False
    spec = [("vrot", "single"), ("dens", "ring"), ("xpos", "single")]
    orc = pipeline.Likelihood(h, par, truth, spec, data, None, region, ((0, 0), (48, 48)), sigma, lf)
    assert not orc.use_norm
    base = -lf * 0.5 * orc.nvox * math.log(2 * math.pi * sigma * sigma)
    th0 = np.concatenate([[120.], truth.dens, [23.5]])
    ll0 = orc.loglike(th0)
    assert abs(ll0 - base) < 1e-9 * abs(base)
    # a moved centre: the model in its own box, the data outside it from the summed-area table
    th = th0.copy()
    th[-1] = 30.25
    ll = orc.loglike(th)
    blo, bhi = orc.last_box
    assert bhi[0] == 48 and blo[0] > 0
    moved = truth.copy()
    moved.xpos = np.full(truth.nr, 30.25)
    full = _model_full(st, kern, sf, moved, blo, bhi, shape)
    used = (data != 0) & region[None]
    S = float(np.sum(np.where(used, (data.astype(np.float64) - full) ** 2, 0.)))
    ref = -lf * (0.5 * S / sigma ** 2 + 0.5 * orc.nvox * math.log(2 * math.pi * sigma ** 2))
    assert abs(ll - ref) <= 1e-9 * abs(ref)
    assert ll < ll0
    # negative free parameters
    bad = th0.copy()
    bad[0] = -1.
    assert orc.loglike(bad) == -np.inf


def test_pipeline_fracerr_limit():
    h, par, truth, st, kern, sf = _setup()
    data = _model_full(st, kern, sf, truth, (0, 0), (48, 48), (32, 48, 48)).astype(np.float32)
    data += np.random.default_rng(3).normal(0., 0.01, data.shape).astype(np.float32)
    region = np.ones((48, 48), bool)
    spec = [("vrot", "single"), ("dens", "ring")]
    g = pipeline.Likelihood(h, par, truth, spec, data, None, region, ((0, 0), (48, 48)), 0.01, 1.)
    e = pipeline.Likelihood(h, par, truth, spec, data, None, region, ((0, 0), (48, 48)), 0.01, 1., fracerr=1e-9)
    th = np.concatenate([[110.], truth.dens])
    a, b = g.loglike(th), e.loglike(th)
    assert abs(a - b) <= 1e-9 * abs(a)


def test_like_maps_sums():
    # small integers: every partial sum exact, so the chunked sums equal the plain ones bit for bit; random data:
    # within FP64 rounding of an exactly rounded sum; counts exact; blanks (NaN, 0) and the mask left out
    from oracle.pipeline import like_maps
    rng = np.random.default_rng(8)
    d = rng.integers(-9, 10, (37, 5, 6)).astype(np.float32)
    d[3, 1, 1] = np.nan
    lm = rng.random(d.shape) < 0.8
    dm2, nv, m0 = like_maps(d, lm)
    v = lm & np.isfinite(d) & (d != 0)
    dd = np.where(v, d, 0.).astype(np.float64)
    assert np.array_equal(dm2, (dd * dd).sum(0)) and np.array_equal(m0, dd.sum(0)) and np.array_equal(nv, v.sum(0))
    r = rng.standard_normal((50, 4, 4)).astype(np.float32)
    dm2, nv, m0 = like_maps(r)
    import math
    exact = np.array([[math.fsum(float(x) ** 2 for x in r[:, i, j]) for j in range(4)] for i in range(4)])
    assert np.allclose(dm2, exact, rtol=1e-14, atol=0) and np.all(nv == 50)
