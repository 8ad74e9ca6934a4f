"""Host logic of the model engine (CPU): ring conventions, header-derived setup, sub-rings, beam kernel."""
import math

import numpy as np
import pytest

from bbarolo2_amd.galmod import (GalmodSetup, GalmodPar, model_rings, outer_edge, edge_weights, SubringPlan,
                                 GalmodError, zstep_for, ZSTEP_FINE, ZSTEP_COARSE, subring_radii)
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.rings import Rings
from bbarolo2_amd.smooth import beam_kernel, fillgauss2d, convpars
from conftest import small_rings


def test_ring_conventions_golden():
    """Golden lists of the reference (tests/test_bbarolo.py:385-402)."""
    r = Rings().setRings(radii=[30., 90, 150, 210], vrot=[60, 100, 120, 130], dens=[8, 4, 2, 1], radsep=60.)
    c = model_rings(r, "CENTRED")
    assert list(c.radii) == [0, 30, 90, 150, 210, 240] and list(c.vrot) == [60, 60, 100, 120, 130, 130]
    assert list(c.dens) == [8, 8, 4, 2, 1, 1]
    s = model_rings(r, "SHIFTED")
    assert list(s.radii) == [0, 90, 150, 240] and list(s.vrot) == [60, 100, 120, 130]
    assert model_rings(r, "NODES") is r and model_rings(r, "centered").nr == 6
    assert outer_edge(r, "CENTRED") == 240 and outer_edge(r, "NODES") == 210 and outer_edge(r, "SHIFTED") == 240
    one = Rings().setRings(radii=[30.], vrot=[60.], radsep=20.)
    assert list(model_rings(one, "CENTRED").radii) == [20, 30, 40]
    with pytest.raises(ValueError):
        model_rings(r, "EDGES")


def test_setup_constants(small_header):
    st = GalmodSetup(small_header, GalmodPar(), densflux=True)
    assert st.axtyp == 4 and st.nsubs == 40
    assert math.isclose(st.pixarea, (10 / 60.) ** 2) and math.isclose(st.radsep, 0.75 * 10 / 60.)
    assert st.chwidth == 10000.0
    assert np.all(st.cd2i == np.float32(3600. / 10000. * 1000.))
    assert math.isclose(st.sig_instr, 10000. / (2 * math.sqrt(2 * math.log(2))))
    # channel edges: the inverse of the channel grid at k - 1/2 (velocity axis: linear)
    k = np.arange(41) - 0.5
    assert np.allclose(st.vedges, (k + 1 - st.crpix3) * st.cdelt3 + st.crval3)
    assert st.chmin == pytest.approx(10000.)


def test_zstep_rule():
    """0.75 pixels for a beam of >= 3 pixels (minor axis) when smoothing, else 0.25 (test_bbarolo.py:5096-5116)."""
    for beam, smooth, want in ((4., True, ZSTEP_COARSE), (3., True, ZSTEP_COARSE), (2.9, True, ZSTEP_FINE),
                               (4., False, ZSTEP_FINE)):
        h = emptyheader([96, 96, 60], [-1 / 3600., 1 / 3600., 5.], beam=beam / 3600.)
        st = GalmodSetup(h)
        assert zstep_for(h, st, smooth) == want


def test_subring_radii_float32():
    """Sub-ring radii accumulate in float32 from radsep/2 (galmod.py:805-819): exact float32 arithmetic."""
    rad = np.array([0., 0.5, 1.0, 1.5])
    idx, dr, const, rads = subring_radii(rad, 0.125, True)
    r, expect = np.float32(0.125 / 2), []
    while r < np.float32(1.5):
        expect.append(float(r))
        r = np.float32(float(r) + 0.125)
    assert list(rads) == expect and not const.any()
    assert idx[0] == 1 and idx[-1] == 3


def test_subrings_interpolation(small_header):
    st = GalmodSetup(small_header)
    r = small_rings(vrot=np.linspace(100, 200, 8), dens=np.linspace(2., 1., 8))
    sr, nv = st.subrings(r)
    R = sr[0] * 60.
    # linear interpolation of the ring values between the radii (measured from the float32 radius of the
    # inner ring: 1e-7 relative offsets)
    assert np.allclose(sr[1] / 1000., np.interp(R, r.radii, r.vrot), rtol=1e-6, atol=0)
    assert np.allclose(sr[5], np.interp(R, r.radii, r.dens), rtol=1e-6, atol=0)
    assert np.all(sr[14] == 1) and np.all(sr[15] == 1) and np.all(nv == 40)
    assert np.allclose(sr[9], np.radians(60.)) and np.allclose(sr[10], np.radians(45.))


def test_subrings_thick_superposition(small_header):
    """A thick layer doubles the sub-rings: thin (dens (1 - f), no lag) + thick (dens f, z0thick, vdispthick, lag)."""
    st = GalmodSetup(small_header)
    r = small_rings(z0=5., fthick=0.3, z0thick=20., vdispthick=25., dvdz=2.)
    sr, nv = st.subrings(r)
    n = sr.shape[1] // 2
    one = small_rings(z0=5.)
    s1, _ = st.subrings(one)
    assert sr.shape[1] == 2 * s1.shape[1]
    assert np.allclose(sr[5, :n], 0.7 * s1[5]) and np.allclose(sr[5, n:], 0.3 * s1[5])
    assert np.all(sr[7, :n] == 0) and np.allclose(sr[7, n:], 2. * 1000 * 60)
    assert np.allclose(sr[6, n:], 20. / 60.) and np.allclose(sr[4, n:], 25000.)


def test_side_edges(small_header):
    """Receding / approaching halves with their own edges: the side weights are the fraction of the sub-ring
    within the edge (galmod.py:834-845, edge_weights)."""
    st = GalmodSetup(small_header, GalmodPar(RADMAXA=80., RADMAXR=-1.))
    sr, _ = st.subrings(small_rings())
    assert np.all(sr[14] == 1.)
    w = edge_weights(sr[0] * 60., st.radsep * 60., 80.)
    assert np.array_equal(sr[15], w) and w.min() == 0. and w.max() == 1.


def test_subrings_validation(small_header):
    st = GalmodSetup(small_header)
    with pytest.raises(GalmodError):
        st.subrings(small_rings(vdisp=-1.))
    with pytest.raises(GalmodError):
        st.subrings(Rings().setRings(radii=[10., 11.]))          # radius separation below a sub-ring
    with pytest.raises(GalmodError):
        st.subrings(Rings().setRings(radii=[30., 10.]))


@pytest.mark.parametrize("thick", [False, True])
def test_subring_plan_matches_subrings(small_header, thick):
    st = GalmodSetup(small_header)
    r = small_rings(z0=5., fthick=0.2 if thick else 0., z0thick=15.)
    plan = SubringPlan(st, r, "CENTRED")
    rng = np.random.default_rng(0)
    vals = dict(vrot=rng.uniform(50, 250, (3, 8)), vdisp=rng.uniform(5, 30, (3, 8)), inc=rng.uniform(40, 80, (3, 8)))
    sr, nv = plan.tables(vals)
    for b in range(3):
        rb = r.copy()
        rb.vrot, rb.vdisp, rb.inc = vals["vrot"][b], vals["vdisp"][b], vals["inc"][b]
        ref, nv2 = st.subrings(model_rings(rb, "CENTRED", st.radsep))
        assert np.array_equal(sr[b], ref) and np.array_equal(nv, nv2)


def test_beam_kernel_closed_form(small_header):
    """The convolution beam from a point source is the beam itself (Wild 1970): a Gaussian of FWHM 3 pixels,
    normalised to unit sum, symmetric, cut below 1e-4 of its peak (smooth.py:62-88)."""
    k, scale = beam_kernel(small_header)
    assert k.shape[0] % 2 == 1 and k.shape == k.shape[::-1]
    assert abs(k.sum() - 1) < 1e-12 and np.array_equal(k, k[::-1, ::-1]) and np.allclose(k, k.T, rtol=1e-12, atol=0)
    c = k.shape[0] // 2
    # FWHM 3 pixels along the axes: k(1.5 px) / k(0) = 1/2 -> k(1) / k(0) = 2^(-4/9)
    assert math.isclose(k[c, c + 1] / k[c, c], 2 ** (-4. / 9.), rel_tol=1e-12)
    assert (k[k > 0].min() / k.max()) >= 1e-4
    assert convpars((10., 5., 30.), (10., 5., 30.))[0] == pytest.approx(0., abs=1e-6)
    # elliptical beam at PA 30: major axis rotated
    h = emptyheader([64, 64, 40], [-10 / 3600., 10 / 3600., 20.], beam=(40 / 3600., 30 / 3600., 30.))
    k2, _ = beam_kernel(h)
    assert abs(k2.sum() - 1) < 1e-12 and (k2.shape[0] != k2.shape[1] or not np.allclose(k2, k2.T, rtol=1e-3))
    assert np.array_equal(k2, k2[::-1, ::-1])


def test_fillgauss_scalefactor_point():
    """Scale factor of a beam convolved from a point: 1/(G_old*K)(0) with G_old the old beam sampled with
    amplitude 1: for a tiny old beam the convolution at the centre is K(0) = peak of the normalised kernel."""
    g = fillgauss2d((3., 3., 0.), (1., 1.), 1.0, norm=False)
    assert g[g.shape[0] // 2, g.shape[1] // 2] == 1.0


def test_profile_prior_rows_match_oracle():
    """The per-fit curvature rows of the device profile equal the oracle's (dense L rebuilt from the 3-column
    form), as do the relative-prior norm, the median spacing and the flags of near radii."""
    from bbarolo2_amd.profile import prior_rows
    from oracle import profile as op
    for R in ([15., 45., 75., 105., 135., 170.], [1., 2., 3., 4., 5., 4.5], [1., 2., 3., 3.],
              list(np.arange(40) * 30. + 15.) + [1230.]):
        lcoef, lcol, near, norm, dmed = prior_rows(R)
        L = np.zeros((len(R) - 2, len(R)))
        for r in range(len(R) - 2):
            L[r, lcol[r]] = lcoef[r]
        Lo, order, Rs, tiny, dmin = op.curvature(R)
        assert np.allclose(L, Lo, rtol=1e-13, atol=0)
        _, s, _ = op.relative_scale(R, 7.0, 0.1)
        assert np.isclose(np.sqrt(0.1 * 7.0 / norm), s, rtol=1e-13)
        assert dmed == np.median(np.diff(np.sort(np.asarray(R)[:-1])))
        gaps = np.diff(Rs)
        assert np.array_equal(near.astype(bool), np.minimum(gaps[:-1], gaps[1:]) < 0.25 * dmed)


def test_azplan_cache_evicts_least_recently_used(monkeypatch):
    # the plan cache of DiskModel keeps the PLAN_CACHE most recently used plans and destroys only the oldest
    import types
    from bbarolo2_amd import engine
    made, destroyed = [], []

    class FakeLib:
        def bb_azplan_create(self, h, r0, n, out):
            made.append(len(made) + 1)
            out._obj.value = made[-1]
            return 0

        def bb_azplan_destroy(self, p):
            destroyed.append(p.value)
            return 0

    fake_torch = types.SimpleNamespace(cuda=types.SimpleNamespace(current_device=lambda: 0))
    monkeypatch.setattr(engine, "_torch", lambda: fake_torch)
    monkeypatch.setattr(engine._lib, "lib", lambda: FakeLib())
    dm = engine.DiskModel.__new__(engine.DiskModel)
    dm._plans = None
    dm.handle = lambda nzq=None: None
    n = engine.PLAN_CACHE
    plans = [dm.azplan(np.array([float(i)])) for i in range(n)]
    assert dm.azplan(np.array([0.])) is plans[0] and not destroyed          # (a hit: plan 0 most recent)
    dm.azplan(np.array([float(n)]))                                        # (a new plan: the oldest, 1, goes)
    assert destroyed == [2] and len(dm._plans) == n
    assert dm.azplan(np.array([0.])) is plans[0]
    dm.azplan(np.array([float(n + 1)]))
    assert destroyed == [2, 3]
