"""Host side of the 3DFIT ring fits (bbarolo2_amd/galfit.py, CPU): FREE parsing, the pixel count of a ring against
lattice counts, the annuli of the rings, the bounds of the trial values, and the simplex on closed-form objectives
(its minimum, its failure at NMAX, and batched evaluations giving the same iterates as one-by-one ones)."""
import math

import numpy as np
import pytest

from bbarolo2_amd import galfit as G
from bbarolo2_amd.rings import Rings


def test_free_flags():
    m = G.free_flags("VROT VDISP")
    assert m[G.VROT] and m[G.VDISP] and sum(m) == 2
    m = G.free_flags("vrot pa inc")
    assert m[G.PA] and m[G.INC] and sum(m) == 3
    assert G.free_flags("phi")[G.PA]
    assert all(G.free_flags("ALL"))
    with pytest.raises(ValueError):
        G.free_flags("dens")                 # (DENS is never free in 3DFIT: no flag, hence no free parameter)


def test_count_ring_pixels_face_on():
    """Face-on annulus: the lattice points with r1 <= r <= r2 of the box edges included; masked pixels among them."""
    mask = np.zeros((50, 50), int)
    mask[20:, :] = 3
    blo, bhi = (5, 6), (40, 44)
    x0, y0 = 20.3, 18.7
    nt, ni = G.count_ring_pixels(mask, blo, bhi, x0, y0, 0., 0., 6., 11.)
    ys, xs = np.mgrid[blo[1]:bhi[1] + 1, blo[0]:bhi[0] + 1]
    r = np.hypot(xs - blo[0] - x0, ys - blo[1] - y0)
    ring = (r >= 6) & (r <= 11)
    assert nt == ring.sum() and ni == (ring & (ys >= 20)).sum() and 0 < ni < nt


def test_count_ring_pixels_inclined():
    """A disk (r1 = 0) at inclination i covers pi R^2 cos i pixels; the count is invariant under a rotation by
    PA = 180 (the ellipse is symmetric); pixels outside the mask array count in nTot only."""
    mask = np.ones((10, 10), int)
    nt, ni = G.count_ring_pixels(mask, (0, 0), (120, 120), 60.2, 60.1, 30., 60., 0., 40.)
    assert abs(nt - math.pi * 1600 * 0.5) < 0.02 * math.pi * 1600 * 0.5
    assert nt == G.count_ring_pixels(mask, (0, 0), (120, 120), 60.2, 60.1, 210., 60., 0., 40.)[0]
    assert ni == 0


def _inr(radii, **kw):
    d = dict(radii=radii, xpos=30., ypos=30., vsys=0., vrot=100., vdisp=10., z0=0., dens=1., inc=60., phi=45.)
    d.update(kw)
    return Rings().setRings(**d)


class Quad(G.RingSimplex):
    """A closed-form objective: sum of ((value - target) / scale)^2 over the free parameters, + 1."""

    def __init__(self, inr, par, target, one_by_one=False, **kw):
        super().__init__(inr, par, 64, 64, np.ones((64, 64)), 10., lambda r: ((0, 0), (64, 64)), **kw)
        self.target = target
        self.one_by_one = one_by_one
        self.calls = []

    def values(self, rings):
        self.calls.append(len(rings))
        self.nevals += len(rings)
        out = []
        for r in rings:
            out.append(1. + sum(((getattr(r, k)[0] - t) / s) ** 2 for k, (t, s) in self.target.items()))
        return np.array(out)

    def func3D_batch(self, dring, points):
        if not self.one_by_one:
            return super().func3D_batch(dring, points)
        return np.concatenate([super(Quad, self).func3D_batch(dring, [p]) for p in points])


def test_make_dring():
    inr = _inr([10.1, 30.3, 50.6, 70.2])
    q = Quad(inr, G.Fit3DPar(), {})
    r0, r1, r3 = q.make_dring(0), q.make_dring(1), q.make_dring(3)
    f, d = np.float32, np.float64                     # (radius in FP64 +- a half width rounded to FP32)
    assert r0.radii[0] == 0. and r0.radii[1] == float(f(d(10.1) + f(10.1)))
    assert r1.radii[0] == float(f(d(30.3) - f(10.1))) and r1.radii[1] == float(f(d(30.3) + f((50.6 - 30.3) / 2.)))
    assert r3.radii[1] == float(f(d(70.2) + f((70.2 - 50.6) / 2.)))
    assert r1.radii[1] != float(f(30.3 + f((50.6 - 30.3) / 2.)))    # (not the all-FP32 sum)
    assert r1.id == 1 and r1.radsep == r1.radii[1] - r1.radii[0] and r1.vrot[0] == 100.
    one = Quad(_inr([20.]), G.Fit3DPar(), {}).make_dring(0)
    assert one.radii[0] == 10. and one.radii[1] == 30.


def test_clamp_bounds():
    """Out-of-range trial values: VROT and VDISP redrawn inside their bands, INC and PA clamped to theirs, Z0 and
    the centre back to the initial values; the trial vector gets the values used."""
    inr = _inr([10., 30., 50.], z0=5.)
    par = G.Fit3DPar(FREE="VROT VDISP Z0 INC PA XPOS", DELTAVROT=40., DELTAINC=5., DELTAPHI=15.)
    q = Quad(inr, par, {})
    d = q.make_dring(1)
    z = np.array([500., -3., 2000., 80., 10., 1000.])
    vals = q.clamp(d, z)
    assert 100. < vals["vrot"] <= 140.               # maxv - u DELTAVROT, maxv = 100 + 40, u in [0, 1)
    assert par.MINVDISP <= vals["vdisp"] <= par.MINVDISP + 10.
    assert vals["z0"] == 5. and vals["inc"] == 65. and vals["phi"] == 30. and vals["xpos"] == 30.
    assert np.array_equal(z, [vals[k] for k in ("vrot", "vdisp", "z0", "inc", "phi", "xpos")])
    z = np.array([-100., 5000., 3., 40., 70., 31.])
    vals = q.clamp(d, z)
    assert 60. <= vals["vrot"] < 100. and par.MAXVDISP - 10. <= vals["vdisp"] <= par.MAXVDISP
    assert (vals["z0"], vals["inc"], vals["phi"], vals["xpos"]) == (3., 55., 60., 31.)


def test_simplex_minimum():
    inr = _inr([10., 30., 50.])
    par = G.Fit3DPar(FREE="VROT VDISP INC", TOL=1e-6)
    q = Quad(inr, par, {"vrot": (123.4, 10.), "vdisp": (17.2, 3.), "inc": (62.5, 2.)})
    ok, fmin, p = q.minimize(q.make_dring(1))
    assert ok and fmin < 1. + 1e-4
    assert abs(p[0] - 123.4) < 0.2 and abs(p[1] - 17.2) < 0.06 and abs(p[2] - 62.5) < 0.04
    # first batch: the ndim + 1 vertices together
    assert q.calls[0] == 4


def test_simplex_nmax():
    """A tolerance no simplex reaches: the fit stops after at most 200 x ndim function-value steps, not ok."""
    inr = _inr([10., 30., 50.])
    q = Quad(inr, G.Fit3DPar(FREE="VROT VDISP", TOL=-1.), {"vrot": (123.4, 10.), "vdisp": (17.2, 3.)})
    ok, fmin, _ = q.minimize(q.make_dring(1))
    assert not ok and fmin >= 1. and 400 <= q.nevals - 3 <= 400 + 2 + 2


@pytest.mark.parametrize("free", ["VROT VDISP", "VROT VDISP INC PA", "VROT"])
def test_batched_equals_one_by_one(free):
    """Batched evaluations (first vertices, shrinks) draw the same random numbers in the same order as
    one-by-one ones: identical iterates, results and evaluation counts, also with redraws at the bounds."""
    inr = _inr([10., 30., 50.], vdisp=3.)
    par = G.Fit3DPar(FREE=free, TOL=1e-5, DELTAVROT=8., DELTAINC=2., DELTAPHI=3.)
    tgt = {"vrot": (109., 4.), "vdisp": (0.05, 1.), "inc": (70., 1.), "phi": (30., 1.)}
    res = []
    for obo in (False, True):
        q = Quad(inr, par, tgt, one_by_one=obo)
        q.rng = np.random.default_rng([12345, 1, 1])
        r = q.minimize(q.make_dring(1))
        res.append((r, q.nevals, q.rng.random()))
    (a, na, ua), (b, nb, ub) = res
    assert a[0] == b[0] and a[1] == b[1] and np.array_equal(a[2], b[2]) and na == nb and ua == ub


def test_fit_skips_rings_outside_mask():
    inr = _inr([10., 30., 50.])
    q = Quad(inr, G.Fit3DPar(FREE="VROT"), {"vrot": (120., 5.)})
    q.mask2d = np.zeros((64, 64))
    out, res = q.fit()
    assert res == [None, None, None] and np.array_equal(out.vrot, inr.vrot)
    q.mask2d = np.ones((64, 64))
    out, res = q.fit()
    assert all(r[0] for r in res) and np.allclose(out.vrot, 120., atol=0.5)


def test_stage_rounds_equal_rings_alone():
    """The stage's lock-step rounds (every ring's simplex one step per round, all their trials evaluated together)
    give each ring the result of fitting it alone with its own generator [12345, stage, ring]: same minimum, point
    and number of evaluations; and each round holds the trials of every ring still running."""
    inr = _inr([10., 30., 50., 70.], vdisp=3.)
    par = G.Fit3DPar(FREE="VROT VDISP INC", TOL=1e-4, DELTAVROT=8., DELTAINC=2.)
    tgt = {"vrot": (109., 4.), "vdisp": (5., 1.), "inc": (62., 1.)}

    class Rounds(Quad):
        def values_rounds(self, groups):
            self.rounds.append([len(g) for _, g in groups])
            return super().values_rounds(groups)

    q = Rounds(inr, par, tgt)
    q.rounds = []
    out, res = q.fit(stage=2)
    assert q.rounds[0] == [4, 4, 4, 4] and len(q.rounds) > 10
    for ir in range(4):
        a = Quad(inr, par, tgt)
        r = a.fit_ring(ir, stage=2)
        assert r[0] == res[ir][0] and r[1] == res[ir][1] and np.array_equal(r[2], res[ir][2])
        if r[0]:
            assert out.vrot[ir] == r[2][0] and out.inc[ir] == r[2][2]
        else:
            assert out.vrot[ir] == inr.vrot[ir]
    assert sum(r[0] for r in res) >= 1


def test_reflection_keeps_the_clamped_point():
    """The reflected, expanded or contracted point enters the simplex as evaluated, clamped to the bounds (func3D
    writes the values back into its trial vector, galfit.py:832-889, 964-973): with the minimum of INC beyond its
    band, the fit's INC stays inside the band (at its edge), not where the unclamped simplex would wander."""
    inr = _inr([10., 30., 50.])
    par = G.Fit3DPar(FREE="VROT INC", TOL=1e-7, DELTAINC=5.)
    q = Quad(inr, par, {"vrot": (110., 10.), "inc": (80., 2.)})
    ok, fmin, p = q.minimize(q.make_dring(1))
    assert ok and 55. <= p[1] <= 65. and abs(p[1] - 65.) < 0.05 and abs(p[0] - 110.) < 0.5


# ---------------------------------------------------------------------- REVERSE (host pieces)
def test_context_dens():
    """The profile scaled to the ring's own density; a ring whose profile is below 1e-3 of the peak is scaled as if
    it were 1e-3 of it (no division by ~0)."""
    prof = np.array([4., 2., 1., 0.])
    assert np.allclose(G.context_dens(prof, 1, 3.), [6., 3., 1.5, 0.])
    assert np.allclose(G.context_dens(prof, 3, 2.), prof * 2. / 4e-3)


def test_segment_nodes_edges_and_status():
    """The rings lo..hi as nodes, extended to the edges of the annuli of rings lo and hi (more than a sub-ring away);
    CS_NONE rings have no emission; a CS_FAILED ring takes the values interpolated from the CS_OK ones; the inner
    part is empty unless the segment starts at the first ring."""
    inr = _inr([10., 30., 50., 70., 90.], vrot=[50., 80., 999., 120., 130.])
    q = Quad(inr, G.Fit3DPar(), {})
    elo, ehi = q.ring_edges()
    assert np.allclose(elo, [0., 20., 40., 60., 80.]) and np.allclose(ehi, [20., 40., 60., 80., 100.])
    cst = np.array([G.CS_OK, G.CS_OK, G.CS_FAILED, G.CS_OK, G.CS_NONE])
    dens = np.array([1., 2., 3., 4., 5.])
    seg = q.segment(1, 4, dens, inr, cst, 6.)
    assert np.allclose(seg.radii, [20., 30., 50., 70., 90., 100.])        # inner edge 20, outer edge 100
    assert np.allclose(seg.dens, [2., 2., 3., 4., 0., 0.])                # ring 4 CS_NONE: no emission
    assert seg.vrot[2] == 100. and seg.vrot[3] == 120.                    # ring 2 failed: halfway between 80 and 120
    assert seg.empty is True
    seg0 = q.segment(0, 1, dens, inr, cst, 6.)
    assert seg0.empty is None and np.allclose(seg0.radii, [0., 10., 30., 40.])    # (from the centre: inner edge 0)
    near = q.segment(1, 1, dens, inr, cst, 15.)                           # (edges within a sub-ring: no extra node)
    assert np.allclose(near.radii, [30.])


def test_strip_region():
    """Half height of the strip: the outer radius projected at the initial inclination + 3 scale heights + half the
    beam, in pixels; the box grown to hold it around the centre, within the cube."""
    inr = _inr([10., 30., 50.], z0=4., inc=80.)
    q = Quad(inr, G.Fit3DPar(), {})
    h = q.strip_hmax(1, 40., 30.)
    assert abs(h - (40. * math.cos(math.radians(80.)) + 12. + 15.) / 10.) < 1e-12
    d = q.make_dring(1)
    lo, hi = q.strip_box(d, (25, 27), (36, 34), h)
    ext = int(math.ceil(max(40. / 10., h))) + 1
    assert lo == [30 - ext, 30 - ext] and hi == [30 + ext + 1, 30 + ext + 1]
    lo, hi = q.strip_box(d, (0, 0), (64, 64), 100.)
    assert lo == [0, 0] and hi == [64, 64]


def test_beam_rings():
    """REVERSE with rings narrower than the beam: rings one beam wide out to the last radius, the initial values
    interpolated at the new radii, the position angle unwrapped across 360."""
    from bbarolo2_amd.classic import beam_rings
    inr = _inr([5., 15., 25., 35., 45., 55.], vrot=[10., 20., 30., 40., 50., 60.], phi=[350., 355., 359., 3., 8., 12.])
    r = beam_rings(inr, 20.)
    assert np.allclose(r.radii, [10., 30., 50.]) and r.radsep == 20.
    assert np.allclose(r.vrot, [15., 35., 55.])
    assert np.allclose(r.phi, [352.5, 361., 370.])


def test_solve_profile_relative_vs_oracle():
    """The relative-prior NNLS of the REVERSE profile (stack then QR) against the oracle's first solution of the
    profile (QR then stack): the same least-squares problem."""
    from scipy.optimize import nnls
    import oracle.profile as op
    from bbarolo2_amd.profile import solve_profile_relative
    rng = np.random.default_rng(5)
    radii = np.r_[np.arange(8) * 30. + 15., 260.]
    A = np.abs(rng.normal(size=(200, 9)))
    b = A @ np.linspace(2., 0.2, 9) + 0.05 * rng.normal(size=200)
    got = solve_profile_relative(A, b, radii, 0.1)
    Q, T = np.linalg.qr(A)
    L, srel, _ = op.relative_scale(radii, float(np.sum(A * A)), 0.1)
    ref = nnls(np.vstack([T, srel * L]), np.concatenate([Q.T @ b, np.zeros(7)]))[0]
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-12)
    assert np.allclose(solve_profile_relative(A, b, radii, 0.), nnls(A, b)[0], rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------- AZIM final model (host pieces)
def _hdr(n=201, pix=1.):
    from bbarolo2_amd.header import emptyheader
    return emptyheader([n, n, 8], [-pix / 3600., pix / 3600., 10.])


def test_face_on_density_uniform_map():
    """A uniform map of c per pixel: the surface density of every annulus is c / pixel area, and its face-on value c
    cos(inc) / pixel area, at any position angle."""
    from bbarolo2_amd.classic import face_on_density
    h = _hdr(201, 2.)
    im = np.full((201, 201), 3., np.float32)
    for inc, pa in ((0., 0.), (60., 30.), (45., 200.)):
        r = _inr([20., 40., 60., 80.], xpos=100., ypos=100., inc=inc, phi=pa)
        got = face_on_density(im, r, h)
        assert np.allclose(got, 3. / 4. * math.cos(math.radians(inc)), rtol=1e-6)


def test_face_on_density_overlapping_annuli():
    """Uneven radii 10, 20, 40: annuli [5, 15], [10, 30], [30, 50] (half the spacing to the next radius on both
    sides); points in two annuli count 1/2 in each, so on a uniform face-on map the densities are the covered area
    fractions: (75 + 125/2) / 200, (800 - 125/2) / 800 and 1, to the pixel sampling."""
    from bbarolo2_amd.classic import face_on_density
    h = _hdr(201, 0.5)
    im = np.ones((201, 201), np.float32)
    r = _inr([10., 20., 40.], xpos=100., ypos=100., inc=0., phi=0.)
    got = face_on_density(im, r, h) * 0.25
    assert np.allclose(got, [137.5 / 200., 737.5 / 800., 1.], rtol=2e-3)


def test_face_on_density_blanks_and_centre():
    """Blank (NaN) pixels are left out of the sums and of the areas; the centre must be inside the map."""
    from bbarolo2_amd.classic import face_on_density
    h = _hdr(101, 1.)
    im = np.full((101, 101), 2., np.float32)
    im[:50] = np.nan
    r = _inr([10., 20.], xpos=50., ypos=50., inc=0., phi=0.)
    assert np.allclose(face_on_density(im, r, h), 2.)
    with pytest.raises(ValueError):
        face_on_density(im, _inr([10., 20.], xpos=150., ypos=50.), h)


def test_pow10_to_unit():
    from bbarolo2_amd.classic import pow10_to_unit
    assert pow10_to_unit([0.003, 5.]) == float(np.float32(100.))      # (0.003 -> 0.3)
    assert pow10_to_unit([25.]) == float(np.float32(1.) / np.float32(10.))
    assert pow10_to_unit([0.5, 9.]) == 1.
