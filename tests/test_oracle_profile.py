"""Pins of the oracle's deconvolved density profile (CPU): the reference's own checks of ring_maps,
profile_system and solve_profile (tests/test_bbarolo.py:90-140, 2946-2995, 4044-4056), and closed forms."""
import numpy as np
import pytest
from scipy.optimize import nnls

import oracle
from oracle import profile as op
from bbarolo2_amd.galmod import GalmodSetup, GalmodPar, model_rings, nzq_for, PHI_TAB
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.smooth import beam_kernel
from conftest import small_rings


def model_cube(setup, kern, sf, rings, blo, bhi, mode="CENTRED"):
    """The final ANALYTIC model of the oracle in a box: (by, bx, nz), convolved, cut at 1e-12, scaled."""
    sr, _ = setup.subrings(model_rings(rings, mode, setup.radsep))
    m = oracle.galmod_analytic(setup, sr, blo, bhi, nzq_for(sr), PHI_TAB)
    return oracle.conv_same(m, kern, sf) if kern is not None else m


@pytest.mark.parametrize("mode,z0", [("CENTRED", 10.), ("NODES", 0.), ("SHIFTED", 20.)])
def test_ring_maps_sum_to_model_map(small_header, mode, z0):
    """sum_r dens_r K_r is the total-intensity map of the model, with the ring conventions, a thick disk and
    a warp, in a box (test_bbarolo.py:90-104)."""
    st = GalmodSetup(small_header, GalmodPar(GALMODTYPE="ANALYTIC"))
    kern, sf = beam_kernel(small_header)
    r = small_rings()
    r.z0[:] = z0
    r.dens = np.linspace(3., 0.5, r.nr)
    r.inc = np.linspace(50., 70., r.nr)
    blo, bhi = (5, 8), (60, 58)
    K = op.ring_maps(st, kern, sf, r, blo, bhi, mode)
    mod = model_cube(st, kern, sf, r, blo, bhi, mode).sum(axis=2)
    assert K.shape == (r.nr, 50, 55) and mod.max() > 0
    assert np.abs(np.tensordot(r.dens, K, axes=1) - mod).max() < 1e-6 * mod.max()


def test_annulus_starts_at_edge(small_header):
    """The exterior annulus is centred one beam beyond the outer edge of the disk: the CENTRED edge of the
    rings, or RADMAX inside the last ring (test_bbarolo.py:4044-4056)."""
    st = GalmodSetup(small_header, GalmodPar(GALMODTYPE="ANALYTIC"))
    r = small_rings(4)
    ones = np.ones((64, 64))
    Rc = op.system(st, None, 1., r, ((0, 0), (64, 64)), ones, ones, ones > 0, beam_arcsec=30.)[3]
    assert np.isclose(Rc[-1], 80. + 30.)
    r.edge = 62.
    Rc = op.system(st, None, 1., r, ((0, 0), (64, 64)), ones, ones, ones > 0, beam_arcsec=30.)[3]
    assert np.isclose(Rc[-1], 62. + 30.)


def test_curvature_rows_closed_form():
    """On any spacing the rows return the second derivative: 2 for R^2, 0 for a linear profile; columns in
    the given (unsorted) order."""
    rng = np.random.default_rng(3)
    R = np.sort(rng.uniform(0., 100., 9))
    perm = rng.permutation(9)
    L = op.curvature(R[perm])[0]
    assert L.shape == (7, 9)
    assert np.allclose(L @ (R[perm] ** 2), 2.)
    assert np.allclose(L @ (3. * R[perm] - 7.), 0., atol=1e-12)
    Lu = op.curvature(np.arange(6.) * 2.)[0]
    assert np.allclose(Lu[1, 1:4], np.array([1., -2., 1.]) / 4.)


def test_relative_scale_uniform():
    """Uniform spacing d: sum of squared coefficients = (n-2) * 6 / d^4, so s = sqrt(alpha tr d^4 / (6 (n-2)))."""
    R = np.arange(7.) * 3.
    _, s, _ = op.relative_scale(R, 5.0, 0.1)
    assert np.isclose(s, np.sqrt(0.1 * 5.0 * 3. ** 4 / (6. * 5)))


def _kkt(M, y, x, tol=1e-8):
    g = M.T @ (y - M @ x)
    scale = np.abs(M.T @ y).max()
    return np.all(g[x == 0] <= tol * scale) and np.all(np.abs(g[x > 0]) <= tol * scale)


def test_solve_noiseless_without_prior():
    """Noiseless data of a positive profile and no prior: the exact densities."""
    rng = np.random.default_rng(4)
    A = rng.random((60, 7))
    x = rng.uniform(0.5, 2., 7)
    s = op.solve(A, A @ x, np.arange(7.), smooth=0.)
    assert np.allclose(s, x, rtol=1e-10)


def test_solve_kkt_and_limits():
    """The final solution is the NNLS optimum of its stacked system (KKT), and the limits with coinciding or
    crossing radii of test_bbarolo.py:2946-2995."""
    rng = np.random.default_rng(1)
    A, b = rng.random((50, 4)), rng.random(50)
    lim = op.solve(A, b, [1., 2., 3., 3. + 1e-9])
    eq = op.solve(A, b, [1., 2., 3., 3.])
    assert np.all(np.isfinite(eq)) and np.allclose(eq, lim)
    rng = np.random.default_rng(1)
    A, b = rng.random((80, 6)), rng.random(80)
    R = [1., 2., 3., 4., 5.]
    s = op.solve(A, b, R + [4.])
    assert np.all(np.isfinite(s))
    assert np.allclose(s, op.solve(A, b, R + [4. + 1e-9])) and np.allclose(s, op.solve(A, b, R + [4. - 1e-9]))
    p = [0, 1, 2, 3, 5, 4]
    assert np.allclose(op.solve(A, b, R + [4.5]), op.solve(A[:, p], b, [1., 2., 3., 4., 4.5, 5.])[p])
    rng = np.random.default_rng(2)
    g = rng.random(80)
    A = g[:, None] + 0.05 * rng.random((80, 6))
    b = 3. * g + 0.05 * rng.random(80)

    def rough(e, smooth=0.1):
        return np.abs(np.diff(op.solve(A, b, R + [e], smooth)[:4], 2)).max()
    assert rough(6., 0.) > 0.3 and rough(6.) < 0.03 and rough(5.) < 0.03 and rough(4.) < 0.1
    # KKT of the first (relative prior) solution, rebuilt from its definition
    Rr = np.array(R + [6.])
    Q, T = np.linalg.qr(A)
    L, srel, _ = op.relative_scale(Rr, float(np.sum(A * A)), 0.1)
    M = np.vstack([T, srel * L])
    y = np.concatenate([Q.T @ b, np.zeros(4)])
    x0 = nnls(M, y)[0]
    assert _kkt(M, y, x0)


@pytest.mark.parametrize("smooth", [0., 0.1])
def test_profile_recovers_model_cube(smooth):
    """Rings narrower than the beam on a noisy model cube: the deconvolved profile is within 8% of the truth
    and the model with it fits the map (test_bbarolo.py:107-140; likemask 'none', map noise
    noiserms sqrt(nz), correlation area = beam area)."""
    h = emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., bunit="JY/BEAM", obj="MOCK")
    st = GalmodSetup(h, GalmodPar(GALMODTYPE="ANALYTIC"))
    kern, sf = beam_kernel(h)
    radii = np.arange(8) * 20. + 10.
    dens = 2. * np.exp(-radii / 50.)
    r = small_rings(8, dens=dens)
    m = model_cube(st, kern, sf, r, (0, 0), (64, 64))
    noise = np.random.default_rng(2).normal(0., m.max() / 300., m.shape)
    data = m + noise
    sig = m.max() / 300.
    ymap = data.sum(axis=2)
    smap = np.full((64, 64), sig * np.sqrt(40.))
    ok = np.ones((64, 64), bool)
    corr = np.pi * 30. * 30. / (4. * np.log(2.)) / 100.
    got = op.deconvolved(st, kern, sf, r, ((0, 0), (64, 64)), ymap, smap, ok, smooth=smooth, beam_arcsec=30.,
                         corr=corr)
    assert np.all(np.abs(got / dens - 1) < 0.08)
    r2 = small_rings(8, dens=got.astype(np.float32).astype(float))
    fit = model_cube(st, kern, sf, r2, (0, 0), (64, 64))
    chi2 = np.sum((data - fit) ** 2) / sig ** 2 / data.size
    assert chi2 < 1.05
