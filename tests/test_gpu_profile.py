"""GPU parity (MI355X) of the deconvolved density profile: ring maps, the non-negative profile and the likelihood
with a per-vector profile, through the C-ABI, against the FP64 oracle (oracle/profile.py).

Tolerances:
  ring maps (FP64 on the device, fixed-point sums at 2^-38 of a sample)   max |gpu - oracle| <= 1e-9 x max
  profile (normal equations on the device, QR in the oracle)              |gpu - oracle| <= 1e-7 x max |oracle|
                                                                          (1e-5 with coinciding radii)
  log-likelihood with the profile (float32 densities and model)          |gpu - oracle| <= 1e-5 x |S|
"""
import math

import numpy as np
import pytest

import oracle
from oracle import profile as op
from bbarolo2_amd.galmod import GalmodSetup, GalmodPar, model_rings, nzq_for, PHI_TAB
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.smooth import beam_kernel
from conftest import small_rings

pytestmark = pytest.mark.gpu


def _dm(h, **kw):
    from bbarolo2_amd.engine import DiskModel
    return DiskModel(h, **kw)


@pytest.mark.parametrize("mode,z0", [("CENTRED", 10.), ("NODES", 0.), ("SHIFTED", 20.)])
def test_ring_maps(small_header, mode, z0):
    dm = _dm(small_header, galmodtype="ANALYTIC", ringmode=mode)
    r = small_rings()
    r.z0[:] = z0
    r.inc = np.linspace(50., 70., r.nr)
    blo, bhi = (5, 8), (60, 58)
    got = dm.ring_maps(r, blo, bhi)
    kern, sf = dm.cfield, dm.scalefactor
    ref = op.ring_maps(dm.setup, kern, sf, r, blo, bhi, mode)
    print(f"ring maps {mode}: max rel {np.abs(got - ref).max() / np.abs(ref).max():.2e}", flush=True)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def _solve_gpu(A, b, radii, smooth, corr=1.):
    """bb_profile_solve on a given system: the maps are the columns of A, the pixel weights 1, the map b."""
    import ctypes as C
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.profile import prior_rows, FLOOR
    npix, n = A.shape
    lcoef, lcol, near, norm, dmed = prior_rows(radii)
    d = _lib.bb_profile_desc(n - 1, npix, npix, 0, smooth, corr, FLOOR, norm, dmed)
    w = np.ones(npix)
    y = np.ascontiguousarray(b, np.float64)
    lcoef, lcol, near = (np.ascontiguousarray(a) for a in (lcoef, lcol, near))
    h = C.c_void_p()
    L = _lib.lib()
    _lib.check(L.bb_profile_create(C.byref(d), w.ctypes.data, y.ctypes.data, lcoef.ctypes.data, lcol.ctypes.data,
                                   near.ctypes.data, C.byref(h)))
    try:
        maps = torch.from_numpy(np.ascontiguousarray(A.T)[None]).cuda()
        wb = int(L.bb_profile_work_size(h, 1))
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        x = torch.empty((1, n), dtype=torch.float64, device="cuda")
        _lib.check(L.bb_profile_solve(h, maps.data_ptr(), 1, work.data_ptr(), wb, x.data_ptr(), None))
        return x.cpu().numpy()[0]
    finally:
        L.bb_profile_destroy(h)


# (seed, shape of A, radii, smooth, tolerance). Coinciding radii (two rings, or the annulus on a ring's radius) put a
# curvature row of weight ~1/(1e-6 extent)^2 in the system: the normal equations square its conditioning, and the
# profile keeps 1e-5 (the rtol with which the reference's own test compares that limit,
# tests/test_bbarolo.py:2954-2956).
SYSTEMS = {
    "annulus_beyond": (1, (80, 6), [1., 2., 3., 4., 5., 6.5], 0.3, 1e-7),
    "annulus_inside": (1, (80, 6), [1., 2., 3., 4., 5., 4.5], 0.3, 1e-7),
    "annulus_on_ring": (1, (80, 6), [1., 2., 3., 4., 5., 4.], 0.3, 1e-5),
    "coinciding": (1, (50, 4), [1., 2., 3., 3.], 0.3, 1e-5),
    "no_prior": (4, (60, 7), list(np.arange(7.)), 0., 1e-7),
    "forty_rings": (9, (400, 41), list(np.arange(40) * 30. + 15.) + [1230.], 0.3, 1e-7),
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_profile_solve(name):
    seed, shape, radii, smooth, tol = SYSTEMS[name]
    rng = np.random.default_rng(seed)
    A, b = rng.random(shape), rng.random(shape[0])
    if name == "forty_rings":
        # a smooth positive profile under noise: most densities positive, the annulus near zero
        x = np.r_[2. * np.exp(-np.arange(40) / 12.), 0.]
        b = A @ x + 0.05 * rng.standard_normal(shape[0])
    ref = op.solve(A, b, radii, smooth, corr=2.)
    got = _solve_gpu(A, b, radii, smooth, corr=2.)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"profile {name}: max rel {err:.2e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("eps", [0., 1e-12, 1e-9])
def test_profile_solve_near_dependent_columns(eps):
    # two ring maps equal to working precision (eps = 0: identical): the solve gives a finite non-negative profile
    # whose fitted map is the oracle's (QR), instead of failing on a pivot at round-off
    rng = np.random.default_rng(12)
    A = rng.random((60, 6))
    A[:, 4] = A[:, 3] * (1. + eps * rng.standard_normal(60))
    x = np.array([1., 0.5, 2., 1., 1., 0.3])
    b = A @ x + 0.01 * rng.standard_normal(60)
    radii = [1., 2., 3., 4., 5., 6.5]
    for smooth in (0., 0.3):
        ref = op.solve(A, b, radii, smooth, corr=2.)
        got = _solve_gpu(A, b, radii, smooth, corr=2.)
        fr, fg = A @ ref, A @ got
        err = np.abs(fg - fr).max() / np.abs(fr).max()
        print(f"near-dependent eps {eps:g} smooth {smooth}: fitted map max rel {err:.2e}, x {np.round(got, 4)}",
              flush=True)
        assert np.all(np.isfinite(got)) and np.all(got >= 0)
        assert err <= 1e-6


def _profile_cube():
    """A noisy model cube of rings narrower than the beam (tests/test_bbarolo.py:107-140)."""
    h = emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., bunit="JY/BEAM", obj="MOCK")
    dm = _dm(h, galmodtype="ANALYTIC")
    st, kern, sf = dm.setup, dm.cfield, dm.scalefactor
    radii = np.arange(8) * 20. + 10.
    dens = 2. * np.exp(-radii / 50.)
    r = small_rings(8, dens=dens, z0=5.)
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    m = oracle.conv_same(oracle.galmod_analytic(st, sr, (0, 0), (64, 64), nzq_for(sr), PHI_TAB), kern, sf)
    sig = m.max() / 300.
    data = (m + np.random.default_rng(2).normal(0., sig, m.shape)).transpose(2, 0, 1).astype(np.float32)
    return dm, st, kern, sf, r, data, sig


def test_likelihood_with_profile():
    """GpuLikelihood with INC and PA free and DENS not fitted: the profile of every vector and the
    log-likelihood agree with the oracle (profile of the float32 rings, model with the float32 densities,
    residual over the whole cube)."""
    from bbarolo2_amd.likelihood import GpuLikelihood, FreePars, float32_rings
    dm, st, kern, sf, r, data, sig = _profile_cube()
    free = FreePars([("vrot", "single"), ("inc", "single"), ("phi", "single")], r.nr)
    like = GpuLikelihood(dm, data, r, free, sigma=sig, likefactor=1., densprof_smooth=0.3)
    theta = np.array([[150., 60., 45.], [140., 55., 50.], [160., 66., 40.]])
    ll = like.loglike(theta)
    dens_gpu = like.last_dens
    nz, ny, nx = data.shape
    ymap = data.astype(np.float64).sum(axis=0)
    smap = np.full((ny, nx), sig * math.sqrt(nz))
    ok = np.ones((ny, nx), bool)
    blo, bhi = like.blo, like.bhi
    for i, th in enumerate(theta):
        rr = float32_rings(r)
        for k, v in free.values(th[None]).items():
            setattr(rr, k, v[0])
        d = op.deconvolved(st, kern, sf, rr, ((0, 0), (nx, ny)), ymap, smap, ok, smooth=0.3, beam_arcsec=30.)
        assert np.abs(dens_gpu[i] - d).max() <= 1e-6 * np.abs(d).max(), (i, dens_gpu[i], d)
        rr.dens = d.astype(np.float32).astype(np.float64)
        sr, _ = st.subrings(model_rings(rr, "CENTRED", st.radsep))
        m = oracle.conv_same(oracle.galmod_analytic(st, sr, blo, bhi, nzq_for(sr), PHI_TAB), kern, sf)
        dbox = data[:, blo[1]:bhi[1], blo[0]:bhi[0]]
        S = oracle.sumsq_residuals(np.ascontiguousarray(m.transpose(2, 0, 1)), dbox, np.ones(dbox.shape, bool), 2,
                                   np.ones(dbox.shape[1:], bool))
        S += like.ss_out
        ref = -(0.5 * S / sig ** 2 + 0.5 * like.nvox * math.log(2 * math.pi * sig ** 2))
        print(f"loglike {i}: gpu {ll[i]:.6f} oracle {ref:.6f}", flush=True)
        assert abs(ll[i] - ref) <= 1e-5 * 0.5 * S / sig ** 2
