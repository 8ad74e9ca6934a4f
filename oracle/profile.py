"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke and bench.py's CPU leg; never by
the product path). NumPy/SciPy FP64 restatement of the deconvolved density profile that Bayesian3DFIT computes
at every evaluation when the geometry is free and DENS is not:

  * unit-density ring maps: the ANALYTIC samples of every sub-ring deposited into one total-intensity map per
    ring through the linear interpolation weights of the ring densities, times mean(cd2i) * sum(relint), then
    convolved with the beam without the 1e-12 cut (bbarolo/galmod.py:1285-1345, 468-536);
  * the exterior annulus two beams wide beyond the outer edge of the disk (bbarolo/profile.py:45-64);
  * the noise-weighted system on the usable pixels (profile.py:33-70, bayesian/density.py:67-88);
  * solve_profile: least squares reduced to the n x n triangle of a QR factorisation, a first non-negative
    solution with the relative curvature prior, then the final one with curvature weights scaled by the local
    density, the misfit and the noise correlation area (profile.py:102-150).

Pinned in tests/test_oracle_profile.py by the reference's own checks (tests/test_bbarolo.py:90-140, 2946-2995,
4044-4056): the model's total-intensity map equals sum_r dens_r K_r, the annulus starts at the edge of the
disk, the profile of a noisy model cube recovers the true one, the limits with coinciding radii; and by
mathematics: the curvature rows return f'' = 2 for f = R^2 on any spacing and 0 for linear f, the NNLS
solutions satisfy the KKT conditions, noiseless data without prior return the true densities.
"""

import math

import numpy as np
from scipy.optimize import nnls

import oracle
from bbarolo2_amd.galmod import model_rings, outer_edge, nzq_for
from bbarolo2_amd.rings import Rings

SMOOTH_DEFAULT = 0.3      # width of the curvature prior (like_kwargs densprof_smooth)
FLOOR = 0.05              # lowest weight of a curvature row, as a fraction of the relative prior's


def interp_weights(setup, rings, mode):
    """The sub-ring densities as linear functions of the ring densities: (widx, wval) of shape (nsub, 2),
    ring index (-1 none) and weight (galmod.py:1285-1314). Built column by column from unit profiles."""
    cols = []
    for r in range(rings.nr):
        unit = rings.copy()
        unit.dens = np.eye(rings.nr)[r]
        cols.append(setup.subrings(model_rings(unit, mode, setup.radsep))[0][5])
    W = np.stack(cols, axis=1)
    if np.any((W != 0).sum(axis=1) > 2):
        raise RuntimeError("ring_maps: a sub-ring depends on more than two rings")
    widx = np.full((W.shape[0], 2), -1, np.int64)
    wval = np.zeros((W.shape[0], 2))
    for s, row in enumerate(W):
        nzr = np.flatnonzero(row)
        widx[s, :nzr.size], wval[s, :nzr.size] = nzr, row[nzr]
    return widx, wval


def ring_maps(setup, kern, scalefactor, rings, blo, bhi, mode="CENTRED", fft=False):
    """(nr, by, bx) FP64 total-intensity maps of the rings with unit density, beam-convolved when kern is
    given (DiskModel.ring_maps, galmod.py:1317-1345); fft: by FFT as the reference (else direct sums)."""
    units = model_rings(rings, mode, setup.radsep).copy()
    units.dens = np.ones(units.nr)
    sr = setup.subrings(units)[0]
    widx, wval = interp_weights(setup, rings, mode)
    maps = oracle.ring_maps_analytic(setup, sr, widx, wval, rings.nr, blo, bhi, nzq_for(sr))
    maps = maps * (float(np.mean(setup.cd2i)) * float(np.sum(setup.relint)))
    if kern is None:
        return maps
    conv = oracle.conv_fft if fft else oracle.conv_same
    yxr = conv(np.ascontiguousarray(np.moveaxis(maps, 0, -1)), kern, 1.0, thresh=False)
    return np.ascontiguousarray(np.moveaxis(yxr, -1, 0)) * scalefactor


def annulus(setup, rings, mode, beam_arcsec):
    """(ring, radius): the exterior annulus, centred half its width beyond the outer edge of the disk, with the
    geometry and kinematics of the last ring (profile.py:45-64)."""
    R = np.asarray(rings.radii, float)
    width = 2. * beam_arcsec if beam_arcsec else 2. * float(np.max(np.diff(R), initial=R[-1]))
    edges = [outer_edge(rings, mode)]
    if rings.edge is not None:
        edges.append(float(rings.edge))
    ea, er = setup.side_edges(rings)
    if ea is not None and er is not None:
        edges.append(max(ea, er))
    rext = min(edges) + 0.5 * width
    last = {k: getattr(rings, k)[-1] for k in ("xpos", "ypos", "vsys", "vrot", "vdisp", "z0", "inc", "phi",
                                              "z0thick", "fthick", "vdispthick")}
    ext = Rings().setRings(radii=[rext], dens=1., radsep=width, **last)
    ext.twolayer = rings.thick()
    ext.edgeapp = ext.edgerec = -1.
    return ext, rext


def system(setup, kern, scalefactor, rings, box, ymap, smap, okmap, beam_arcsec=None, mode="CENTRED", fft=False):
    """(A, b, ok, radii): the columns of the rings and of the annulus on the usable pixels, divided by the map
    noise, or None with fewer than 2 nr usable pixels (profile.py:33-70)."""
    blo, bhi = box
    ext, rext = annulus(setup, rings, mode, beam_arcsec)
    K = np.concatenate([ring_maps(setup, kern, scalefactor, rings, blo, bhi, mode, fft),
                        ring_maps(setup, kern, scalefactor, ext, blo, bhi, "CENTRED", fft)])
    ok = okmap & (smap > 0) & np.isfinite(ymap)
    if np.count_nonzero(ok) < 2 * rings.nr:
        return None
    inv = 1. / smap[ok]
    return K[:, ok].T * inv[:, None], ymap[ok] * inv, ok, np.append(np.asarray(rings.radii, float), rext)


def curvature(radii):
    """Second differences on the radii in order of radius (any spacing; coinciding radii separated by 1e-6 of
    the extent): (L (n-2, n) on the columns in the given order, order, sorted radii, tiny, dmin), with dmin a
    quarter of the smallest ring spacing, the last radius (the annulus) excluded (profile.py:81-99)."""
    R = np.asarray(radii, float)
    n = R.size
    order = np.argsort(R, kind="stable")
    Rs = R[order]
    tiny = 1e-6 * max(float(np.ptp(R)), 1e-300)
    dmin = max(0.25 * float(np.diff(np.sort(R[:-1])).min()), tiny) if n > 2 else tiny
    gaps = np.maximum(np.diff(Rs), tiny)
    L = np.zeros((max(n - 2, 0), n))
    for i in range(n - 2):
        a, b = gaps[i], gaps[i + 1]
        L[i, order[i:i + 3]] = (2. / (a + b)) * np.array([1. / a, -(1. / a + 1. / b), 1. / b])
    return L, order, Rs, tiny, dmin


def relative_scale(radii, trace_ata, alpha):
    """(L, s, order): the curvature rows and the weight s = sqrt(alpha trace(A^T A) / sum of the squared row
    coefficients), the spacings in that sum at least dmin (profile.py:141-150)."""
    L, order, Rs, tiny, dmin = curvature(radii)
    g = np.maximum(np.maximum(np.diff(Rs), tiny), dmin)
    a, b = g[:-1], g[1:]
    norm = float(np.sum((2. / (a + b)) ** 2 * (1. / a ** 2 + (1. / a + 1. / b) ** 2 + 1. / b ** 2)))
    return L, math.sqrt(alpha * trace_ata / max(norm, 1e-300)), order


def solve(A, b, radii, smooth=SMOOTH_DEFAULT, corr=1., floor=None):
    """The non-negative profile of profile.py:102-138 (all columns: rings, then the annulus)."""
    n = A.shape[1]
    Q, T = np.linalg.qr(A)
    c = Q.T @ b
    if not smooth or n < 3:
        return nnls(T, c)[0]
    L, srel, order = relative_scale(radii, float(np.sum(A * A)), 0.1)
    pad = np.zeros(n - 2)
    first = nnls(np.vstack([T, srel * L]), np.concatenate([c, pad]))[0]
    mean_rings = float(np.mean(first[:-1]))
    if not mean_rings > 0:
        return first
    misfit = float(np.sum((b - A @ first) ** 2)) / max(A.shape[0] - n, 1)
    inflate = max(float(corr), 1.) * max(misfit, 1.)
    d = float(np.median(np.diff(np.sort(np.asarray(radii, float)[:-1]))))
    local = np.maximum(first[order[1:-1]], 0.1 * mean_rings)
    w = np.maximum(math.sqrt(inflate) * d * d / (float(smooth) * local),
                   math.sqrt(FLOOR if floor is None else float(floor)) * srel)
    Rs = np.asarray(radii, float)[order]
    gaps = np.diff(Rs)
    w[np.minimum(gaps[:-1], gaps[1:]) < 0.25 * d] = srel
    return nnls(np.vstack([T, w[:, None] * L]), np.concatenate([c, pad]))[0]


def deconvolved(setup, kern, scalefactor, rings, box, ymap, smap, okmap, smooth=SMOOTH_DEFAULT, beam_arcsec=None,
                corr=1., mode="CENTRED", fft=False):
    """The densities of the rings (profile.py:188-197), or None with too few pixels."""
    s = system(setup, kern, scalefactor, rings, box, ymap, smap, okmap, beam_arcsec, mode, fft)
    if s is None:
        return None
    A, b, _, radii = s
    return solve(A, b, radii, smooth, corr)[:rings.nr]
