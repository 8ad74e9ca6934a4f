"""The deconvolved density profile of the rings on the GPU: the densities whose beam-convolved model reproduces
the total-intensity map of the data, recomputed for every parameter vector when the geometry is free and DENS
is not (Bayesian3DFIT._set_density, bbarolo/bayesian/density.py:13-88, called at likelihood.py:595-596).

Per fit (the radii are fixed) the host prepares what does not depend on the parameters: the interpolation
weights of the ring densities in the sub-rings (DiskModel._ring_weights, galmod.py:1285-1314), the exterior
annulus (profile.py:45-64), the second-difference rows of the smoothness prior on the radii (profile.py:81-99,
141-150) and the weighted map pixels (profile.py:66-70). Per batch of parameter vectors the device deposits
the unit-density ring maps (bb_ring_maps_deposit / bb_ring_maps_finish), solves the non-negative profile
(bb_profile_solve) and writes the float32 densities into the sub-ring tables of the model
(bb_profile_set_dens), inside the likelihood context (bb_loglike_launch): nothing returns to the host between the
maps and the model. ProfileSetup holds the per-fit constants that the context takes.
"""

import math

import numpy as np

from . import _lib
from .engine import naz_of
from .galmod import SubringPlan, outer_edge
from .rings import Rings

SMOOTH_DEFAULT = 0.3      # like_kwargs densprof_smooth (profile.py:72)
FLOOR = 0.05              # DENSPROF_FLOOR (profile.py:100)
QBITS = 38                # fixed-point resolution of the ring maps: 2^-38 of the largest sample flux

# the values of the last ring that the exterior annulus takes (profile.py:56-58)
ANNULUS_KEYS = ("xpos", "ypos", "vsys", "vrot", "vdisp", "z0", "inc", "phi", "z0thick", "fthick", "vdispthick")


def weights_of(W):
    """(widx, wval) (nsub, 2) int32 / float64 of a dense (nsub, nr) weight matrix with <= 2 rings per row."""
    if np.any((W != 0).sum(axis=1) > 2):
        raise RuntimeError("ring_maps: a sub-ring depends on more than two rings.")
    widx = np.full((W.shape[0], 2), -1, np.int32)
    wval = np.zeros((W.shape[0], 2))
    for s, row in enumerate(W):
        nzr = np.flatnonzero(row)
        widx[s, :nzr.size] = nzr
        wval[s, :nzr.size] = row[nzr]
    return widx, wval


def exterior_annulus(setup, rings, mode, beam_arcsec):
    """(Rings, radius) of the exterior annulus: two beams wide (or twice the largest ring spacing), centred half
    its width beyond the outer edge of the disk, with the values of the last ring (profile.py:45-64)."""
    R = np.asarray(rings.radii, float)
    width = 2. * beam_arcsec if beam_arcsec else 2. * float(np.max(np.diff(R), initial=R[-1]))
    edge = outer_edge(rings, mode)
    if rings.edge is not None:
        edge = min(edge, float(rings.edge))
    ea, er = setup.side_edges(rings)
    if ea is not None and er is not None:
        edge = min(edge, max(ea, er))
    rext = edge + width / 2.
    vals = {k: getattr(rings, k)[-1] for k in ANNULUS_KEYS}
    ext = Rings().setRings(radii=[rext], dens=1., radsep=width, **vals)
    ext.twolayer = rings.thick()
    ext.edgeapp = ext.edgerec = -1.
    return ext, rext


def prior_rows(radii):
    """The second differences of the profile at the radii (rings, then the annulus) in order of radius:
    (lcoef (n-2, 3), lcol (n-2, 3) int32, near (n-2,) int32, norm, dmed). Coinciding radii are 1e-6 of the
    extent apart; norm sums the squared coefficients with spacings of at least a quarter of the smallest ring
    spacing; near flags the rows across two radii closer than a quarter of the median ring spacing dmed
    (profile.py:81-150)."""
    R = np.asarray(radii, float)
    n = R.size
    order = np.argsort(R, kind="stable")
    Rs = R[order]
    tiny = 1e-6 * max(float(np.ptp(R)), 1e-300)
    dmin = max(0.25 * float(np.diff(np.sort(R[:-1])).min()), tiny) if n > 2 else tiny
    gap = np.maximum(np.diff(Rs), tiny)
    lcoef = np.zeros((max(n - 2, 0), 3))
    lcol = np.zeros((max(n - 2, 0), 3), np.int32)
    norm = 0.
    for i in range(n - 2):
        a, b = gap[i], gap[i + 1]
        lcoef[i] = (2. / (a + b)) * np.array([1. / a, -(1. / a + 1. / b), 1. / b])
        lcol[i] = order[i:i + 3]
        a, b = max(a, dmin), max(b, dmin)
        norm += float(np.sum(((2. / (a + b)) * np.array([1. / a, -(1. / a + 1. / b), 1. / b])) ** 2))
    dmed = float(np.median(np.diff(np.sort(R[:-1])))) if n > 2 else 0.
    raw = np.diff(Rs)
    near = (np.minimum(raw[:-1], raw[1:]) < 0.25 * dmed).astype(np.int32) if n > 2 else np.zeros(0, np.int32)
    return lcoef, lcol, near, norm, dmed


def solve_profile_relative(A, b, radii, smooth=0.1):
    """Non-negative least squares of A dens = b (the rings, then the exterior annulus) with the curvature prior of
    weight smooth relative to the data's (smooth x trace(A^T A) / the squared coefficients of the rows, prior_rows),
    through the QR of the stacked system (profile.py:153-185): the profile of the models around the ring fits of
    REVERSE (a host solve once per pass)."""
    from scipy.optimize import nnls
    A = np.asarray(A, float)
    b = np.asarray(b, float)
    n = A.shape[1]
    alpha = float(smooth or 0.)
    if alpha > 0 and n >= 3:
        lcoef, lcol, _, norm, _ = prior_rows(radii)
        L = np.zeros((n - 2, n))
        for i in range(n - 2):
            L[i, lcol[i]] = lcoef[i]
        L *= math.sqrt(alpha * float(np.sum(A * A)) / max(norm, 1e-300))
        A, b = np.vstack([A, L]), np.r_[b, np.zeros(n - 2)]
    q, r = np.linalg.qr(A)
    return nnls(r, q.T @ b)[0]


class ProfileSetup:
    """The per-fit constants of the deconvolved density profile (host arrays, handed to the likelihood context).

    dm: the DiskModel; rings0: the rings of the fit (float32 values, as the ring table of Bayesian3DFIT); plan: the
    SubringPlan of the model (its radii and ring convention); box ((blo), (bhi)): the box of the maps (the likelihood
    box); ymap, smap, okmap (by, bx): the total-intensity map of the data, its noise and the usable pixels; smooth,
    corr, floor: as solve_profile; beam_arcsec: the beam FWHM (the width of the annulus is two of them)."""

    def __init__(self, dm, rings0, plan, box, ymap, smap, okmap, smooth=SMOOTH_DEFAULT, corr=1., floor=None,
                 beam_arcsec=None):
        st = dm.setup
        self.nr = rings0.nr
        self.n = self.nr + 1
        (blo, bhi) = box
        self.blo, self.bhi = tuple(blo), tuple(bhi)
        self.by, self.bx = bhi[1] - blo[1], bhi[0] - blo[0]
        self.plan = plan
        W = plan.tables({"dens": np.eye(self.nr)})[0][:, 5, :].T
        self.widx, self.wval = weights_of(W)
        self.ext, self.rext = exterior_annulus(st, rings0, dm.ringmode, beam_arcsec)
        self.plan_ext = SubringPlan(st, self.ext, "CENTRED", final=True)
        We = self.plan_ext.tables({"dens": np.ones((1, 1))})[0][:, 5, :].T
        self.widx_e, self.wval_e = weights_of(We)
        self.radii = np.append(np.asarray(rings0.radii, float), self.rext)
        self.lcoef, self.lcol, self.near, self.norm, self.dmed = prior_rows(self.radii)
        ok = np.asarray(okmap, bool) & (np.asarray(smap) > 0) & np.isfinite(ymap)
        self.nok = int(ok.sum())
        if self.nok < 2 * self.nr:
            raise ValueError("3DFIT ERROR: too few pixels with data in the likelihood region to compute the density "
                             "profile.\n  Check the mask and the rings, or fit DENS.")
        with np.errstate(divide="ignore", invalid="ignore"):
            self.wpix = np.ascontiguousarray(np.where(ok, 1. / np.asarray(smap, float), 0.).ravel())
            self.ypix = np.ascontiguousarray(np.where(ok, np.asarray(ymap, float) / np.asarray(smap, float), 0.).ravel())
        self.smooth = float(smooth or 0.)
        self.corr = float(corr)
        self.floor = FLOOR if floor is None else float(floor)
        # fixed-point unit: every contribution is at most the flux 2 pi R0 radsep / (2 naz) of its sub-ring
        rads = np.concatenate([plan.rads, self.plan_ext.rads])
        flux = [2 * math.pi * R * st.radsep / st.pixarea * st.pixarea / (naz_of(R, st.pixarea) * 2) for R in rads]
        self.qunit = max(max(flux, default=1.), 1e-300) * 2. ** -QBITS
        self.unit = self.qunit * float(np.mean(st.cd2i)) * float(np.sum(st.relint))

    def prior(self):
        """True if the curvature prior is on (its rows are needed)."""
        return self.smooth > 0 and self.n >= 3

    def desc(self):
        return _lib.bb_profile_desc(self.nr, self.by * self.bx, self.nok, 0, self.smooth, self.corr, self.floor,
                                    self.norm, self.dmed)
