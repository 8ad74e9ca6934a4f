"""The deconvolved density profile of the rings on the GPU: the densities whose beam-convolved model reproduces
the total-intensity map of the data, recomputed for every parameter vector when the geometry is free and DENS
is not (Bayesian3DFIT._set_density, bbarolo/bayesian/density.py:13-88, called at likelihood.py:595-596).

Per fit (the radii are fixed) the host prepares what does not depend on the parameters: the interpolation
weights of the ring densities in the sub-rings (DiskModel._ring_weights, galmod.py:1285-1314), the exterior
annulus (profile.py:45-64), the second-difference rows of the smoothness prior on the radii (profile.py:81-99,
141-150) and the weighted map pixels (profile.py:66-70). Per batch of parameter vectors the device deposits
the unit-density ring maps (bb_ring_maps_deposit / bb_ring_maps_finish), solves the non-negative profile
(bb_profile_solve) and writes the float32 densities into the sub-ring tables of the model
(bb_profile_set_dens): nothing returns to the host between the maps and the model.
"""

import ctypes as C
import math

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr, naz_of
from .galmod import SubringPlan, model_rings, outer_edge, nzq_for
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


class DensityProfile:
    """The deconvolved density profile of a fit on the current device.

    dm: the DiskModel (ANALYTIC); rings0: the rings of the fit (float32 values, as the ring table of
    Bayesian3DFIT); plan: the SubringPlan of the model (its radii and ring convention); box ((blo), (bhi)): the
    box of the maps (the likelihood box); ymap, smap, okmap (by, bx): the total-intensity map of the data, its
    noise and the usable pixels; smooth, corr, floor: as solve_profile; beam_arcsec: the beam FWHM (the width of
    the annulus is two of them)."""

    def __init__(self, dm, rings0, plan, box, ymap, smap, okmap, smooth=SMOOTH_DEFAULT, corr=1., floor=None,
                 beam_arcsec=None):
        torch = _torch()
        self.torch = torch
        self.dm = dm
        st = dm.setup
        self.nr = rings0.nr
        self.n = self.nr + 1
        (blo, bhi) = box
        self.blo, self.bhi = tuple(blo), tuple(bhi)
        self.by, self.bx = bhi[1] - blo[1], bhi[0] - blo[0]
        self.plan = plan
        W = plan.tables({"dens": np.eye(self.nr)})[0][:, 5, :].T
        self.widx, self.wval = weights_of(W)
        ext, self.rext = exterior_annulus(st, rings0, dm.ringmode, beam_arcsec)
        self.plan_ext = SubringPlan(st, ext, "CENTRED", final=True)
        We = self.plan_ext.tables({"dens": np.ones((1, 1))})[0][:, 5, :].T
        self.widx_e, self.wval_e = weights_of(We)
        self.radii = np.append(np.asarray(rings0.radii, float), self.rext)
        lcoef, lcol, near, norm, dmed = prior_rows(self.radii)
        ok = np.asarray(okmap, bool) & (np.asarray(smap) > 0) & np.isfinite(ymap)
        self.nok = int(ok.sum())
        if self.nok < 2 * self.nr:
            raise ValueError("3DFIT ERROR: too few pixels with data in the likelihood region to compute the density "
                             "profile.\n  Check the mask and the rings, or fit DENS.")
        with np.errstate(divide="ignore", invalid="ignore"):
            wpix = np.where(ok, 1. / np.asarray(smap, float), 0.).ravel()
            ypix = np.where(ok, np.asarray(ymap, float) / np.asarray(smap, float), 0.).ravel()
        self.smooth = float(smooth or 0.)
        d = _lib.bb_profile_desc(self.nr, self.by * self.bx, self.nok, 0, self.smooth, float(corr),
                                 FLOOR if floor is None else float(floor), norm, dmed)
        lcoef = np.ascontiguousarray(lcoef)
        lcol = np.ascontiguousarray(lcol)
        near = np.ascontiguousarray(near)
        wpix = np.ascontiguousarray(wpix)
        ypix = np.ascontiguousarray(ypix)
        out = C.c_void_p()
        prior = self.smooth > 0 and self.n >= 3
        _lib.check(_lib.lib().bb_profile_create(C.byref(d), wpix.ctypes.data, ypix.ctypes.data,
                                                lcoef.ctypes.data if prior else None,
                                                lcol.ctypes.data if prior else None,
                                                near.ctypes.data if prior else None, C.byref(out)))
        self.handle = out
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.d_widx = torch.from_numpy(self.widx).to(self.dev)
        self.d_wval = torch.from_numpy(self.wval).to(self.dev)
        self.d_widx_e = torch.from_numpy(self.widx_e).to(self.dev)
        self.d_wval_e = torch.from_numpy(self.wval_e).to(self.dev)
        # fixed-point unit: every contribution is at most the flux 2 pi R0 radsep / (2 naz) of its sub-ring
        rads = np.concatenate([plan.rads, self.plan_ext.rads])
        flux = [2 * math.pi * R * st.radsep / st.pixarea * st.pixarea / (naz_of(R, st.pixarea) * 2) for R in rads]
        self.qunit = max(max(flux, default=1.), 1e-300) * 2. ** -QBITS
        self.unit = self.qunit * float(np.mean(st.cd2i)) * float(np.sum(st.relint))
        self._bufs = {}

    def __del__(self):
        try:
            _lib.lib().bb_profile_destroy(self.handle)
        except Exception:
            pass

    def annulus_values(self, values):
        """The values of the annulus per parameter vector: those of the last ring (dict key -> (B, 1))."""
        return {k: np.asarray(v)[..., -1:] for k, v in values.items() if k in ANNULUS_KEYS}

    def _buffers(self, B, nsub_e, slot=0):
        """Device buffers of a batch of B vectors; slot: an independent set (batches in flight on two streams)."""
        b = self._bufs.get((B, slot))
        if b is None:
            torch = self.torch
            L = _lib.lib()
            P = B * self.n
            b = dict(counts=torch.empty((B, self.n, self.by, self.bx), dtype=torch.int64, device=self.dev),
                     maps=torch.empty((B, self.n, self.by, self.bx), dtype=torch.float64, device=self.dev),
                     x=torch.empty((B, self.n), dtype=torch.float64, device=self.dev),
                     dens=torch.empty((B, self.nr), dtype=torch.float32, device=self.dev),
                     work=torch.empty(int(L.bb_profile_work_size(self.handle, B)), dtype=torch.uint8,
                                      device=self.dev), P=P)
            self._bufs = {k: v for k, v in self._bufs.items() if k[0] == B}
            self._bufs[(B, slot)] = b
        return b

    def ext_tables(self, values, B):
        """Host sub-ring tables (B, 16, nsub_ext) of the annulus of B parameter vectors."""
        sr_ext = self.plan_ext.tables(self.annulus_values(values))[0]
        if sr_ext.shape[0] != B:
            sr_ext = np.ascontiguousarray(np.broadcast_to(sr_ext, (B,) + sr_ext.shape[1:]))
        return sr_ext

    def maps(self, sr, d_sr, sr_ext, stream=None, d_sr_ext=None, slot=0):
        """Device ring maps (B, nr+1, by, bx) float64 of B parameter vectors: sr / d_sr the host / device sub-ring
        tables of the model (their density row is not read), sr_ext (d_sr_ext) the host (device) tables of the
        annulus."""
        torch = self.torch
        L = _lib.lib()
        B = sr.shape[0]
        b = self._buffers(B, sr_ext.shape[2], slot)
        st = _stream_ptr(torch, stream)
        box = _lib.bb_box(self.blo[0], self.blo[1], self.bhi[0], self.bhi[1])
        for tab, dtab, widx, wval, map0, nw in ((sr, d_sr, self.d_widx, self.d_wval, 0, self.nr),
                                                (sr_ext, d_sr_ext, self.d_widx_e, self.d_wval_e, self.nr, 1)):
            nsub = tab.shape[2]
            if dtab is None:
                dtab = torch.from_numpy(np.ascontiguousarray(tab)).to(self.dev)
            nzq = nzq_for(tab.reshape(16, -1))
            h = self.dm.handle(nzq)
            p = self.dm.azplan(tab[0, 0], nzq)
            wb = int(L.bb_ring_maps_work_size(p, B))
            ws = torch.empty(wb, dtype=torch.uint8, device=self.dev)
            _lib.check(L.bb_ring_maps_deposit(h, p, box, dtab.data_ptr(), nsub, B, widx.data_ptr(), wval.data_ptr(),
                                              self.n, map0, nw, self.qunit, ws.data_ptr(), wb, b["counts"].data_ptr(),
                                              st))
        h = self.dm.handle(nzq_for(sr.reshape(16, -1)))
        _lib.check(L.bb_ring_maps_finish(h, b["counts"].data_ptr(), b["P"], self.by, self.bx, self.unit,
                                         int(self.dm.cfield is not None), b["maps"].data_ptr(), st))
        return b["maps"]

    def solve(self, maps, stream=None, slot=0):
        """x (B, nr+1) float64 device tensor: the profile of the rings and the annulus from the maps."""
        torch = self.torch
        B = maps.shape[0]
        b = self._buffers(B, 0, slot)
        st = _stream_ptr(torch, stream)
        _lib.check(_lib.lib().bb_profile_solve(self.handle, maps.data_ptr(), B, b["work"].data_ptr(),
                                               b["work"].numel(), b["x"].data_ptr(), st))
        return b["x"]

    def set_dens(self, x, d_sr, stream=None, slot=0):
        """float32 densities of the solutions x into row 5 of the device tables d_sr (B, 16, nsub) of the model;
        returns them (B, nr) (device)."""
        B = x.shape[0]
        b = self._buffers(B, 0, slot)
        _lib.check(_lib.lib().bb_profile_set_dens(x.data_ptr(), B, self.n, self.nr, self.d_widx.data_ptr(),
                                                  self.d_wval.data_ptr(), d_sr.shape[2], d_sr.data_ptr(),
                                                  b["dens"].data_ptr(), _stream_ptr(self.torch, stream)))
        return b["dens"]

    def run(self, sr, d_sr, values, stream=None, slot=0, d_sr_ext=None, sr_ext=None):
        """The profile of B parameter vectors written into row 5 of the device tables d_sr (B, 16, nsub) of the
        model; returns the float32 densities (B, nr) (device)."""
        if sr_ext is None:
            sr_ext = self.ext_tables(values, sr.shape[0])
        maps = self.maps(sr, d_sr, sr_ext, stream, d_sr_ext, slot)
        x = self.solve(maps, stream, slot)
        return self.set_dens(x, d_sr, stream, slot)
