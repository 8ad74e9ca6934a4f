"""3DFIT CLASSIC ring fits with the models and residuals on the GPU (bbarolo/galfit.py).

The host driver mirrors the reference's first-stage fit of each ring: the annulus of the ring (_make_dring,
galfit.py:1006-1029), the clamping and redraws of the trial values (func3D, galfit.py:832-889), the downhill simplex
(minimize and _mtry, galfit.py:891-1003), the pixel count that skips empty rings (_count_ring_pixels,
galfit.py:199-216), and the generator of each ring fit seeded [12345, stage, ring] (galfit.py:1097-1100). Every
evaluation is a device model of the ring in its own box (ANALYTIC or MC; bb_galmod_analytic / bb_galmod_mc, then the
beam, as DiskModel.array with final=False, galmod.py:1242-1271) and its residual (bb_norm3dfit, with the support of
ANALYTIC models, galfit.py:685-733). The points of a simplex that are independent (its first vertices, the
shrink) are evaluated as one batch.
"""

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr
from .header import arcsconv
from .rings import Rings

VROT, VDISP, DENS, Z0, INC, PA, XPOS, YPOS, VSYS, VRAD, MAXPAR = range(11)
NORMCODE = {"LOCAL": 0, "AZIM": 1, "NONE": 2}
SIDECODE = {"R": 1, "A": 2}
# expected number of MC sub-clouds below which a voxel of an ANALYTIC model is not a blank (galfit.py:191-194)
MC_SUPPORT_NMIN = math.exp(-0.5772156649015329)
KEYS = {VROT: "vrot", VDISP: "vdisp", DENS: "dens", Z0: "z0", INC: "inc", PA: "phi", XPOS: "xpos", YPOS: "ypos",
        VSYS: "vsys", VRAD: "vrad"}
# REVERSE: the status of each ring for the models of the rings around a ring fit (galfit.py:1010-1011): not fitted (or
# too few pixels in the mask), fitted, failed (not converged or rejected)
CS_NONE, CS_OK, CS_FAILED = 0, 1, 2
# the values a failed ring of a segment takes from the fitted ones (galfit.py:1030-1032)
SEG_KEYS = ("xpos", "ypos", "vsys", "vrot", "vdisp", "vrad", "vvert", "dvdz", "zcyl", "z0", "inc", "phi", "z0thick",
            "fthick", "vdispthick")


def context_dens(prof, ir, dens):
    """The densities of the rings in the model around ring ir whose own density is dens: the profile prof scaled to
    dens at ring ir (Galfit._ctx_dens, galfit.py:1069-1074; only the ratios matter with NORM LOCAL or AZIM)."""
    prof = np.asarray(prof, float)
    return dens * prof / max(prof[ir], 1.E-3 * float(prof.max()))


@dataclass
class Fit3DPar:
    """The 3DFIT parameters of the ring fits (the defaults of bbarolo/params.py:177-216)."""
    FREE: str = "VROT VDISP INC PA"
    DELTAVROT: float = 1000.0
    DELTAINC: float = 5.0
    DELTAPHI: float = 15.0
    MINVDISP: float = 0.1
    MAXVDISP: float = 1000.0
    FTYPE: int = 2
    WFUNC: int = 2
    TOL: float = 1e-3
    BWEIGHT: float = 1.0
    NORM: str = "AZIM"
    SIDE: str = "B"
    REGTYPE: str = "auto"
    REVERSE: str = "false"
    SWEEPS: int = 1
    RINGREGION: str = "STRIP"


def free_flags(free):
    """Galfit.setFree (galfit.py:658-674)."""
    f = free.lower()
    m = [False] * MAXPAR
    for key, idx in (("vrot", VROT), ("disp", VDISP), ("z0", Z0), ("inc", INC), ("xpos", XPOS), ("ypos", YPOS),
                     ("vsys", VSYS), ("vrad", VRAD)):
        if key in f:
            m[idx] = True
    if "pa" in f or "phi" in f:
        m[PA] = True
    if "all" in f:
        m = [True] * MAXPAR
    if not any(m):
        raise ValueError(" 3DFIT ERROR: NO free parameters!")
    return m


def count_ring_pixels(mask2d, blo, bhi, x0, y0, phi, inc, r1, r2):
    """(pixels of the annulus in the box, those with some mask) (galfit.py:199-216; the box edges included)."""
    F = math.pi / 180.
    sphi, cphi, cinc = math.sin(F * phi), math.cos(F * phi), math.cos(F * inc)
    ys, xs = np.mgrid[blo[1]:bhi[1] + 1, blo[0]:bhi[0] + 1]
    xx, yy = xs - blo[0], ys - blo[1]
    xr = -(xx - x0) * sphi + (yy - y0) * cphi
    yr = (-(xx - x0) * cphi - (yy - y0) * sphi) / cinc
    r = np.sqrt(xr * xr + yr * yr)
    ring = (r >= r1) & (r <= r2)
    inm = (ys < mask2d.shape[0]) & (xs < mask2d.shape[1])
    m = np.zeros(ring.shape, bool)
    m[inm] = mask2d[ys[inm], xs[inm]] > 0
    return int(ring.sum()), int((ring & m).sum())


class RingSimplex:
    """The host side of the first-stage 3DFIT of the rings: bounds, trial clamps, the simplex and the ring loop.
    The residuals come from values(rings) (a list of single-ring Rings -> array of residuals), which RingFit
    evaluates on the device.

    inr: the initial Rings; par: Fit3DPar; nx, ny: the cube's spatial size; mask2d (ny, nx): the number of masked
    channels of each pixel; pixscale: arcsec per pixel; box: rings -> (blo, bhi), the box of a ring's model;
    vsys_range: (min, max) of VSYS (the velocity range of the cube)."""

    def __init__(self, inr, par, nx, ny, mask2d, pixscale, box, vsys_range=(-1e9, 1e9)):
        self.par = par or Fit3DPar()
        self.inr = inr
        self.nx, self.ny = nx, ny
        self.mask2d = np.asarray(mask2d)
        self.pixscale = float(pixscale)
        self.box = box
        self.mpar = free_flags(self.par.FREE)
        self.nfree = sum(self.mpar)
        nm = (self.par.NORM.split() or ["LOCAL"])[0].upper()
        nm = "AZIM" if nm == "AZIMUTHAL" else nm
        self.norm = NORMCODE.get(nm, NORMCODE["LOCAL"])
        p = self.par
        self.maxs = np.zeros(MAXPAR)
        self.mins = np.zeros(MAXPAR)
        self.maxs[VROT] = inr.vrot.max() + p.DELTAVROT
        self.mins[VROT] = max(inr.vrot.min() - p.DELTAVROT, 0)
        self.maxs[VDISP], self.mins[VDISP] = p.MAXVDISP, p.MINVDISP
        self.maxs[Z0], self.mins[Z0] = 1000, 0.
        self.maxs[INC] = min(inr.inc.max() + p.DELTAINC, 90)
        self.mins[INC] = max(inr.inc.min() - p.DELTAINC, 0)
        self.maxs[PA] = min(inr.phi.max() + p.DELTAPHI, 360)
        self.mins[PA] = max(inr.phi.min() - p.DELTAPHI, -360)
        self.maxs[XPOS] = min(inr.xpos.max() + 10, nx)
        self.mins[XPOS] = max(inr.xpos.min() - 10, 0)
        self.maxs[YPOS] = min(inr.ypos.max() + 10, ny)
        self.mins[YPOS] = max(inr.ypos.min() - 10, 0)
        self.mins[VSYS], self.maxs[VSYS] = sorted(vsys_range)
        self.maxs[VRAD], self.mins[VRAD] = 500, -500
        self.rng = np.random.default_rng([12345, 0])
        self.nevals = 0

    def values(self, rings):
        raise NotImplementedError

    # ------------------------------------------------------------------ rings
    def make_dring(self, ir, rings=None):
        """Galfit._make_dring (galfit.py:1006-1029): the annulus halfway to the neighbouring radii."""
        inr = self.inr
        vr = inr if rings is None else rings
        R = np.asarray(inr.radii, float)
        if inr.nr == 1:
            w1 = w2 = inr.radsep / 2. if inr.radsep > 0 else R[0] / 2.
        elif ir == 0:
            w1 = w2 = (R[1] - R[0]) / 2.
        elif ir == inr.nr - 1:
            w1 = w2 = (R[ir] - R[ir - 1]) / 2.
        else:
            w1, w2 = (R[ir] - R[ir - 1]) / 2., (R[ir + 1] - R[ir]) / 2.
        w1, w2 = np.float32(w1), np.float32(w2)
        drads = [max(float(np.float32(R[ir] - w1)), 0.), max(float(np.float32(R[ir] + w2)), 0.)]
        vals = {k: getattr(vr, k)[ir] for k in ("xpos", "ypos", "vsys", "vrot", "vdisp", "vrad", "vvert", "dvdz", "zcyl",
                                                "dens", "z0", "inc", "phi", "z0thick", "fthick", "vdispthick")}
        d = Rings().setRings(radii=drads, **vals)
        d.twolayer = vr.thick()
        d.radsep = drads[1] - drads[0]
        d.id = ir
        return d

    def ring_edges(self):
        """The inner and outer edges (arcsec) of the annuli of the ring fits (Galfit._ring_edges, galfit.py:1005-1008)."""
        e = np.array([self.make_dring(ir).radii for ir in range(self.inr.nr)])
        return e[:, 0], e[:, 1]

    def segment(self, lo, hi, dens, rings, cstat, sub):
        """The rings lo..hi of rings as one model for the passes of REVERSE (Galfit._segment, galfit.py:1013-1046):
        nodes at the radii, linear in between, from the inner edge of the annulus of ring lo to the outer edge of that
        of ring hi (the end values kept to the edges), so that the segments and the annulus of a ring fit tile the
        disk. dens: the densities of all the rings; cstat: the status of every ring (CS_NONE rings have no emission,
        CS_FAILED ones take the values interpolated from the CS_OK ones); sub: the width of the model's sub-rings
        (arcsec). The inner part of a segment that does not start at the first ring is empty."""
        cstat = np.asarray(cstat)
        idx = np.arange(lo, hi + 1)
        seg = rings.subset(idx)
        seg.dens = np.where(cstat[idx] == CS_NONE, 0., np.asarray(dens, float)[idx])
        good = np.flatnonzero(cstat == CS_OK)
        bad = cstat[idx] == CS_FAILED
        if bad.any() and good.size:
            R = np.asarray(rings.radii, float)
            for k in SEG_KEYS:
                getattr(seg, k)[bad] = np.interp(seg.radii[bad], R[good], np.asarray(getattr(rings, k), float)[good])
        elo, ehi = self.ring_edges()
        nodes, rad = list(range(seg.nr)), list(seg.radii)
        if elo[lo] < rad[0] - sub:
            nodes.insert(0, 0)
            rad.insert(0, elo[lo])
        if ehi[hi] > rad[-1] + sub:
            nodes.append(seg.nr - 1)
            rad.append(ehi[hi])
        seg = seg.subset(np.array(nodes))
        seg.radii = np.array(rad, float)
        seg.empty = None if lo == 0 else True
        return seg

    def strip_hmax(self, ir, rout, beam_arcsec):
        """The half height (pixels) of the strip of ring ir whose annulus ends at rout (arcsec): the projected ring at
        its initial inclination plus three scale heights and half the beam (Galfit._region, galfit.py:760-778)."""
        inr = self.inr
        hm = (rout * abs(math.cos(math.radians(inr.inc[ir]))) + 3 * float(inr.layer_height()[ir]) + beam_arcsec / 2.)
        return hm / self.pixscale

    def strip_box(self, dring, blo, bhi, hmax):
        """The box of a ring grown to hold its strip (Galfit._strip_box, galfit.py:780-787)."""
        ext = int(math.ceil(max(dring.radii[-1] / self.pixscale, hmax))) + 1
        xpos, ypos = int(math.ceil(dring.xpos[-1])), int(math.ceil(dring.ypos[-1]))
        lo = [max(min(blo[0], xpos - ext), 0), max(min(blo[1], ypos - ext), 0)]
        hi = [min(max(bhi[0], xpos + ext + 1), self.nx), min(max(bhi[1], ypos + ext + 1), self.ny)]
        return lo, hi

    # ------------------------------------------------------------------ trials
    def clamp(self, dring, zpar, rng=None):
        """func3D's bounds (galfit.py:832-889): the values of a trial point, clamped or redrawn (with rng, default the
        fit's generator), written back into zpar; returns the dict of ring values."""
        p, inr, mins, maxs, m = self.par, self.inr, self.mins, self.maxs, self.mpar
        w = dring.id
        fran = (self.rng if rng is None else rng).random
        vals = {k: float(getattr(dring, k)[0]) for k in KEYS.values()}
        n = 0
        if m[VROT]:
            minv, maxv = max(inr.vrot[w] - p.DELTAVROT, mins[VROT]), min(inr.vrot[w] + p.DELTAVROT, maxs[VROT])
            z = zpar[n]
            v = maxv - fran() * p.DELTAVROT if z > maxv else (minv + fran() * p.DELTAVROT if z < minv else z)
            vals["vrot"] = zpar[n] = v
            n += 1
        if m[VDISP]:
            z = zpar[n]
            span = min(inr.vdisp[w], maxs[VDISP] - mins[VDISP])
            v = maxs[VDISP] - fran() * span if z > maxs[VDISP] else (mins[VDISP] + fran() * span
                                                                        if z < mins[VDISP] else z)
            vals["vdisp"] = zpar[n] = v
            n += 1
        if m[DENS]:
            z = zpar[n]
            vals["dens"] = zpar[n] = inr.dens[w] if z < 0 else z
            n += 1
        if m[Z0]:
            z = zpar[n]
            vals["z0"] = zpar[n] = inr.z0[w] if (z < mins[Z0] or z > maxs[Z0]) else z
            n += 1
        if m[INC]:
            lo, hi = max(inr.inc[w] - p.DELTAINC, mins[INC]), min(inr.inc[w] + p.DELTAINC, maxs[INC])
            z = zpar[n]
            vals["inc"] = zpar[n] = hi if z > hi else (lo if z < lo else z)
            n += 1
        if m[PA]:
            lo, hi = max(inr.phi[w] - p.DELTAPHI, mins[PA]), min(inr.phi[w] + p.DELTAPHI, maxs[PA])
            z = zpar[n]
            vals["phi"] = zpar[n] = hi if z > hi else (lo if z < lo else z)
            n += 1
        for idx, key in ((XPOS, "xpos"), (YPOS, "ypos"), (VSYS, "vsys"), (VRAD, "vrad")):
            if m[idx]:
                z = zpar[n]
                vals[key] = zpar[n] = getattr(inr, key)[w] if (z < mins[idx] or z > maxs[idx]) else z
                n += 1
        return vals

    def trials(self, dring, points, rng=None):
        """The ring of each point of a simplex (func3D's clamps in order, galfit.py:832-889)."""
        rings = []
        for zpar in points:
            vals = self.clamp(dring, zpar, rng)
            r = dring.copy()
            r.id = dring.id
            for k, v in vals.items():
                getattr(r, k)[:] = v
            rings.append(r)
        return rings

    def func3D_batch(self, dring, points):
        """func3D of several points of one ring (the clamps in order, then one batch of evaluations)."""
        return self.values(self.trials(dring, points))

    def func3D(self, dring, zpar):
        return float(self.func3D_batch(dring, [zpar])[0])

    # ------------------------------------------------------------------ minimisation
    def simplex(self, dring, rng=None):
        """The downhill simplex of Galfit.minimize / _mtry (galfit.py:891-1003) as a generator: it yields the list of
        trial rings it needs evaluated and receives their residuals (an array), so that the simplexes of all the rings
        advance together, one batch of evaluations per round; it returns (ok, minimum, pmin). rng: the generator of
        the clamps' redraws (the ring's own, seeded [12345, stage, ring] by the caller)."""
        TINY = 1.0e-10
        tol = self.par.TOL
        ndim = self.nfree
        m = self.mpar
        NMAX = 200 * ndim
        mpts = ndim + 1
        point, dels = [], []
        if m[VROT]:
            dels.append(30. if dring.vrot[0] > 40 else 5)
            point.append(dring.vrot[0])
        if m[VDISP]:
            v = dring.vdisp[0]
            dels.append(20. if v > 80 else (10. if 20. < v <= 80. else (0.2 if v <= 2 else 5)))
            point.append(v)
        if m[DENS]:
            dels.append(0.2 * dring.dens[0])
            point.append(dring.dens[0])
        if m[Z0]:
            dels.append(0.2 * dring.z0[0])
            point.append(dring.z0[0])
        if m[INC]:
            dels.append(5.)
            point.append(dring.inc[0])
        if m[PA]:
            dels.append(10.)
            point.append(dring.phi[0])
        for idx, key in ((XPOS, "xpos"), (YPOS, "ypos"), (VSYS, "vsys")):
            if m[idx]:
                dels.append((self.maxs[idx] - self.mins[idx]) / 4.)
                point.append(getattr(dring, key)[0])
        if m[VRAD]:
            dels.append(25)
            point.append(dring.vrad[0])
        p = np.tile(np.array(point, float), (mpts, 1))
        for i in range(1, mpts):
            p[i, i - 1] += dels[i - 1]
        y = np.array((yield self.trials(dring, [p[i].copy() for i in range(mpts)], rng)), float)
        nfunc = 0
        psum = p.sum(axis=0)

        def mtry(ihi, fac):
            # the reflection, expansion or contraction of the worst point through the face opposite it (_mtry)
            nonlocal psum
            fac1 = (1.0 - fac) / ndim
            fac2 = fac1 - fac
            ptry = psum * fac1 - p[ihi] * fac2
            # (the clamps write the values evaluated into ptry, which replaces the worst point: func3D's zpar,
            # galfit.py:964-973)
            ytry = float((yield self.trials(dring, [ptry], rng))[0])
            if ytry < y[ihi]:
                y[ihi] = ytry
                psum += ptry - p[ihi]
                p[ihi] = ptry
            return ytry

        while True:
            ilo = 0
            ihi, inhi = (0, 1) if y[0] > y[1] else (1, 0)
            for i in range(mpts):
                if y[i] <= y[ilo]:
                    ilo = i
                if y[i] > y[ihi]:
                    inhi = ihi
                    ihi = i
                elif y[i] > y[inhi] and i != ihi:
                    inhi = i
            rtol = 2.0 * abs(y[ihi] - y[ilo]) / (abs(y[ihi]) + abs(y[ilo]) + TINY)
            if rtol < tol:
                y[0], y[ilo] = y[ilo], y[0]
                p[[0, ilo]] = p[[ilo, 0]]
                return True, float(y[0]), p[0].copy()
            if nfunc >= NMAX:
                return False, float(y[ilo]), p[ilo].copy()
            nfunc += 2
            ytry = yield from mtry(ihi, -1.0)
            if ytry <= y[ilo]:
                ytry = yield from mtry(ihi, 2.0)
            elif ytry >= y[inhi]:
                ysave = y[ihi]
                ytry = yield from mtry(ihi, 0.5)
                if ytry >= ysave:
                    idx = [i for i in range(mpts) if i != ilo]
                    for i in idx:
                        p[i] = 0.5 * (p[i] + p[ilo])
                    # (clamped copies are evaluated; the simplex keeps the unclamped points, galfit.py:978-981)
                    y[idx] = np.asarray((yield self.trials(dring, [p[i].copy() for i in idx], rng)), float)
                    nfunc += ndim
                    psum = p.sum(axis=0)
            else:
                nfunc -= 1

    def minimize(self, dring):
        """The simplex of one ring evaluated on its own: (ok, minimum, pmin)."""
        gen = self.simplex(dring)
        try:
            req = next(gen)
            while True:
                req = gen.send(self.values(req))
        except StopIteration as e:
            return e.value

    # ------------------------------------------------------------------ the fit
    def ring_start(self, ir, stage=1, start=None, pas=0):
        """The annulus of ring ir (its values from the rings start, default the initial rings) and its generator
        (seeded [12345, stage, ring], and the pass of REVERSE after the first, galfit.py:1097-1100), or None when the
        ring has too few pixels in the mask (galfit.py:1118-1125)."""
        dring = self.make_dring(ir, start)
        blo, bhi = self.box(dring)
        ps = self.pixscale
        nTot, nIn = count_ring_pixels(self.mask2d, blo, bhi, dring.xpos[-1] - blo[0], dring.ypos[-1] - blo[1],
                                      dring.phi[-1], dring.inc[-1], dring.radii[0] / ps, dring.radii[-1] / ps)
        if nIn < 20 and (nIn == 0 or nIn < 0.7 * nTot):
            return None
        return dring, np.random.default_rng([12345, stage, ir] + ([pas] if pas else []))

    def fit_ring(self, ir, stage=1):
        """The first-stage fit of ring ir alone (Galfit._fit_ring / _fit_one_ring, galfit.py:1097-1142): None if the
        ring has too few pixels in the mask, else (ok, minimum, values of the free parameters)."""
        st = self.ring_start(ir, stage)
        if st is None:
            return None
        self.rng = st[1]
        return self.minimize(st[0])

    def values_rounds(self, groups):
        """The residuals of the trial rings of several rings, groups = [(ring index, [trial rings])]: a list of
        arrays. Here one ring after the other; RingFit evaluates all of them in one launch."""
        return [self.values(rings) for _, rings in groups]

    def fit(self, stage=1, rings=None):
        """Fit the rings (all, or the indices given) of one stage: every ring's simplex advances one step per round
        and the trial rings of all of them are evaluated together. Returns (Rings with the fitted values, the result
        of each ring: None when skipped, else (ok, minimum, values of the free parameters))."""
        out = self.inr.copy()
        idx = list(range(self.inr.nr)) if rings is None else list(rings)
        res = {ir: None for ir in idx}
        gens, req = {}, {}
        for ir in idx:
            st = self.ring_start(ir, stage)
            if st is None:
                continue
            gens[ir] = self.simplex(*st)
            req[ir] = next(gens[ir])
        self.begin_stage({ir: self.make_dring(ir) for ir in gens})
        try:
            while req:
                order = sorted(req)
                vals = self.values_rounds([(ir, req[ir]) for ir in order])
                for ir, v in zip(order, vals):
                    try:
                        req[ir] = gens[ir].send(v)
                    except StopIteration as e:
                        res[ir] = e.value
                        del req[ir]
        finally:
            self.begin_stage(None)
        for ir in idx:
            r = res[ir]
            if r is None or not r[0]:
                continue
            n = 0
            for k in range(MAXPAR):
                if self.mpar[k]:
                    getattr(out, KEYS[k])[ir] = r[2][n]
                    n += 1
        return out, [res[ir] for ir in idx]

    def begin_stage(self, drings):
        """Called with the annuli of the rings of a stage before its rounds (None after them)."""


class RingContext:
    """The model of the rings around a ring fit in a pass of REVERSE, on the device (Galfit._context and _Context,
    galfit.py:436-452, 786-808): conv (by, bx, nz) float32, the model of the segments around the ring convolved with
    the beam, in the box [blo, bhi); counts (by, bx, nz) float32, the expected numbers of Monte-Carlo sub-clouds of
    its non-convolved ANALYTIC model (the support of the ring's residual; None for MC models); grow (x, y): the box
    of a trial ring grows by the kernel's half sizes (ANALYTIC with a beam), within the context's box; hmax: the half
    height (pixels) of the ring's strip (0: the annulus, RINGREGION ANNULUS)."""

    def __init__(self, blo, bhi, conv, counts, grow=(0, 0), hmax=0.):
        self.blo, self.bhi, self.conv, self.counts, self.grow, self.hmax = list(blo), list(bhi), conv, counts, grow, hmax

    def ringbox(self, box):
        (rlo, rhi), (gx, gy) = box, self.grow
        blo = [max(rlo[0] - gx, self.blo[0]), max(rlo[1] - gy, self.blo[1])]
        bhi = [min(rhi[0] + gx, self.bhi[0]), min(rhi[1] + gy, self.bhi[1])]
        return blo, [max(bhi[0], blo[0] + 1), max(bhi[1], blo[1] + 1)]

    def slices(self, blo, bhi):
        return (slice(blo[1] - self.blo[1], bhi[1] - self.blo[1]), slice(blo[0] - self.blo[0], bhi[0] - self.blo[0]))


class RingFit(RingSimplex):
    """The first-stage 3DFIT of the rings of a cube, the residuals evaluated on the current device.

    dm: the DiskModel (DENSFLUX models, as 3DFIT sets); data (nz, ny, nx) float32; mask (nz, ny, nx) bool; inr: the
    initial Rings; par: Fit3DPar; data_noise: the spread of the data (obs where the data are <= 0); chan_noise (nz)
    (default ones); vsys_range: (min, max) of VSYS (the velocity range of the cube)."""

    def __init__(self, dm, data, mask, inr, par=None, data_noise=1.0, chan_noise=None, vsys_range=(-1e9, 1e9)):
        torch = _torch()
        nz, ny, nx = data.shape
        h = dm.header
        self._layout = None
        ps = (abs(h.Cdelt(0)) * arcsconv(h.Cunit(0)) + abs(h.Cdelt(1)) * arcsconv(h.Cunit(1))) / 2.
        mask = np.asarray(mask, bool)
        super().__init__(inr, par, nx, ny, mask.sum(axis=0), ps, lambda r: dm.box(r, final=False), vsys_range)
        self.torch = torch
        self.dm = dm
        self.nz = nz
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.d_data = torch.from_numpy(np.ascontiguousarray(np.asarray(data, np.float32).transpose(1, 2, 0))).to(self.dev)
        self.d_mask = torch.from_numpy(np.ascontiguousarray(mask.transpose(1, 2, 0)).astype(np.uint8)).to(self.dev)
        cn = np.ones(nz) if chan_noise is None else np.asarray(chan_noise, float)
        self.d_cn = torch.from_numpy(np.ascontiguousarray(cn, np.float64)).to(self.dev)
        self.data_noise = float(data_noise)
        self.ctx = None            # (the model of the rings around the ring being fitted: a RingContext, REVERSE)

    # ------------------------------------------------------------------ REVERSE
    def context(self, segs, rings):
        """The model of the segments segs (Rings, see segment) in the box of the rings, built on the device and
        convolved once (the model is linear in the rings): a RingContext, or None without segments (Galfit._context,
        galfit.py:786-808)."""
        torch = self.torch
        dm, st = self.dm, self.dm.setup
        blo, bhi = self.box(rings)
        analytic = st.galmodtype != "MC"
        mod = counts = None
        for seg in segs:
            sr, nv = st.subrings(seg)
            if not sr.shape[1]:
                continue
            if analytic:
                m = dm.array_batch(sr[None], blo, bhi, convolve=False)[0]
                unit = st.subcloud_flux(sr, nv)
                if unit > 0:
                    counts = m / unit if counts is None else counts + m / unit
            else:
                m = dm.array_batch_mc(sr[None], blo, bhi, convolve=False)[0]
            mod = m.clone() if mod is None else mod + m
        if mod is None:
            return None
        conv = dm.convolve_batch(mod[None])[0] if dm.cfield is not None else mod
        grow = (dm.cfield.shape[1] // 2, dm.cfield.shape[0] // 2) if (analytic and dm.cfield is not None) else (0, 0)
        return RingContext(blo, bhi, conv, counts, grow)

    def relative_profile(self, rings, beam_arcsec=None):
        """The deconvolved density profile of the data for the geometry of rings, the densities of the models around
        the ring fits of REVERSE (Galfit._profile, galfit.py:1048-1067): the beam-convolved unit-density maps of the
        rings and of the exterior annulus (ring maps on the device) against the total-intensity map of the data in
        the mask, in the box of the rings, solved with the relative smoothness prior 0.1 (solve_profile_relative);
        ones if there are too few pixels or no positive density."""
        from .profile import exterior_annulus, solve_profile_relative
        torch = self.torch
        dm, st = self.dm, self.dm.setup
        blo, bhi = self.box(rings)
        sl = (slice(blo[1], bhi[1]), slice(blo[0], bhi[0]))
        m = self.d_mask[sl]
        mom0 = (self.d_data[sl].double() * m.double()).sum(dim=2).cpu().numpy()
        nvox = m.sum(dim=2).cpu().numpy()
        rms = float(self.data_noise) * np.sqrt(nvox)
        K = dm.ring_maps(rings, blo, bhi)
        ext, rext = exterior_annulus(st, rings, dm.ringmode, beam_arcsec)
        K = np.concatenate([K, dm.ring_maps(ext, blo, bhi, ringmode="CENTRED")])
        ok = (nvox > 0) & (rms > 0) & np.isfinite(mom0)
        prof = None
        if ok.sum() >= 2 * rings.nr:
            A = K[:, ok].T / rms[ok, None]
            b = mom0[ok] / rms[ok]
            prof = solve_profile_relative(A, b, np.r_[np.asarray(rings.radii, float), rext], 0.1)[:rings.nr]
        return np.ones(rings.nr) if prof is None or not np.any(prof > 0) else np.asarray(prof, float)

    # ------------------------------------------------------------------ device evaluation
    def _support(self, unit):
        """ilo, ihi (kernel rows' footprint where one sub-cloud stays above the 1e-12 cut) and thr of
        Galfit._support_args (galfit.py:698-716), or None without an ANALYTIC support."""
        dm, st = self.dm, self.dm.setup
        if st.galmodtype != "ANALYTIC" or not unit > 0:
            return None
        cd = np.asarray(st.cd2i, float)
        if dm.cfield is not None:
            K = dm.cfield
            kmin = 1.E-12 / (unit * float(cd.max()) * float(dm.scalefactor))
            fp = K > max(kmin, 0.)
            ky, kx = K.shape
            hx = kx // 2
            anyr = fp.any(axis=1)
            i0 = np.where(anyr, fp.argmax(axis=1), 0)
            i1 = np.where(anyr, kx - 1 - fp[:, ::-1].argmax(axis=1), -1)
            ilo = np.where(anyr, hx - i1, 1).astype(np.int32)
            ihi = np.where(anyr, hx - i0, 0).astype(np.int32)
        else:
            ilo = np.zeros(1, np.int32)
            ihi = np.zeros(1, np.int32)
        return ilo, ihi, MC_SUPPORT_NMIN * unit * cd

    def begin_stage(self, drings):
        """The layout of a stage's rounds: the sub-rings of every ring's annulus, one block per ring, in one table whose
        radii (and azimuth plan) stay the same through the stage (None: the layout of each call's rings)."""
        self._layout = None
        if drings:
            st = self.dm.setup
            self._layout = {ir: st.subrings(d)[0] for ir, d in sorted(drings.items())}

    def values(self, rings):
        """The 3DFIT residual of each ring of the list (Galfit.getFuncValue without the models of the other rings,
        galfit.py:736-745), all evaluated in one launch."""
        groups, order = {}, []
        for i, r in enumerate(rings):
            groups.setdefault(r.id, []).append(i)
        out = np.empty(len(rings))
        keys = sorted(groups)
        vals = self.values_rounds([(k, [rings[i] for i in groups[k]]) for k in keys])
        for k, v in zip(keys, vals):
            out[groups[k]] = v
        return out

    def values_rounds(self, groups):
        """The residuals of the trial rings of several rings, groups = [(ring index, [trial rings])], in one launch:
        every trial is a table of the stage's layout whose density is zero outside its own ring's block, modelled in its
        own box (bb_galmod_analytic_boxes / bb_galmod_mc_boxes) within the union of the boxes and convolved there, then
        each ring's trials go through the 3DFIT residual (bb_norm3dfit) with that ring's support."""
        torch = self.torch
        dm, st = self.dm, self.dm.setup
        layout = self._layout
        if layout is None or any(ir not in layout for ir, _ in groups):
            layout = {ir: st.subrings(rs[0])[0] for ir, rs in groups}
        order = sorted(layout)
        offs, n = {}, 0
        for ir in order:
            offs[ir] = n
            n += layout[ir].shape[1]
        trials = [(ir, r) for ir, rs in groups for r in rs]
        B = len(trials)
        sr = np.empty((B, 16, n))
        units, vbox = [], np.empty((B, 4), np.int64)
        mc = st.galmodtype == "MC"
        for b, (ir, r) in enumerate(trials):
            own, nv = st.subrings(r)
            if own.shape[1] != layout[ir].shape[1]:
                raise RuntimeError("the sub-rings of a trial differ from those of its ring's annulus")
            # ANALYTIC: the blocks in the layout's order (one azimuth plan for the batch); MC: the trial's own block
            # first, so that its sub-rings keep the indices of the ring's own model (the keys of their random streams)
            blocks = ([ir] + [j for j in order if j != ir]) if mc else order
            a = 0
            for jr in blocks:
                w = layout[jr].shape[1]
                if jr == ir:
                    sr[b, :, a:a + w] = own
                else:
                    sr[b, :, a:a + w] = layout[jr]
                    sr[b, 5, a:a + w] = 0.
                a += w
            units.append(st.subcloud_flux(own, nv) if own.shape[1] else 0.)
            blo, bhi = self.box(r)
            if self.ctx is not None:
                # (REVERSE: the ring's box grown to hold its strip and the kernel, within the context's box)
                if self.ctx.hmax > 0:
                    blo, bhi = self.strip_box(r, blo, bhi, self.ctx.hmax)
                blo, bhi = self.ctx.ringbox((blo, bhi))
            vbox[b] = (blo[0], blo[1], bhi[0], bhi[1])
        ublo = (int(vbox[:, 0].min()), int(vbox[:, 1].min()))
        ubhi = (int(vbox[:, 2].max()), int(vbox[:, 3].max()))
        by, bx, nz = ubhi[1] - ublo[1], ubhi[0] - ublo[0], st.nsubs
        dep = torch.empty((B, by, bx, nz), dtype=torch.float32, device=self.dev)
        if st.galmodtype == "MC":
            mod = dm.array_batch_mc(sr, ublo, ubhi, convolve=dm.cfield is not None, work=dep, vbox=vbox)
        else:
            mod = dm.array_batch(sr, ublo, ubhi, convolve=dm.cfield is not None, work=dep, vbox=vbox)
        self.nevals += B
        ctx = self.ctx
        if ctx is not None:
            # the model of the rings around the ring added to each trial in its own box, and their sub-clouds to the
            # support of an ANALYTIC trial (getFuncValue with a context, galfit.py:746-758)
            if mod.data_ptr() == dep.data_ptr():
                mod = mod.clone()
            for b in range(B):
                lo, hi = vbox[b, :2], vbox[b, 2:]
                ys, xs = slice(lo[1] - ublo[1], hi[1] - ublo[1]), slice(lo[0] - ublo[0], hi[0] - ublo[0])
                cy, cx = ctx.slices(lo, hi)
                mod[b, ys, xs] += ctx.conv[cy, cx]
                if ctx.counts is not None and units[b] > 0:
                    dep[b, ys, xs] += float(units[b]) * ctx.counts[cy, cx]
        out, a = [], 0
        for ir, rs in groups:
            k = len(rs)
            out.append(self._norm_rows(mod, dep, a, rs, units[a:a + k], ublo, ubhi))
            a += k
        return out

    def _norm_rows(self, mod, dep, a, rings, units, blo, bhi):
        """bb_norm3dfit of the models a .. a + len(rings) of mod (the box [blo, bhi)): the trials of one ring, sharing
        the support of its unit (trials of different units are split)."""
        torch = self.torch
        st = self.dm.setup
        L = _lib.lib()
        B = len(rings)
        if st.galmodtype != "MC" and len(set(units)) > 1:
            return np.concatenate([self._norm_rows(mod, dep, a + i, [r], [u], blo, bhi)
                                   for i, (r, u) in enumerate(zip(rings, units))])
        sup = self._support(units[0]) if st.galmodtype != "MC" else None
        by, bx, nz = bhi[1] - blo[1], bhi[0] - blo[0], st.nsubs
        vox = by * bx * nz
        ps = self.pixscale
        ring = np.array([[r.xpos[-1] - blo[0], r.ypos[-1] - blo[1], r.phi[-1], r.inc[-1], r.radii[0] / ps,
                          r.radii[-1] / ps] for r in rings], np.float64)
        d_ring = torch.from_numpy(ring).to(self.dev)
        hmax = self.ctx.hmax if self.ctx is not None else 0.
        par = _lib.bb_norm_par(SIDECODE.get(self.par.SIDE, 0), int(self.par.FTYPE), self.norm, int(hmax > 0),
                               float(self.par.WFUNC), self.data_noise, float(self.par.BWEIGHT), float(hmax))
        wl = int(L.bb_norm_work_size(B, by, bx, nz, int(sup is not None)))
        work = torch.empty(wl, dtype=torch.float64, device=self.dev)
        val = torch.empty(B, dtype=torch.float64, device=self.dev)
        args = (None, None, None, 0, None)
        if sup is not None:
            ilo, ihi, thr = sup
            d_ilo = torch.from_numpy(ilo).to(self.dev)
            d_ihi = torch.from_numpy(ihi).to(self.dev)
            d_thr = torch.from_numpy(np.ascontiguousarray(thr, np.float64)).to(self.dev)
            args = (dep.data_ptr() + 4 * a * vox, d_ilo.data_ptr(), d_ihi.data_ptr(), len(ilo), d_thr.data_ptr())
        _lib.check(L.bb_norm3dfit(mod.data_ptr() + 4 * a * vox, B, by, bx, nz, self.d_data.data_ptr(),
                                  self.d_mask.data_ptr(), self.ny, self.nx, int(blo[0]), int(blo[1]), d_ring.data_ptr(),
                                  self.d_cn.data_ptr(), C.byref(par), *args, work.data_ptr(), wl, val.data_ptr(), None,
                                  None, _stream_ptr(torch)))
        return val.cpu().numpy()
