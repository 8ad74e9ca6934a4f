"""3DFIT METHOD CLASSIC of a cube on the GPU: the first-stage fit of every ring and the final model
(Galfit.galfit and _fit_one_ring, bbarolo/galfit.py:1097-1260; finalModel with LOCAL normalisation,
bbarolo/galfit_out.py:128-159).

A stage fits all the rings together (RingFit.fit): every ring's downhill simplex advances one step per round and the
trial rings of all of them are modelled in one launch, each in its own box. As the reference does:
  * the data and their spread are scaled by 10 / max(data) during the fit (NORMALCUBE, galfit.py:1172-1181);
  * a ring with too few pixels in the mask is skipped, a fit that does not converge or whose VROT does not fit in the
    band of the cube is not ok (galfit.py:1118-1142, _max_vrot galfit.py:1085-1093);
  * after the last stage the rings that are not ok get DENS 0 (galfit.py:1250-1254);
  * the final model is the model of the fitted rings with the ring convention, normalised LOCAL (each spectrum scaled
    to the flux of the data in the mask, bb_local_norm) or not at all (NONE).
  * with TWOSTAGE and a free geometry, the second stage: the fitted INC, PA, VSYS, XPOS, YPOS, Z0 regularised
    (median, Bezier or polynomial: regularize.py) and fixed, the rings fitted again (galfit.py:1300-1393).
  * REVERSE (true, or auto for a median inclination above 75 degrees): the rings regridded to the beam's width if they
    are narrower (galfit.py:503-532), then after the first pass of each stage SWEEPS more passes from the outermost
    ring inwards, each ring fitted alone with the model of all the others around it (Gauss-Seidel: the rings outside
    at their values of this pass, those inside at those of the previous one; densities from the deconvolved profile
    of the previous pass), in the strip of the sky between the ring's radii along the major axis (RINGREGION STRIP)
    or its annulus (_sweep, _ring_context, _segment, _context, galfit.py:786-808, 1013-1074, 1263-1294).
  * the AZIM final model: the density of every fitted ring set to the face-on surface density of the data's
    intensity map in its annulus (face_on_density, the radial profile of Ellprof), scaled by a power of ten to 0.1-10
    and the model divided by it again (radialProfile and finalModel, galfit_out.py:86-159).

The mask: the reference builds it with its source finder (MASK SEARCH); search_mask is a simpler stand-in (the largest
connected region above a growth threshold holding a voxel above the detection threshold).
"""

import math

import numpy as np

from . import _lib
from .engine import DiskModel, _torch, _stream_ptr
from .galfit import (RingFit, Fit3DPar, free_flags, context_dens, INC, PA, Z0, XPOS, YPOS, VSYS, KEYS, MAXPAR,
                     CS_NONE, CS_OK, CS_FAILED)
from .galmod import GalmodPar, model_rings
from .header import arcsconv
from .rings import Rings


def robust_spread(data):
    """1.4826 x the median absolute deviation of the non-blank voxels (the cube's robust spread, stats.py:81-82)."""
    v = np.asarray(data, np.float64)
    v = v[np.isfinite(v) & (v != 0)]
    if v.size == 0:
        return 0.
    return float(1.4826 * np.median(np.abs(v - np.median(v))))


def beam_rings(inr, beam):
    """The rings of REVERSE when they are narrower than the beam (beam FWHM, arcsec): as wide as the beam out to the
    last radius, with the initial values interpolated at the new radii (the position angle unwrapped;
    galfit.py:516-532)."""
    R = np.asarray(inr.radii, float)
    nr = int(math.ceil(R[-1] / beam))
    rad = np.arange(nr) * beam + beam / 2.
    vals = {k: np.interp(rad, R, np.asarray(getattr(inr, k), float))
            for k in ("xpos", "ypos", "vsys", "vrot", "vdisp", "vrad", "vvert", "dvdz", "zcyl", "dens", "z0", "inc",
                      "z0thick", "fthick", "vdispthick")}
    vals["phi"] = np.interp(rad, R, np.degrees(np.unwrap(np.radians(np.asarray(inr.phi, float)))))
    r = Rings().setRings(radii=rad, radsep=beam, **vals)
    r.twolayer = inr.twolayer
    return r


def face_on_density(im, rings, h, subpix=2):
    """The face-on surface density of the map im (ny, nx) in the annulus of every ring (Ellprof.RadialProfile and
    getSurfDensFaceOn, ellprof.py:78-226): the annuli halfway to the next radius (the last as wide as the one before),
    around the mean centre of the rings, in the plane of each ring's inclination and position angle; every pixel
    sampled at subpix x subpix points (float32 offsets), a point inside several annuli counting 1/n in each; per
    ring the sum of the valid values over the points inside it, over their area (arcsec^2), times cos(inc) (an
    inclination with |cos| <= 1e-4 taken as 89 deg). Returns (nr,) float64."""
    im = np.asarray(im)
    nr = rings.nr
    R = np.asarray(rings.radii, float)
    W = np.empty(nr)
    for r in range(nr):
        if nr == 1:
            W[r] = rings.radsep if rings.radsep > 0 else R[0] / 2.
        elif r == nr - 1:
            W[r] = R[r] - R[r - 1]
        else:
            W[r] = R[r + 1] - R[r]
    ang = np.radians(-1.0 * (np.asarray(rings.phi, float) + 90.0))
    cp, sp = np.cos(ang), np.sin(ang)
    inc = np.asarray(rings.inc, float).copy()
    ci = np.cos(np.radians(inc))
    bad = np.abs(ci) <= 0.0001
    ci[bad] = math.cos(math.radians(89.))
    ann = np.stack([np.maximum(R - 0.5 * W, 0.), np.maximum(R + 0.5 * W, 0.)], 1).astype(np.float32)
    rmax = float(np.maximum(R + 0.5 * W, 0.).max())
    P = [float(np.mean(rings.xpos)), float(np.mean(rings.ypos))]
    ny, nx = im.shape
    if not (0 <= P[0] < nx and 0 <= P[1] < ny):
        raise ValueError("ELLPROF ERROR: The center of galaxy must be inside the map. Exiting...")
    dx = h.Cdelt(0) * arcsconv(h.Cunit(0))
    dy = h.Cdelt(1) * arcsconv(h.Cunit(1))
    adx, ady = abs(dx), abs(dy)
    x0, x1 = max(0, int(P[0] - rmax / adx - 1)), min(nx - 1, int(P[0] + rmax / adx + 1))
    y0, y1 = max(0, int(P[1] - rmax / ady - 1)), min(ny - 1, int(P[1] + rmax / ady + 1))
    ys, xs = np.mgrid[y0:y1 + 1, x0:x1 + 1]
    xs, ys = xs.ravel(), ys.ravel()
    v = im[ys, xs].astype(np.float32)
    Xr = (adx * (xs - P[0])).astype(np.float32)
    Yr = (ady * (ys - P[1])).astype(np.float32)

    def offsets(step, absd):
        o, pos = [], np.float32(0.5 * (step - absd))
        while pos < np.float32(0.5 * absd):
            o.append(float(pos))
            pos = np.float32(pos + np.float32(step))
        return np.array(o)
    ox, oy = offsets(adx / float(subpix), adx), offsets(ady / float(subpix), ady)
    OX, OY = np.meshgrid(ox, oy, indexing="ij")
    px = (Xr[:, None] + OX.ravel()[None]).astype(np.float32)
    py = (Yr[:, None] + OY.ravel()[None]).astype(np.float32)

    def inside(r):
        Xx = px * cp[r] - py * sp[r]
        Yy = (px * sp[r] + py * cp[r]) / ci[r]
        R2 = (Xx * Xx + Yy * Yy).astype(np.float32)
        return (ann[r, 0] * ann[r, 0] <= R2) & (R2 < ann[r, 1] * ann[r, 1])
    over = np.zeros(px.shape, np.int64)
    for r in range(nr):
        over += inside(r)
    val = np.broadcast_to(v[:, None], px.shape).astype(np.float64)
    val = np.where(over > 1, val / np.maximum(over, 1), val)
    valid = np.broadcast_to((~np.isnan(v))[:, None], px.shape)
    sub = float(len(ox) * len(oy))
    out = np.zeros(nr)
    for r in range(nr):
        good = inside(r) & valid
        area = good.sum() / sub * abs(dx * dy)
        out[r] = ci[r] * (val[good].sum() / sub / area) if area > 0 else 0.
    return out


def pow10_to_unit(vals):
    """The power of ten (float32 steps from 1) that brings the smallest of the positive values vals between 0.1 and 10
    (galfit_out.py:93-106); 1 with none."""
    vmin = min(vals) if len(vals) else np.finfo(np.float32).max
    s = np.float32(1.)
    while vmin < 0.1:
        vmin *= 10
        s *= np.float32(10)
    while vmin > 10:
        vmin /= 10
        s /= np.float32(10)
    return float(s)


def search_mask(data, snr=5., grow=3.):
    """A mask in place of the reference's source finder: the largest connected region (26-neighbours) of the voxels
    above grow sigma that holds a voxel above snr sigma (sigma: the robust spread)."""
    from scipy import ndimage
    sig = robust_spread(data)
    d = np.nan_to_num(np.asarray(data, np.float64))
    lab, n = ndimage.label(d > grow * sig, structure=np.ones((3, 3, 3), bool))
    if n == 0:
        raise ValueError("MASKING error: No sources detected in the datacube. Cannot build mask!!!")
    peaks = np.unique(lab[d > snr * sig])
    peaks = peaks[peaks > 0]
    if peaks.size == 0:
        raise ValueError("MASKING error: No sources detected in the datacube. Cannot build mask!!!")
    sizes = ndimage.sum(np.ones_like(lab), lab, index=peaks)
    return lab == peaks[int(np.argmax(sizes))]


class ClassicFit:
    """3DFIT CLASSIC of the cube (h, data (nz, ny, nx) float32) from the initial rings inr.

    par: Fit3DPar (FREE, FTYPE, WFUNC, NORM, ...); gpar: GalmodPar of the models (GALMODTYPE, LTYPE, ...); mask
    (nz, ny, nx) bool (default search_mask); normalcube: scale the data during the fit; twostage: the reference's
    TWOSTAGE (a second stage when the geometry is free); chan_noise (nz): the noise of each channel (default ones)."""

    def __init__(self, h, data, inr, par=None, gpar=None, mask=None, normalcube=True, twostage=True, chan_noise=None):
        self.torch = _torch()
        self.h = h
        self.par = par or Fit3DPar()
        self.gpar = gpar or GalmodPar()
        self.data = np.ascontiguousarray(data, np.float32)
        R = np.asarray(inr.radii, float)
        if np.any(np.diff(R) <= 0):
            raise ValueError("3DFIT ERROR: Radii not in increasing order.")
        if np.any(R < 0):
            raise ValueError("3DFIT ERROR: Negative radius!!!")
        self.inr = inr.copy()
        self.beam = abs(float(h.Bmaj())) * 3600. if h.Bmaj() else 0.        # (arcsec)
        rev = str(self.par.REVERSE).lower()
        self.reverse = "true" in rev or ("auto" in rev and float(np.median(self.inr.inc)) > 75.)
        if self.reverse and self.beam > 0 and self.inr.radsep < self.beam:
            self.inr = beam_rings(self.inr, self.beam)
        self.sweeps = max(int(self.par.SWEEPS), 1) if self.reverse else 0
        self.strip = self.reverse and str(self.par.RINGREGION).upper() == "STRIP"
        self.mask = search_mask(self.data) if mask is None else np.asarray(mask, bool)
        self.normalcube = bool(normalcube)
        self.twostage = bool(twostage)
        self.chan_noise = chan_noise
        self.dm = DiskModel(h, self.gpar, densflux=True)
        self.rings = None
        self.fitok = None
        self.results = None

    def _band_vrot(self, r, ir):
        """The largest VROT whose projection stays in the band of the cube for ring ir of r (_max_vrot)."""
        st = self.dm.setup
        ax = st.axis
        nz = st.nsubs
        v = ax.velocity(ax.world(np.array([0., nz - 1.])), ax.freq0) / 1000.
        vmin, vmax = abs(v[0] - r.vsys[ir]), abs(v[1] - r.vsys[ir])
        sini = math.sin(r.inc[ir] * math.pi / 180.)
        return max(vmax / sini, vmin / sini) if sini != 0 else np.inf

    def _stage(self, fit, stage):
        """One stage of the ring fits: (rings, ok per ring, results). The first pass fits all the rings together; with
        REVERSE, the passes after it (_sweep)."""
        out, res = fit.fit(stage=stage)
        ok = [bool(r is not None and r[0] and out.vrot[ir] <= self._band_vrot(out, ir)) for ir, r in enumerate(res)]
        cstat = np.array([CS_NONE if r is None else (CS_OK if o else CS_FAILED) for r, o in zip(res, ok)])
        for pas in range(1, self.sweeps + 1):
            res = self._sweep(fit, out, cstat, res, stage, pas)
            ok = [bool(c == CS_OK) for c in cstat]
        return out, ok, res

    def _sweep(self, fit, out, cstat, res, stage, pas):
        """One pass of REVERSE after the first (Galfit._sweep and _ring_context, galfit.py:1263-1294): the rings
        fitted again one at a time from the outermost inwards, each from its values of the previous pass and with the
        model of all the other rings on the device (those outside at their values of this pass, those inside at the
        previous pass's; the densities: the deconvolved profile of the previous pass scaled to the ring's own). out
        (the rings), cstat (their status) and the results are updated in place; returns the results."""
        nr = fit.inr.nr
        prev = out.copy()
        for ir in range(nr):
            if cstat[ir] != CS_OK:
                # (a ring not fitted starts again from its initial values)
                for k in KEYS.values():
                    getattr(prev, k)[ir] = getattr(fit.inr, k)[ir]
        prof = fit.relative_profile(prev, self.beam or None)
        sub = self.dm.setup.radsep * 60.
        res = list(res)
        for ir in range(nr - 1, -1, -1):
            dens = context_dens(prof, ir, prev.dens[ir])
            segs = []
            if ir > 0:
                segs.append(fit.segment(0, ir - 1, dens, out, cstat, sub))
            if ir < nr - 1:
                segs.append(fit.segment(ir + 1, nr - 1, dens, out, cstat, sub))
            cstat[ir] = CS_NONE
            st = fit.ring_start(ir, stage, start=prev, pas=pas)
            if st is None:
                res[ir] = None
                continue
            ctx = fit.context(segs, out) if segs else None
            if ctx is not None and self.strip:
                ctx.hmax = fit.strip_hmax(ir, st[0].radii[-1], self.beam)
            fit.ctx, fit.rng = ctx, st[1]
            try:
                r = fit.minimize(st[0])
            finally:
                fit.ctx = None
            res[ir] = r
            if not r[0]:
                cstat[ir] = CS_FAILED
                continue
            n = 0
            for k in range(MAXPAR):
                if fit.mpar[k]:
                    getattr(out, KEYS[k])[ir] = r[2][n]
                    n += 1
            cstat[ir] = CS_OK if out.vrot[ir] <= self._band_vrot(out, ir) else CS_FAILED
        return res

    def run(self):
        """Fit the rings: the first stage and, with TWOSTAGE and a free geometry, the second (the geometry of the
        first regularised, then fixed: SecondStage, galfit.py:1300-1393); set rings, fitok, results (those of the last
        stage), stage1 (rings, fitok, results of the first) and regularised (the regularisation used, or None);
        returns self."""
        m = free_flags(self.par.FREE)
        geometry = [k for k in (INC, PA, VSYS, XPOS, YPOS, Z0) if m[k]]
        scaling = 10. / float(self.data.max()) if self.normalcube else 1.
        data = (self.data * np.float32(scaling)).astype(np.float32) if self.normalcube else self.data
        noise = robust_spread(self.data) * scaling
        fit = RingFit(self.dm, data, self.mask, self.inr, par=self.par, data_noise=noise, chan_noise=self.chan_noise,
                      vsys_range=self._vsys_range())
        out, ok, res = self._stage(fit, 1)
        self.stage1 = (out.copy(), list(ok), res)
        self.regularised = None
        if self.twostage and geometry:
            inr2, how = self._regularise(out, ok, m)
            if inr2 is not None:
                self.regularised = how
                fit.inr = inr2
                fit.mpar = [v and k not in (INC, PA, VSYS, XPOS, YPOS, Z0) for k, v in enumerate(m)]
                fit.nfree = sum(fit.mpar)
                out, ok, res = self._stage(fit, 2)
        for ir in range(out.nr):
            if not ok[ir]:
                out.dens[ir] = 0.
        self.rings, self.fitok, self.results, self.nevals = out, ok, res, fit.nevals
        return self

    def _regularise(self, out, ok, m):
        """The initial rings of the second stage: the first stage's rings, the rings not fitted back at their initial
        VROT, VDISP, DENS and VRAD, the free geometry regularised (regularize.regtypes / regularize): (rings, [(key,
        type, auto)]), or (None, None) when a regularisation fails or no ring was fitted (the first stage is final)."""
        from .regularize import regtypes, regularize
        good = np.asarray(ok, bool)
        if not good.any():
            return None, None
        types = regtypes(self.par.REGTYPE, out.inc, out.phi, good)
        r2 = out.copy()
        for k in ("vrot", "vdisp", "dens", "vrad"):
            getattr(r2, k)[~good] = getattr(self.inr, k)[~good]
        how = []
        for (flag, key), (rtype, auto) in zip(((INC, "inc"), (PA, "phi"), (VSYS, "vsys"), (XPOS, "xpos"),
                                                 (YPOS, "ypos"), (Z0, "z0")), types):
            if not m[flag]:
                continue
            x, y = np.asarray(r2.radii, float), np.asarray(getattr(r2, key), float)
            if good.all():
                done, vals = regularize(x, y, rtype)
            else:
                done, vals = regularize(x[good], y[good], rtype, xout=x)
            if not done:
                return None, None
            setattr(r2, key, np.asarray(vals, float))
            how.append((key, rtype, auto))
        return r2, how

    def _vsys_range(self):
        st = self.dm.setup
        ax = st.axis
        v = ax.velocity(ax.world(np.array([0., st.nsubs - 1.])), ax.freq0) / 1000.
        return (float(v.min()), float(v.max()))

    def model(self, norm=None, rings=None):
        """The final model (nz, ny, nx) float32 of the fitted rings (or of rings): the model of the disk with the ring
        convention, convolved with the beam, normalised as norm (default the first NORM of par: LOCAL or NONE)."""
        torch = self.torch
        r = self.rings if rings is None else rings
        norm = (norm or (self.par.NORM.split() or ["LOCAL"])[0]).upper()
        norm = "AZIM" if norm == "AZIMUTHAL" else norm
        if norm not in ("LOCAL", "NONE", "AZIM"):
            raise ValueError(f"NORM {norm}: LOCAL, AZIM or NONE")
        dm, st = self.dm, self.dm.setup
        nz, ny, nx = self.data.shape
        sf = 1.
        if norm == "AZIM":
            r = r.copy()
            surf = face_on_density(self.intensity_map(), r, self.h)
            sf = pow10_to_unit([v for v in surf if np.isfinite(v) and v > 0])
            on = np.asarray(r.dens, float) > 0
            r.dens = np.where(on, sf * np.abs(surf), r.dens)
        sr, _ = st.subrings(model_rings(r, dm.ringmode, st.radsep))
        if st.galmodtype == "MC":
            m = dm.array_batch_mc(sr[None], (0, 0), (nx, ny), convolve=dm.cfield is not None)
        else:
            m = dm.array_batch(sr[None], (0, 0), (nx, ny), convolve=dm.cfield is not None)
        if norm == "LOCAL":
            dev = m.device
            d = torch.from_numpy(np.ascontiguousarray(self.data.transpose(1, 2, 0))).to(dev)
            k = torch.from_numpy(np.ascontiguousarray(self.mask.transpose(1, 2, 0)).astype(np.uint8)).to(dev)
            _lib.check(_lib.lib().bb_local_norm(m.data_ptr(), d.data_ptr(), k.data_ptr(), ny, nx, nz, m.data_ptr(),
                                                _stream_ptr(torch)))
        out = np.ascontiguousarray(m[0].cpu().numpy().transpose(2, 0, 1))
        return (out / np.float32(sf)).astype(np.float32) if norm == "AZIM" else out

    def intensity_map(self):
        """The total-intensity map (ny, nx) float32 of the data in the mask: the sum over the channels times the
        channel width (km/s), NaN where the sum is 0 (the moment-0 map without flux conversion, moments.py:86-110)."""
        from .header import delta_vel
        s = (self.data.astype(np.float64) * self.mask).sum(axis=0)
        m0 = s * abs(float(delta_vel(self.h)))
        return np.where(s == 0, np.nan, m0).astype(np.float32)
