"""The per-fit set-up of the likelihood of Bayesian3DFIT from a cube, its mask, the initial rings and the priors:
the likelihood mask, the noise, the likelihood region and its box, as LikelihoodMixin sets them up
(bbarolo/bayesian/likelihood.py:83-128 _setup_likelihood / _setup_noise, 130-205 _setup_likemask, 294-355
_setup_likebox, 357-411 _side_region / _disk_region, 413-427 _z0_reach), handed to GpuLikelihood.

The passes over the cube run on the device: the dilation of the mask (noise.dilate_mask_device), the noise rms and
the correlation of the noise (noise.setup_noise), and the data of the likelihood with the squared-data sums of the
region (the likelihood context). The region is a geometric map of the sky plane (ny, nx): the pixels within the
projected outer disk of the rings, for every position angle and the lowest inclination the priors allow, widened by
3 Z0 + half a beam + the allowed shift of the centre (a Euclidean distance transform of the disk, scipy's, as the
reference computes it), within the union of the model boxes of the extreme centres, restricted to one half of the
galaxy with SIDE. It is computed once on the host from the rings alone.
"""

import math

import numpy as np

from .engine import _torch
from .galmod import model_box
from .likelihood import GpuLikelihood, float32_rings


def _bounds(free, lo, hi):
    """(upper, lower)(key, default): the ends of the prior of a free key (uniform priors lo, hi over theta), default
    when the key is not free (likelihood.py:310-322)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)

    def upper(key, default):
        return float(hi[free.index[key]].max()) if key in free.index else default

    def lower(key, default):
        return float(lo[free.index[key]].min()) if key in free.index else default

    return upper, lower


def z0_reach(rings, free, upper):
    """The largest scale height (arcsec) the priors allow: Z0's, and Z0THICK's with a thick layer (likelihood.py:413-427,
    fitter.py:479-481)."""
    z = upper("z0", float(np.max(rings.z0)))
    thick = "fthick" in free.index or bool(np.any(np.asarray(rings.fthick, float) > 0))
    if thick:
        z = max(z, upper("z0thick", float(np.max(rings.z0thick))))
    return z


def _outer_half_width(rings):
    """Half the width of the last ring (galmod.py:1055-1063): half the last spacing; a single ring is RADSEP wide, or
    extends from R/2 to 3R/2 without RADSEP."""
    R = np.asarray(rings.radii, float)
    if len(R) == 1:
        return rings.radsep / 2. if rings.radsep > 0 else R[0] / 2.
    return (R[-1] - R[-2]) / 2.


def disk_region(header, rings, blo, bhi, free, upper, lower):
    """The pixels of the box [blo, bhi) within the projected outer disk, widened (likelihood.py:373-411)."""
    from scipy.ndimage import distance_transform_edt
    pix = 0.5 * (abs(header.Cdelt(0)) + abs(header.Cdelt(1))) * 3600.
    radii = np.asarray(rings.radii, float)
    radsep = float(radii[-1] - radii[-2]) if len(radii) > 1 else 2. * _outer_half_width(rings)
    rout = float(radii[-1]) + radsep / 2.
    z0max = z0_reach(rings, free, upper)
    inc = np.asarray(rings.inc, float)
    incmin = min(float(inc.min()), lower("inc", float(inc.min())))
    incmin = min(max(incmin, 0.), 89.)
    phis = np.asarray(rings.phi, float)
    if "phi" in free.index:
        plo, phi_hi = lower("phi", float(phis.min())), upper("phi", float(phis.max()))
        if phi_hi - plo < 180.:
            phis = np.linspace(plo, phi_hi, 73)
        else:
            phis = np.array([0.])
            incmin = 0.                                 # (any orientation: a circle)
    xc, yc = float(rings.xpos[-1]), float(rings.ypos[-1])
    shift = max(abs(upper("xpos", xc) - xc), abs(lower("xpos", xc) - xc),
                abs(upper("ypos", yc) - yc), abs(lower("ypos", yc) - yc)) * pix
    yy, xx = np.mgrid[blo[1]:bhi[1], blo[0]:bhi[0]]
    dx, dy = (xx - xc) * pix, (yy - yc) * pix
    cinc = math.cos(math.radians(incmin))
    disk = np.zeros(xx.shape, bool)
    for ph in np.unique(phis):
        sp, cp = math.sin(math.radians(ph)), math.cos(math.radians(ph))
        xr = -dx * sp + dy * cp
        yr = (-dx * cp - dy * sp) / cinc
        disk |= xr * xr + yr * yr <= rout * rout
    if not disk.any():
        return np.ones(xx.shape, bool)
    margin = 3. * z0max + 0.5 * abs(header.Bmaj()) * 3600. + shift + 0.5 * pix
    return distance_transform_edt(~disk) * pix <= margin


def side_region(region, rings, side):
    """region restricted to the approaching ('A') or receding ('R') half of the galaxy: the side of the minor axis of
    the rings (median centre and position angle; the axis in both halves) (likelihood.py:357-371)."""
    side = str(side or "").strip().upper()[:1]
    if side not in ("A", "R"):
        return region
    xc, yc = float(np.median(rings.xpos)), float(np.median(rings.ypos))
    pa = float(np.median(np.unwrap(np.radians(np.asarray(rings.phi, float)))))
    ys, xs = region.shape
    yy, xx = np.mgrid[0:ys, 0:xs]
    xr = -(xx - xc) * np.sin(pa) + (yy - yc) * np.cos(pa)
    return region & (xr >= 0 if side == "R" else xr <= 0)


def likelihood_region(header, rings, free, lo, hi, ringmode="CENTRED", side=None):
    """(likebox ((blo), (bhi)), region (ny, nx) bool) of _setup_likebox (likelihood.py:294-355) for the initial
    rings (float32 ring values, as the reference keeps them), the free parameters (likelihood.FreePars) and their
    uniform priors (lo, hi over theta)."""
    upper, lower = _bounds(free, lo, hi)
    xs, ys = header.DimAx(0), header.DimAx(1)
    r0 = float32_rings(rings)
    rings = r0.copy()
    rings.z0 = np.array(rings.z0, float)
    rings.z0[-1] = z0_reach(r0, free, upper)
    xp, yp = float(rings.xpos[-1]), float(rings.ypos[-1])
    rings.xpos, rings.ypos = np.array(rings.xpos, float), np.array(rings.ypos, float)
    corners = []
    for x in (lower("xpos", xp), upper("xpos", xp)):
        for y in (lower("ypos", yp), upper("ypos", yp)):
            rings.xpos[-1], rings.ypos[-1] = x, y
            corners.append(model_box(header, rings, ringmode, final=True))
    blo = [max(int(min(c[0][0] for c in corners)), 0), max(int(min(c[0][1] for c in corners)), 0)]
    bhi = [min(int(max(c[1][0] for c in corners)), xs), min(int(max(c[1][1] for c in corners)), ys)]
    region = np.zeros((ys, xs), bool)
    region[blo[1]:bhi[1], blo[0]:bhi[0]] = disk_region(header, r0, blo, bhi, free, upper, lower)
    return (tuple(blo), tuple(bhi)), side_region(region, r0, side)


def likelihood_mask(header, mask, likemask="dilated", dilation=(1., 3)):
    """The voxels of the likelihood (_setup_likemask, likelihood.py:155-175) as a uint8 device tensor: the mask
    dilated by dilation[0] beam FWHM in each channel and +-dilation[1] channels ('dilated'), or every voxel ('none')."""
    from .noise import dilate_mask_device, _flags
    torch = _torch()
    mode = str(likemask).lower()
    if mode not in ("dilated", "none"):
        raise ValueError(f"ERROR! likemask must be 'dilated' or 'none', not {likemask!r}.")
    nbeam, nchan = float(dilation[0]), int(round(float(dilation[1])))
    if not (nbeam >= 0 and nchan >= 0):
        raise ValueError(f"ERROR! dilation must be >= 0: {dilation!r}.")
    shape = (header.DimAx(2), header.DimAx(1), header.DimAx(0))
    m = _flags(torch, mask, shape)
    if mode == "none":
        return torch.ones_like(m)
    bmaj, cd = header.Bmaj(), header.Cdelt(0)
    bpix = abs(bmaj / cd) if bmaj and cd else 1.
    return dilate_mask_device(m, nbeam * bpix, nchan)


class LikelihoodSetup:
    """The set-up of one fit: likemask (uint8 device tensor), likebox, region, the noise (noise.setup_noise's dict:
    noiserms, beam_area, noise_correlation, likefactor, likefactor2d, note), sigma (LIKEOPTS sigma, or the noise rms),
    fracerr, and what a free RADMAX is (LIKEOPTS radmax, radmax_width)."""

    def __init__(self, dm, data, mask, rings, free, lo, hi, options=None, side=None):
        from .noise import setup_noise
        from .params import LIKE_DEFAULTS
        from .noise import _cube, _flags
        torch = _torch()
        opts = dict(LIKE_DEFAULTS)
        opts.update(options or {})
        h = dm.header
        d = _cube(torch, data)                           # (the cube and the mask on the device once)
        m = _flags(torch, mask, d.shape)
        self.noise = setup_noise(h, d, m, opts["noisecorr"])
        self.likemask = likelihood_mask(h, m, opts["likemask"], opts["dilation"])
        del d, m
        self.likebox, self.region = likelihood_region(h, rings, free, lo, hi, dm.ringmode, side)
        self.sigma = float(opts["sigma"]) if opts.get("sigma") else self.noise["noiserms"]
        self.fracerr = opts.get("fracerr")
        self.densprof_smooth = opts["densprof_smooth"]
        self.radmax, self.radmax_width = opts.get("radmax") or "edge", opts.get("radmax_width")

    def likelihood(self, dm, data, rings, free, **kw):
        """The GpuLikelihood of the fit (kw: chunk, graphs, halves)."""
        lm = self.likemask.cpu().numpy().astype(bool)
        return GpuLikelihood(dm, data, rings, free, likemask=lm, region=self.region, likebox=self.likebox,
                             sigma=self.sigma, likefactor=self.noise["likefactor"],
                             likefactor2d=self.noise["likefactor2d"], fracerr=self.fracerr,
                             densprof_smooth=self.densprof_smooth, radmax=self.radmax,
                             radmax_width=self.radmax_width, **kw)


def setup_likelihood(dm, data, mask, rings, free, lo, hi, options=None, side=None, **kw):
    """GpuLikelihood of a cube (nz, ny, nx) float32 with the reference's per-fit set-up: (likelihood, LikelihoodSetup).
    options: LIKEOPTS (params.like_options: likemask, dilation, noisecorr, sigma, fracerr, densprof_smooth)."""
    s = LikelihoodSetup(dm, data, mask, rings, free, lo, hi, options, side)
    return s.likelihood(dm, np.ascontiguousarray(data, np.float32), rings, free, **kw), s
