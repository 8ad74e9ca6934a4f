"""The beam kernel of the model convolution (computed once per fit, on the host).

Follows the reference's beam algebra: ``convpars`` (Wild 1970; bbarolo/smooth.py:25-51), ``fillgauss2d``
(a rotated Gaussian sampled on the pixel grid, cut below 1e-4 of the peak and normalised to unit sum;
smooth.py:62-88), ``gaussian_scalefactor`` (smooth.py:91-101) and ``Smooth3D.defineBeam_Gaussian``
(smooth.py:203-231). The convolution itself runs on the GPU (csrc/bb_conv.hip).
"""

import math
import numpy as np

from .header import arcsconv

CUTOFFRATIO = 1.0 / 10000.0
MAXCONV = 512 * 512


class BeamError(ValueError):
    pass


def convpars(oldbeam, newbeam):
    """Convolution beam from oldbeam to newbeam (bmaj, bmin, pa in deg w.r.t. +X), or None."""
    d2r = math.atan(1.) / 45.
    a2, b2 = oldbeam[0] / 2, oldbeam[1] / 2
    a0, b0 = newbeam[0] / 2, newbeam[1] / 2
    D0 = a0 * a0 - b0 * b0
    D2 = a2 * a2 - b2 * b2
    th2, th0 = oldbeam[2] * d2r, newbeam[2] * d2r
    D1 = math.sqrt(max(D0 * D0 + D2 * D2 - 2 * D0 * D2 * math.cos(2 * (th0 - th2)), 0.))
    s = a0 * a0 + b0 * b0 - a2 * a2 - b2 * b2
    arg = 0.5 * (s + D1)
    if arg < 0:
        return None
    a1 = math.sqrt(arg)
    arg = 0.5 * (s - D1)
    if arg < 0:
        if arg > -1e-9 * max(a0 * a0, 1e-30):
            arg = 0.
        else:
            return None
    b1 = math.sqrt(arg)
    nom = D0 * math.sin(2 * th0) - D2 * math.sin(2 * th2)
    den = D0 * math.cos(2 * th0) - D2 * math.cos(2 * th2)
    th1 = 0. if (den == 0 and nom == 0) else math.atan2(nom, den) / 2
    return (2 * a1, 2 * b1, th1 / d2r)


def fillgauss2d(beam, gridspace, ampl=1.0, norm=True, cutoff=CUTOFFRATIO):
    """Rotated 2D Gaussian on the pixel grid: beam (bmaj, bmin, pa [deg w.r.t. +X]) in the units of
    gridspace. Returns an array (2 Ymax + 1, 2 Xmax + 1), or None if it is too large."""
    amplitude = 1.0 if norm else ampl
    phi = beam[2] * math.pi / 180.
    cs, sn = math.cos(phi), math.sin(phi)
    bm = (abs(beam[0]), abs(beam[1]))
    extend = math.sqrt(-1.0 * math.log(cutoff) / math.log(2.0))
    xr, yr = 0.5 * bm[0] * extend, 0.5 * bm[1] * extend
    x = max(abs(xr * cs), abs(-yr * sn))
    y = max(abs(xr * sn), abs(yr * cs))
    gs = (abs(gridspace[0]), abs(gridspace[1]))
    Xmax, Ymax = int(round(x / gs[0])), int(round(y / gs[1]))
    if (2 * Xmax + 1) * (2 * Ymax + 1) > MAXCONV:
        return None
    jj, ii = np.mgrid[-Ymax:Ymax + 1, -Xmax:Xmax + 1]
    xx, yy = ii * gs[0], jj * gs[1]
    xrr = xx * cs + yy * sn
    yrr = -xx * sn + yy * cs
    argX = xrr / bm[0] if bm[0] != 0 else 0 * xrr
    argY = yrr / bm[1] if bm[1] != 0 else 0 * yrr
    c = np.exp(-4.0 * math.log(2.0) * (argX * argX + argY * argY))
    c = np.where(c >= cutoff, c * amplitude, 0.)
    if norm:
        c = c / c.sum()
    return c


def gaussian_scalefactor(oldbeam_rot, confie, gridspace):
    """1 / (G_old * K)(0): the flux scaling of column-density models (smooth.py:91-101)."""
    g = fillgauss2d(oldbeam_rot, gridspace, 1.0, norm=False)
    ky, kx = confie.shape
    gy, gx = g.shape
    hy, hx = min(ky, gy) // 2, min(kx, gx) // 2
    cg = g[gy // 2 - hy:gy // 2 + hy + 1, gx // 2 - hx:gx // 2 + hx + 1]
    ck = confie[ky // 2 - hy:ky // 2 + hy + 1, kx // 2 - hx:kx // 2 + hx + 1]
    val = float(np.sum(cg * ck[::-1, ::-1]))
    return 1.0 / val if val != 0 else 1.0


def beam_kernel(h, oldbeam=(0., 0., 0.), newbeam=None):
    """The convolution kernel of the models of a cube (DiskModel._set_kernel -> Smooth3D.setfromCube ->
    defineBeam_Gaussian): from oldbeam (default a point) to the beam of the header (arcsec, arcsec, deg).
    Returns (kernel (ky, kx) float64, scalefactor), or None if the kernel cannot be built."""
    unittoarc = arcsconv(h.Cunit(0))
    gridspace = (h.Cdelt(0) * unittoarc, h.Cdelt(1) * unittoarc)
    crota = h.Crota()
    if newbeam is None:
        newbeam = (h.Bmaj() * 3600., h.Bmin() * 3600., h.Bpa())
    ob, nb = list(oldbeam), list(newbeam)
    for b in (ob, nb):
        if b[0] < b[1]:
            b[0], b[1], b[2] = b[1], b[0], b[2] + 90.
    if not (nb[0] >= ob[0] and nb[1] >= ob[1]):
        raise BeamError(f"the new beam ({nb[0]:g} x {nb[1]:g} arcsec) is smaller than the old beam "
                        f"({ob[0]:g} x {ob[1]:g} arcsec)")
    cb = convpars(ob, nb)
    if cb is None:
        raise BeamError(f"cannot calculate the convolution from the beam {ob} to the beam {nb}")
    cb = [cb[0], cb[1], cb[2] - (crota - 90)]
    conf = fillgauss2d(cb, gridspace, 1.0, norm=True)
    if conf is None:
        return None
    obr = (ob[0], ob[1], ob[2] + 90 - crota)
    return conf, gaussian_scalefactor(obr, conf, gridspace)
