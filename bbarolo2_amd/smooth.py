"""The beam kernel of the model convolution (computed once per fit, on the host).

A Gaussian beam of FWHM axes (a, b) at position angle t has the second-moment matrix (in FWHM^2 units)
S = R(t) diag(a^2, b^2) R(t)^T, and convolving two Gaussians adds their matrices. The kernel that turns an old beam
into a new one is therefore the Gaussian of S_new - S_old, whose axes and angle are the eigenvalues and eigenvector
of a symmetric 2x2 matrix: the result of Wild (1970) that the reference uses (convpars, bbarolo/smooth.py:25-51). It
is sampled on the pixel grid out to where it falls to CUTOFFRATIO of its peak, cut there and normalised to unit sum
(fillgauss2d, smooth.py:62-88); column-density models are scaled by 1 / (G_old * K)(0) (gaussian_scalefactor,
smooth.py:91-101). The kernel of a cube is that of Smooth3D.defineBeam_Gaussian (smooth.py:203-231): from a point
source to the beam of the header, the angle relative to CROTA. The convolution itself runs on the GPU
(csrc/bb_conv.hip).
"""

import math

import numpy as np

from .header import arcsconv

CUTOFFRATIO = 1.0 / 10000.0
MAXCONV = 512 * 512
_LN2 = math.log(2.0)


class BeamError(ValueError):
    pass


def _moments(fwhm_major, fwhm_minor, pa_deg):
    """(sxx, syy, sxy) of a Gaussian beam, FWHM^2 units, the major axis at pa_deg from +X."""
    t = math.radians(pa_deg)
    c, s = math.cos(t), math.sin(t)
    a2, b2 = fwhm_major * fwhm_major, fwhm_minor * fwhm_minor
    return a2 * c * c + b2 * s * s, a2 * s * s + b2 * c * c, (a2 - b2) * c * s


def convpars(oldbeam, newbeam):
    """The Gaussian (bmaj, bmin, pa [deg from +X]) that convolves oldbeam into newbeam (same units), or None when
    newbeam is not wider than oldbeam in every direction."""
    n = _moments(*newbeam)
    o = _moments(*oldbeam)
    dxx, dyy, dxy = n[0] - o[0], n[1] - o[1], n[2] - o[2]
    half_tr = 0.5 * (dxx + dyy)
    rad = math.hypot(0.5 * (dxx - dyy), dxy)
    big, small = half_tr + rad, half_tr - rad
    scale = max(abs(n[0]) + abs(n[1]), 1e-30)
    if big < 0:
        return None
    if small < 0:
        if small > -1e-9 * scale:
            small = 0.
        else:
            return None
    pa = 0. if (dxy == 0 and dxx == dyy) else 0.5 * math.degrees(math.atan2(2. * dxy, dxx - dyy))
    return math.sqrt(big), math.sqrt(small), pa


def fillgauss2d(beam, gridspace, ampl=1.0, norm=True, cutoff=CUTOFFRATIO):
    """A rotated 2D Gaussian (bmaj, bmin, pa [deg from +X], in the units of gridspace) sampled on the pixel grid out
    to the cutoff (a fraction of its peak), zero below it; amplitude 1 (ampl without norm) or unit sum (norm).
    Returns an array (2 Y + 1, 2 X + 1), or None if it would exceed MAXCONV pixels."""
    t = beam[2] * math.pi / 180.
    c, s = math.cos(t), math.sin(t)
    fx, fy = abs(beam[0]), abs(beam[1])
    # half extent of the ellipse's axes at the cutoff, and the half sizes of the sampled rectangle
    reach = math.sqrt(-math.log(cutoff) / _LN2)
    ex, ey = 0.5 * fx * reach, 0.5 * fy * reach
    gx, gy = abs(gridspace[0]), abs(gridspace[1])
    nx = int(round(max(abs(ex * c), abs(ey * s)) / gx))
    ny = int(round(max(abs(ex * s), abs(ey * c)) / gy))
    if (2 * nx + 1) * (2 * ny + 1) > MAXCONV:
        return None
    ys, xs = np.meshgrid(np.arange(-ny, ny + 1) * gy, np.arange(-nx, nx + 1) * gx, indexing="ij")
    u = xs * c + ys * s              # along the major axis
    v = ys * c - xs * s              # along the minor axis
    qu = u / fx if fx != 0 else np.zeros_like(u)
    qv = v / fy if fy != 0 else np.zeros_like(v)
    g = np.exp(-4.0 * _LN2 * (qu * qu + qv * qv))
    g = np.where(g >= cutoff, g if norm else g * ampl, 0.)
    return g / g.sum() if norm else g


def gaussian_scalefactor(oldbeam_rot, kernel, gridspace):
    """1 / (G_old * kernel)(0): the old beam of unit amplitude convolved with the kernel at the centre, over the
    pixels where both are sampled."""
    g = fillgauss2d(oldbeam_rot, gridspace, 1.0, norm=False)
    hy = min(kernel.shape[0], g.shape[0]) // 2
    hx = min(kernel.shape[1], g.shape[1]) // 2
    gc = (g.shape[0] // 2, g.shape[1] // 2)
    kc = (kernel.shape[0] // 2, kernel.shape[1] // 2)
    total = 0.
    for dy in range(-hy, hy + 1):
        grow = g[gc[0] + dy, gc[1] - hx:gc[1] + hx + 1]
        krow = kernel[kc[0] - dy, kc[1] - hx:kc[1] + hx + 1][::-1]
        total += float(np.dot(grow, krow))
    return 1.0 / total if total != 0 else 1.0


def _as_major_first(b):
    """(bmaj, bmin, pa) with bmaj >= bmin: a beam given with its axes swapped is turned by 90 degrees."""
    return (b[0], b[1], b[2]) if b[0] >= b[1] else (b[1], b[0], b[2] + 90.)


def beam_kernel(h, oldbeam=(0., 0., 0.), newbeam=None):
    """The convolution kernel of the models of a cube: from oldbeam (default a point) to the beam of the header
    (arcsec, arcsec, deg). Returns (kernel (ky, kx) float64, scalefactor), or None if the kernel cannot be built."""
    to_as = arcsconv(h.Cunit(0))
    grid = (h.Cdelt(0) * to_as, h.Cdelt(1) * to_as)
    if newbeam is None:
        newbeam = (h.Bmaj() * 3600., h.Bmin() * 3600., h.Bpa())
    ob, nb = _as_major_first(oldbeam), _as_major_first(newbeam)
    if nb[0] < ob[0] or nb[1] < ob[1]:
        raise BeamError(f"the new beam ({nb[0]:g} x {nb[1]:g} arcsec) is smaller than the old beam "
                        f"({ob[0]:g} x {ob[1]:g} arcsec)")
    cb = convpars(ob, nb)
    if cb is None:
        raise BeamError(f"cannot calculate the convolution from the beam {ob} to the beam {nb}")
    # position angles of the header are from North; the grid's +X axis is 90 - CROTA from North
    rot = 90. - h.Crota()
    kernel = fillgauss2d((cb[0], cb[1], cb[2] + rot), grid, 1.0, norm=True)
    if kernel is None:
        return None
    return kernel, gaussian_scalefactor((ob[0], ob[1], ob[2] + rot), kernel, grid)
