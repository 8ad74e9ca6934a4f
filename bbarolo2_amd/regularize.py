"""The regularisation of the geometry between the two stages of 3DFIT (Galfit.SecondStage and regularizeParams,
bbarolo/galfit.py:1300-1433): the fitted INC, PA, VSYS, XPOS, YPOS and Z0 of the rings replaced by their median, a
Bezier curve through them or a least-squares polynomial of the radius, chosen by REGTYPE or, by default, from the
scatter of INC and PA.

  * median: numpy's median of the values;
  * Bezier: the curve of degree n - 1 whose control points are the n (radius, value) pairs (Bernstein polynomials);
    at the radii when they are evenly spaced (the curve's points at uniform t), else at each radius by solving
    x(t) = radius for t (x(t) is increasing) by bisection;
  * polynomial of k coefficients: least squares on the columns of the Vandermonde matrix scaled to unit norm, the
    largest and the smallest value left out first when more than k + 2 values are given, a constant when there are
    not more values than coefficients;
  * REGTYPE: one word (median, bezier or a degree n: n + 1 coefficients) for INC and PA, or key=value words for any of
    inc, pa, vsys, xpos, ypos, z0; a parameter without a type (auto): INC and PA a constant with at most 4 fitted
    rings, else the median when the robust scatter (MADFM / 0.6744888) is at most 3 (deg or pix), else a line below 10
    fitted rings and a Bezier curve from 10; the others their median.
"""

import math

import numpy as np

MEDIAN, BEZIER, AUTO = -2, -1, -3
REG_KEYS = ("inc", "pa", "vsys", "xpos", "ypos", "z0")
MADFM_TO_SIGMA = 1. / 0.6744888          # (stats.py:8-12: the MADFM of a Gaussian is 0.6744888 sigma)


def _type(word):
    """The regularisation of a REGTYPE word: coefficients of the polynomial (degree + 1), BEZIER, MEDIAN or AUTO."""
    if word.isdigit():
        return 1 + int(word)
    return {"bezier": BEZIER, "median": MEDIAN}.get(word, AUTO)


def regtypes(regtype, inc, pa, good):
    """The regularisation of (inc, pa, vsys, xpos, ypos, z0) for REGTYPE and the first-stage INC and PA of the rings
    (good: the rings fitted): a list of 6 (type, chosen automatically)."""
    words = str(regtype).lower().split()
    kv = [w.split("=", 1) for w in words if "=" in w]
    reg = [AUTO] * 6
    if len(words) == 1 and not kv:
        reg[0] = reg[1] = _type(words[0])
    else:
        for k, v in kv:
            if k in REG_KEYS:
                reg[REG_KEYS.index(k)] = _type(v)
    good = np.asarray(good, bool)
    ngood = int(good.sum())
    out = []
    for i in range(6):
        if reg[i] != AUTO:
            out.append((reg[i], False))
            continue
        if i >= 2:
            out.append((MEDIAN, True))
        elif ngood <= 4:
            out.append((1, True))
        else:
            v = np.asarray(inc if i == 0 else pa, float)[good]
            med = np.median(v)
            spread = float(np.median(np.abs(v - med))) * MADFM_TO_SIGMA
            out.append(((2 if ngood < 10 else BEZIER) if spread > 3 else MEDIAN, True))
    return out


def polyfit(x, y, ncoeff):
    """(coefficients c_0..c_{k-1} of sum c_j x^j, ok): least squares on the unit-norm columns of the Vandermonde matrix;
    not ok when the columns are not independent or there are not more points than coefficients."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    if ncoeff < 1 or ncoeff >= x.size:
        return np.zeros(max(ncoeff, 1)), False
    V = x[:, None] ** np.arange(ncoeff)[None, :]
    s = np.sqrt((V * V).sum(axis=0))
    if not np.all(np.isfinite(s)) or np.any(s == 0):
        return np.zeros(ncoeff), False
    c, _, rank, _ = np.linalg.lstsq(V / s, y, rcond=None)
    if rank < ncoeff:
        return np.zeros(ncoeff), False
    return c / s, True


def _bernstein(n, t):
    """(len(t), n + 1) Bernstein polynomials b_{j,n}(t)."""
    t = np.asarray(t, float)[:, None]
    j = np.arange(n + 1)[None, :]
    binom = np.array([math.comb(n, k) for k in range(n + 1)], float)[None, :]
    return binom * t ** j * (1. - t) ** (n - j)


def bezier_uniform(x, y, ns=None):
    """The points of the Bezier curve of the control points (x, y) at ns uniform t in [0, 1] (default: as many as the
    points): (x(t), y(t)); the ends are the first and last points."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    n = x.size - 1
    ns = x.size if ns is None else ns
    t = np.arange(ns) / (ns - 1.) if ns > 1 else np.zeros(1)
    B = _bernstein(n, t)
    xo, yo = B @ x, B @ y
    xo[t == 0.], yo[t == 0.] = x[0], y[0]
    xo[t == 1.], yo[t == 1.] = x[-1], y[-1]
    return xo, yo


def bezier_at(x, y, xq):
    """The Bezier curve of the control points (x, y) (x increasing) at the abscissae xq (clipped to [x0, xn]): the t
    of each by bisection on the increasing x(t), then y(t)."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    xq = np.clip(np.atleast_1d(np.asarray(xq, float)), x[0], x[-1])
    n = x.size - 1
    lo, hi = np.zeros(xq.size), np.ones(xq.size)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        below = _bernstein(n, mid) @ x < xq
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return _bernstein(n, 0.5 * (lo + hi)) @ y


def regularize(x, y, rtype, xout=None):
    """(ok, the values of the regularised curve at the radii x, or at xout) for the values y at the radii x:
    MEDIAN, BEZIER or a polynomial of rtype coefficients (see the module docstring)."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    xo = x if xout is None else np.asarray(xout, float)
    if rtype == MEDIAN:
        return True, np.full(xo.size, float(np.median(y)))
    if rtype == BEZIER:
        d = np.diff(x)
        if d.size and np.allclose(d, d[0], rtol=1e-6):
            xb, yb = bezier_uniform(x, y)
            return True, (yb if xout is None else np.interp(xo, xb, yb))
        return True, bezier_at(x, y, xo)
    k = rtype if x.size > rtype else 1
    xx, yy = list(x), list(y)
    if x.size - 2 > k:
        # the largest and then the smallest value left out
        i = int(np.argmax(yy))
        del xx[i], yy[i]
        i = int(np.argmin(yy))
        del xx[i], yy[i]
    c, ok = polyfit(np.array(xx), np.array(yy), k)
    if not ok:
        return False, y
    return True, np.polynomial.polynomial.polyval(xo, c)
