"""Pins of the oracle's 3DFIT ring residual (oracle/norm.c, CPU): values worked out by hand on a one-pixel ring
for every ftype, normalisation, blank and support rule; the ring pixels against a count of the annulus; the
weights' mean over a ring; +inf without voxels."""
import math

import numpy as np
import pytest

import oracle

M = np.array([1., 2., 3.])
D = np.array([2., 0., 4.], np.float32)
NW = np.array([1., 2., 1.])


def _one_pixel(model, data, mask, **kw):
    """A 3x3 box whose only pixel inside the ring (r <= 0.5) is the centre."""
    m = np.zeros((3, 3, 3))
    m[1, 1] = model
    d = np.zeros((3, 3, 3), np.float32)
    d[1, 1] = data
    mk = np.zeros((3, 3, 3), np.uint8)
    mk[1, 1] = mask
    args = dict(blo=(0, 0), x0=1., y0=1., phi=0., inc=0., r1=0., r2=0.5, wpow=0., side=0, ftype=4,
                chan_noise=NW, data_noise=0.5, bweight=0., normtype=2)
    args.update(kw)
    return oracle.norm3dfit(m, d, mk, **args)


@pytest.mark.parametrize("ftype,total", [
    (4, 1. + 1.5 ** 2 / 2 + 1.),                                   # (m - o)^2 / n, obs = 0.5 where data <= 0
    (2, 1. + 1.5 / 2 + 1.),                                        # |m - o| / n
    (3, 1. / 3 + 1.5 / 2.5 / 2 + 1. / 7),                          # |m - o| / (m + o) / n
    (1, 1. / math.sqrt(2) + 1.5 ** 2 / math.sqrt(0.5) / 2 + 1. / 2)])   # (m - o)^2 / sqrt(o) / n
def test_ftypes(ftype, total):
    v, _, nb, nt = _one_pixel(M, D, [1, 1, 1], ftype=ftype)
    assert nb == 0 and nt == 3 and math.isclose(v, total / 3, rel_tol=1e-12)


def test_local_and_azim_normalisation():
    """LOCAL: the model x sum(positive masked data) / sum(model) = 6 / 12; AZIM: the same over the ring."""
    for nt_ in (0, 1):
        v, m, _, _ = _one_pixel(2 * M, D, [1, 1, 1], normtype=nt_)
        assert np.allclose(m[1, 1], M) and math.isclose(v, (1. + 1.5 ** 2 / 2 + 1.) / 3, rel_tol=1e-12)
    v, m, _, _ = _one_pixel(2 * M, D, [1, 1, 1], normtype=2)
    assert np.allclose(m[1, 1], 2 * M)


def test_blanks_and_support():
    """A model voxel outside the mask is a blank: f = (1 + Nb/N)^bweight sum / (N - Nb); with a support it counts
    only where the expected sub-clouds lam reach thr."""
    v, _, nb, nt = _one_pixel(M, D, [1, 1, 0], bweight=1.)
    assert nb == 1 and nt == 3 and math.isclose(v, (1 + 1 / 3) * 3.125 / 2, rel_tol=1e-12)
    pre = np.zeros((3, 3, 3))
    pre[1, 1] = [5., 5., 0.2]
    kw = dict(pre=pre, ilo=[0], ihi=[0], thr=np.full(3, 0.5), bweight=1.)
    v, _, nb, nt = _one_pixel(M, D, [1, 1, 0], **kw)
    assert nb == 0 and nt == 2 and math.isclose(v, 2.125 / 2, rel_tol=1e-12)
    pre[1, 1, 2] = 0.7
    v, _, nb, nt = _one_pixel(M, D, [1, 1, 0], **kw)
    assert nb == 1 and nt == 3
    # the footprint: rows y + ky//2 - j, columns x + ilo[j] .. x + ihi[j] of pre
    pre2 = np.zeros((3, 3, 3))
    pre2[0, 2, 2] = 0.7                       # row y - 1 (j = 2), column x + 1: in the footprint
    kw2 = dict(pre=pre2, ilo=[0, 0, 1], ihi=[0, 0, 1], thr=np.full(3, 0.5), bweight=1.)
    assert _one_pixel(M, D, [1, 1, 0], **kw2)[2] == 1
    pre2[:] = 0
    pre2[0, 0, 2] = 0.7                       # row y - 1, column x - 1: outside it
    assert _one_pixel(M, D, [1, 1, 0], **kw2)[2] == 0
    kw2["ilo"], kw2["ihi"] = [1, 0, 0], [1, 0, 0]      # (j = 0 is row y + 1)
    pre2[:] = 0
    pre2[2, 2, 2] = 0.7
    assert _one_pixel(M, D, [1, 1, 0], **kw2)[2] == 1


def test_ring_pixels_and_weights():
    """Face-on annulus: the pixels with 2 <= r <= 4 of the centre; |cos th| averages 2/pi over a ring, and a
    side keeps half of it; an inclined ring keeps the pixels of its ellipse."""
    n = 21
    m = np.ones((n, n, 1))
    d = np.zeros((n, n, 1), np.float32)
    mk = np.ones((n, n, 1), np.uint8)
    yy, xx = np.mgrid[0:n, 0:n]
    r = np.hypot(xx - 10., yy - 10.)
    args = dict(blo=(0, 0), x0=10., y0=10., phi=0., inc=0., r1=6., r2=9., side=0, ftype=2, chan_noise=[1.],
                data_noise=1e-9, bweight=0., normtype=2)
    v, _, _, nt = oracle.norm3dfit(m, d, mk, wpow=0., **args)
    assert nt == np.count_nonzero((r >= 6) & (r <= 9))
    v1, _, _, _ = oracle.norm3dfit(m, d, mk, wpow=1., **args)
    assert abs(v1 / v - 2 / math.pi) < 0.02
    _, _, _, nh = oracle.norm3dfit(m, d, mk, wpow=0., **dict(args, side=1))
    assert abs(nh - nt / 2) <= 8
    args_inc = dict(args, inc=60., r1=0., r2=8.)
    _, _, _, ne = oracle.norm3dfit(m, d, mk, wpow=0., **args_inc)
    assert abs(ne - math.pi * 8 * 8 * 0.5) < 0.05 * math.pi * 32


def test_no_voxels_is_inf():
    v, _, _, nt = _one_pixel(M, D, [1, 1, 1], r1=2., r2=3.)
    assert nt == 0 and v == math.inf
