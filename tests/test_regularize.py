"""The regularisation of the geometry between the 3DFIT stages (bbarolo2_amd/regularize.py, galfit.py:1300-1433),
pinned by closed forms: a median; a polynomial recovered exactly from its values, with the largest and smallest value
left out first; a constant when there are too few values; the quadratic Bezier curve's midpoint (y0 + 2 y1 + y2) / 4;
collinear control points give their line at any abscissa; the automatic choices of REGTYPE."""
import numpy as np

from bbarolo2_amd.regularize import regularize, regtypes, bezier_uniform, bezier_at, polyfit, MEDIAN, BEZIER


def test_median():
    y = np.array([3., 1., 7., 5.])
    ok, v = regularize(np.arange(4.), y, MEDIAN, xout=np.arange(6.))
    assert ok and np.array_equal(v, np.full(6, 4.))


def test_polynomial_exact_and_outliers_left_out():
    x = np.linspace(30., 300., 9)
    q = 60. + 0.05 * x - 1e-4 * x * x
    ok, v = regularize(x, q, 3)
    assert ok and np.allclose(v, q, rtol=0, atol=1e-9)
    # a line with its largest and smallest values spoiled: both left out, the line recovered
    y = 10. + 0.02 * x
    y[2], y[6] = 90., -40.
    ok, v = regularize(x, y, 2)
    assert ok and np.allclose(v, 10. + 0.02 * x, atol=1e-9)


def test_too_few_values_give_a_constant():
    ok, v = regularize(np.array([10., 20.]), np.array([40., 50.]), 3)
    assert ok and np.allclose(v, 45.)
    ok, v = regularize(np.array([10.]), np.array([40.]), 2)       # (one value, one coefficient: no fit)
    assert not ok and np.array_equal(v, [40.])
    c, ok = polyfit(np.array([1., 2., 3.]), np.array([2., 4., 6.]), 2)
    assert ok and np.allclose(c, [0., 2.], atol=1e-12)


def test_bezier_closed_forms():
    x, y = np.array([0., 1., 2.]), np.array([1., 5., 2.])
    xo, yo = bezier_uniform(x, y)
    assert np.allclose(xo, x) and np.allclose(yo, [1., (1. + 10. + 2.) / 4., 2.])
    ok, v = regularize(x, y, BEZIER)
    assert ok and np.allclose(v, yo)
    # evenly spaced radii: the curve at the radii by bisection is the curve at uniform t
    xe = np.arange(7.) * 30. + 15.
    ye = np.array([50., 70., 65., 80., 72., 75., 74.])
    assert np.allclose(bezier_at(xe, ye, xe), bezier_uniform(xe, ye)[1], atol=1e-9)
    # collinear control points at uneven radii: the curve is their line
    xu = np.array([10., 13., 30., 31., 60.])
    ok, v = regularize(xu, 3. - 0.5 * xu, BEZIER, xout=np.array([12., 25., 59.]))
    assert ok and np.allclose(v, 3. - 0.5 * np.array([12., 25., 59.]), atol=1e-9)


def test_regtypes():
    inc = np.array([60., 61., 59., 60.5, 60.2, 59.8])
    scattered = np.array([40., 60., 50., 70., 45., 65.])
    good = np.ones(6, bool)
    t = regtypes("auto", inc, inc, good)
    assert t[0] == (MEDIAN, True) and t[1] == (MEDIAN, True) and all(x == (MEDIAN, True) for x in t[2:])
    assert regtypes("auto", scattered, inc, good)[0] == (2, True)                 # (scatter > 3, below 10 rings)
    many = np.ones(12, bool)
    assert regtypes("auto", np.r_[scattered, scattered], np.r_[inc, inc], many)[0] == (BEZIER, True)
    few = np.array([True, True, True, False, False, True])
    assert regtypes("auto", scattered, inc, few)[:2] == [(1, True), (1, True)]    # (at most 4 fitted rings)
    assert regtypes("2", inc, inc, good)[:3] == [(3, False), (3, False), (MEDIAN, True)]
    t = regtypes("inc=bezier pa=1 vsys=0", inc, inc, good)
    assert t[:3] == [(BEZIER, False), (2, False), (1, False)] and t[3] == (MEDIAN, True)
