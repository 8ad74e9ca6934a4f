"""CPU: the host side of the edges of the disk as free parameters of the likelihood (RADMAX, RADMAXA, RADMAXR;
_update_rings, bayesian/fitter.py:1109-1136; _zero_reason, bayesian/likelihood.py:547-559).

  * FreePars takes the edges as single values that are not ring values, at the context's keys, and flags a negative
    RADMAX (only) as a zero likelihood;
  * the oracle's rings of a vector carry the edges as floats (not float32) and a negative RADMAX is a zero;
  * the prior's rows with the exterior annulus inside the rings (an edge inside the disk: its radius sorts between
    the rings' radii), as the device builds them per vector (prior_rows), equal the oracle's curvature rows and the
    sum of their squared coefficients that scales the relative prior (profile.py:81-150).
"""
import math

import numpy as np

from oracle import profile as op
from oracle import pipeline
from bbarolo2_amd import _lib
from bbarolo2_amd.galmod import GalmodPar, LIKE_KEYS
from bbarolo2_amd.header import emptyheader
from bbarolo2_amd.likelihood import FreePars, EDGE_KEYS
from bbarolo2_amd.profile import prior_rows
from bbarolo2_amd.rings import Rings


def test_free_edges_are_single_values():
    spec = [("vrot", "ring"), ("radmax", "single"), ("radmaxa", "single"), ("radmaxr", "single")]
    f = FreePars(spec, 4)
    assert f.ndim == 7 and len(LIKE_KEYS) == _lib.NKEY == 21
    th = np.array([[100., 110., 120., 130., 80., -1., 50.], [100., 110., 120., 130., -0.5, 40., -2.]])
    assert set(f.values(th)) == {"vrot"}
    assert list(f.negative(th)) == [False, True]          # (RADMAXA, RADMAXR negative: no edge, not a zero)
    off, ring = f.offsets()
    for k, o in zip(EDGE_KEYS, (4, 5, 6)):
        assert off[LIKE_KEYS.index(k)] == o and ring[LIKE_KEYS.index(k)] == 0


def test_oracle_rings_carry_the_edges():
    h = emptyheader([16, 16, 8], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600.)
    r = Rings().setRings(radii=np.array([10., 30., 50.]), xpos=8., ypos=8., vsys=0., vrot=100., vdisp=10., dens=1.,
                         inc=60., phi=30.)
    spec = [("vrot", "single"), ("radmax", "single"), ("radmaxa", "single"), ("radmaxr", "single")]
    data = np.ones((8, 16, 16), np.float32)
    orc = pipeline.Likelihood(h, GalmodPar(), r, spec, data, None, np.ones((16, 16), bool), ((0, 0), (16, 16)), 1.,
                              1., radmax_width=2.5)
    rr, g = orc.rings_of(np.array([100., 41.123456789, -1., 33.3]))
    assert rr.edge == 41.123456789 and rr.edgewidth == 2.5           # (not rounded to float32)
    assert rr.edgeapp == -1. and rr.edgerec == 33.3
    assert orc.st.side_edges(rr) == (None, 33.3)
    assert not orc.zero(rr, g)
    rr, g = orc.rings_of(np.array([100., -0.25, 10., 20.]))
    assert orc.zero(rr, g)


def test_prior_rows_with_the_annulus_inside():
    R = np.array([10., 30., 50., 70., 90., 110., 130., 150.])
    for rext in (42.5, 70., 7.5, 150., 175.):          # (between radii, equal to one, before all, at the last, beyond)
        radii = np.append(R, rext)
        lcoef, lcol, near, norm, dmed = prior_rows(radii)
        n = radii.size
        L = np.zeros((n - 2, n))
        for i in range(n - 2):
            L[i, lcol[i]] = lcoef[i]
        Lo, order, Rs, tiny, dmin = op.curvature(radii)
        assert np.allclose(L, Lo, rtol=1e-14, atol=0.), rext
        _, s, _ = op.relative_scale(radii, 1.0, 0.1)
        assert math.isclose(norm, 0.1 / s ** 2, rel_tol=1e-13), rext
        assert dmed == 20.
        gaps = np.diff(Rs)
        assert np.array_equal(near.astype(bool), np.minimum(gaps[:-1], gaps[1:]) < 0.25 * dmed), rext
