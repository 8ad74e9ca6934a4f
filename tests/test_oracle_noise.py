"""The oracle of the likelihood mask (oracle/noise.py) against the reference's own case and a brute-force search."""
import numpy as np
import pytest

from oracle.noise import dilate_mask, dilate_mask_brute


def test_reference_case():
    # test_bbarolo.py:834-842: one pixel, radius 3, +-2 channels
    m = np.zeros((9, 21, 21), bool)
    m[4, 10, 10] = True
    dm = dilate_mask(m, 3., 2)
    yy, xx = np.mgrid[:21, :21]
    assert np.array_equal(dm[2:7], np.broadcast_to((xx - 10) ** 2 + (yy - 10) ** 2 <= 9, (5, 21, 21)))
    assert not dm[:2].any() and not dm[7:].any()


@pytest.mark.parametrize("radius,nchan", [(0., 0), (0.5, 1), (1., 0), (2.5, 2), (3., 3), (4.2, 0), (7.9, 1)])
def test_brute_force(radius, nchan):
    rng = np.random.default_rng(int(radius * 10) + nchan)
    m = rng.random((7, 23, 30)) < 0.01
    m[3] = False                                     # (a channel without mask pixels)
    m[0, 0, 0] = m[6, 22, 29] = True                 # (corners)
    assert np.array_equal(dilate_mask(m, radius, nchan), dilate_mask_brute(m, radius, nchan))
