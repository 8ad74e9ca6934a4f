"""GPU (MI355X): the likelihood mask (bb_dilate_mask) against the oracle (oracle/noise.py: scipy's Euclidean
distance transform per channel, then the channel maximum filter), bit for bit: the reference's own case, random
masks with integer and fractional radii up to the largest supported, radii below one pixel, no channel dilation,
ragged cube sizes, channels without mask pixels, an empty mask, a single channel."""
import numpy as np
import pytest

from oracle.noise import dilate_mask as oracle_dilate

pytestmark = pytest.mark.gpu


def _gpu(m, r, n):
    import torch
    from bbarolo2_amd.noise import dilate_mask
    torch.cuda.set_device(0)
    return dilate_mask(m, r, n)


def test_reference_case():
    m = np.zeros((9, 21, 21), bool)
    m[4, 10, 10] = True
    assert np.array_equal(_gpu(m, 3., 2), oracle_dilate(m, 3., 2))


@pytest.mark.parametrize("shape,radius,nchan", [
    ((7, 23, 30), 0., 0), ((7, 23, 30), 0.5, 1), ((7, 23, 30), 1., 0), ((7, 23, 30), 2.5, 2), ((7, 23, 30), 3., 3),
    ((5, 70, 45), 4.2, 0), ((5, 70, 45), 7.9, 1), ((3, 100, 131), 20.3, 2), ((2, 64, 64), 56.9, 1),
    ((1, 33, 65), 5.0, 4), ((12, 96, 96), 12.0, 3)])
def test_random_masks(shape, radius, nchan):
    rng = np.random.default_rng(shape[1] + int(radius * 10) + nchan)
    m = rng.random(shape) < 0.004
    if shape[0] > 2:
        m[shape[0] // 2] = False                       # (a channel without mask pixels)
    m[0, 0, 0] = m[-1, -1, -1] = True                  # (corners)
    got, want = _gpu(m, radius, nchan), oracle_dilate(m, radius, nchan)
    print(shape, radius, nchan, int(want.sum()), int((got != want).sum()), flush=True)
    assert np.array_equal(got, want)


def test_empty_and_blob_masks():
    m = np.zeros((6, 40, 50), bool)
    assert not _gpu(m, 4.0, 2).any()
    yy, xx = np.mgrid[:40, :50]
    m[2] = (xx - 20) ** 2 / 81 + (yy - 18) ** 2 / 25 <= 1       # (an elliptical blob)
    m[3, 5:9, 40:47] = True
    for r, n in ((3.7, 1), (6.0, 3), (0.0, 2)):
        assert np.array_equal(_gpu(m, r, n), oracle_dilate(m, r, n)), (r, n)
