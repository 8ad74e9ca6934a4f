"""ORACLE -- test infrastructure only (imported by tests/; never by the product path). NumPy / SciPy restatement
of the likelihood mask of Bayesian3DFIT, dilate_mask (bbarolo/bayesian/noise.py:18-30): in every channel that
holds mask pixels, the pixels whose Euclidean distance transform to the mask (scipy.ndimage.distance_transform_edt
of the complement) is within the radius; then the maximum over +-nchan channels (maximum_filter1d, mode
'constant', cval 0). Pinned in tests/test_oracle_noise.py by the reference's own case (test_bbarolo.py:834-842:
one pixel, radius 3, 2 channels: the disk x^2 + y^2 <= 9 in channels 2-6 and nothing else) and by a brute-force
search of the mask pixels within the radius on random masks.
"""

import numpy as np


def dilate_mask(mask, radius, nchan):
    from scipy.ndimage import distance_transform_edt, maximum_filter1d
    m = np.asarray(mask, bool)
    out = m.copy()
    if radius > 0:
        for k in np.flatnonzero(m.any(axis=(1, 2))):
            out[k] = distance_transform_edt(~m[k]) <= radius
    if nchan > 0:
        out = maximum_filter1d(out.view(np.uint8), size=2 * int(nchan) + 1, axis=0, mode="constant",
                               cval=0).astype(bool)
    return out


def dilate_mask_brute(mask, radius, nchan):
    """The same set by direct search: (k, y, x) is in it when a mask pixel (k', y', x') has |k - k'| <= nchan and
    sqrt((x - x')^2 + (y - y')^2) <= radius (small masks only)."""
    m = np.asarray(mask, bool)
    nz, ny, nx = m.shape
    out = np.zeros_like(m)
    yy, xx = np.mgrid[:ny, :nx]
    for k, y, x in zip(*np.nonzero(m)):
        disk = np.sqrt((xx - x) ** 2.0 + (yy - y) ** 2.0) <= radius if radius > 0 else (xx == x) & (yy == y)
        out[max(k - nchan, 0):k + nchan + 1] |= disk
    return out
