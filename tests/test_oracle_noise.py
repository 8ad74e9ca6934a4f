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


# ---- the noise rms and the correlation of the noise (oracle/noise.py) --------------------------------------------------
from oracle.noise import noise_rms, noise_correlation, noise_flags, autocorr_fft_xy, autocorr_fft_z, lag_sums_brute


def _smoothed(shape, b, seed, scale=1.):
    from scipy.ndimage import gaussian_filter
    s = scale * b / (2 * np.sqrt(2 * np.log(2)))
    return gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma=(0, s, s), mode="wrap")


def test_reference_known_answers():
    # test_bbarolo.py:677-704: twice the beam area for white noise smoothed by the beam, one beam area when the
    # correlation is the beam itself, 8/3 channels after Hanning smoothing; the emission is excluded
    b = 4.
    barea = np.pi * b * b / (4 * np.log(2))
    n = _smoothed((40, 96, 96), b, 5)
    nc = noise_correlation(n, None, beam_fwhm=b)
    assert abs(nc["spatial"] / (2 * barea) - 1) < 0.12 and abs(nc["spectral"] - 1) < 0.05
    m = _smoothed((40, 96, 96), b, 6, 1 / np.sqrt(2))
    assert abs(noise_correlation(m, None, beam_fwhm=b)["spatial"] / barea - 1) < 0.12
    h = 0.25 * np.roll(n, 1, 0) + 0.5 * n + 0.25 * np.roll(n, -1, 0)
    assert abs(noise_correlation(h, None, beam_fwhm=b)["spectral"] / (8 / 3) - 1) < 0.06
    e = n.copy()
    mask = np.zeros(n.shape, bool)
    mask[10:30, 30:60, 30:60] = True
    e[mask] += 10.
    ne = noise_correlation(e, mask, beam_fwhm=b)
    assert abs(ne["spatial"] / (2 * barea) - 1) < 0.15 and ne["spatial_from"] == "free channels"
    eh = h.copy()
    eh[mask] += 10.
    ne = noise_correlation(eh, mask, beam_fwhm=b)
    assert abs(ne["spectral"] / (8 / 3) - 1) < 0.06 and ne["spectral_from"] == "free channels"
    assert noise_correlation(n[:2, :20, :20], None, beam_fwhm=b) is None


def test_white_noise_is_one_voxel():
    # independent voxels: the autocorrelation is 1 at lag 0 and ~0 elsewhere (the sums over 300 lags of ~N^-1/2 terms)
    w = np.random.default_rng(2).standard_normal((30, 80, 80))
    nc = noise_correlation(w, None, beam_fwhm=2.)
    assert abs(nc["spatial"] - 1) < 0.1 and abs(nc["spectral"] - 1) < 0.05 and abs(nc["spatial_sq"] - 1) < 0.01


def test_fft_sums_equal_direct_sums():
    # the zero-padded FFT autocorrelations are the direct sums of the lag products (brute force)
    rng = np.random.default_rng(3)
    d = rng.standard_normal((9, 13, 17)).astype(np.float32)
    w = rng.random(d.shape) < 0.8
    med = float(np.median(d[w]))
    c = np.where(w, d.astype(np.float64) - med, 0.)
    for L, K in ((1, 0), (4, 3), (12, 8)):
        cn, cw, cz, wz = lag_sums_brute(d, w, med, L, K)
        assert np.allclose(autocorr_fft_xy(c, L), cn, rtol=1e-10, atol=1e-9)
        assert np.allclose(autocorr_fft_xy(w.astype(float), L), cw, atol=1e-8)
        assert np.allclose(autocorr_fft_z(c, K), cz, rtol=1e-10, atol=1e-9)
        assert np.allclose(autocorr_fft_z(w.astype(float), K), wz, atol=1e-8)
        assert np.allclose(cn, cn[::-1, ::-1]) and cn[L, L] == pytest.approx((c * c).sum())


def test_noise_flags_known_case():
    # one emission voxel: the box (+-2 channels, +-ceil(beam) pixels) leaves the noise; the channel is not free
    d = np.ones((12, 20, 20), np.float32)
    d[0, 0, 0] = np.nan
    d[1, 0, 0] = 0.
    mask = np.zeros(d.shape, bool)
    mask[6, 10, 10] = True
    noise, free, pixfree = noise_flags(d, mask, 2.5, zero_blank=True)
    emis = np.zeros(d.shape, bool)
    emis[4:9, 7:14, 7:14] = True
    assert np.array_equal(noise, ~emis & np.isfinite(d) & (d != 0))
    # channel 6 has 1/400 of its pixels in the mask: above 0.1%, not free; below 1%: its neighbours stay free
    assert np.array_equal(free, np.arange(12) != 6) and np.array_equal(pixfree, ~emis.any(axis=0))
    mask[3, 0:3, 0:2] = True                            # 6/400 > 1%: channels 2-4 are not free
    noise, free, _ = noise_flags(d, mask, 2.5, zero_blank=True)
    assert np.array_equal(free, ~np.isin(np.arange(12), [2, 3, 4, 6]))


def _median_sorted(v):
    """The median by a sort: the middle value, or the float32 mean of the two middle values."""
    s = np.sort(np.asarray(v, np.float32))
    n = len(s)
    return s[n // 2] if n % 2 else np.float32((s[n // 2 - 1] + s[n // 2]) * np.float32(0.5))


@pytest.mark.parametrize("n", [1, 2, 101, 1000])
def test_noise_rms_by_sort(n):
    rng = np.random.default_rng(n)
    d = (rng.standard_normal((1, 1, n)) * 3).astype(np.float32)
    d[0, 0, ::7] = np.round(d[0, 0, ::7])               # (ties)
    d[0, 0, 0] = np.nan
    mask = np.zeros(d.shape, bool)
    mask[0, 0, 1::5] = True
    out = d[np.isfinite(d) & (d != 0) & ~mask]
    if out.size == 0:
        out = d[np.isfinite(d) & (d != 0)]
    if out.size == 0:
        assert noise_rms(d, mask) == 1.
        return
    med = _median_sorted(out)
    mad = _median_sorted(np.abs(out - med))
    expect = float(np.float32(1.4826) * mad) if mad > 0 else float(np.nanstd(out))
    assert noise_rms(d, mask) == expect
