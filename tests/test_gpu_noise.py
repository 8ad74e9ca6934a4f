"""GPU (MI355X): the per-fit set-up of the noise against the oracle (oracle/noise.py).

  * the likelihood mask (bb_dilate_mask) against scipy's Euclidean distance transform per channel and the channel
    maximum filter, bit for bit: the reference's own case, random masks with integer and fractional radii up to the
    largest supported, radii below one pixel, no channel dilation, ragged cube sizes, channels without mask pixels, an
    empty mask, a single channel;
  * the noise rms (bb_noise_flags, bb_masked_median), bit for bit: blanks (NaN and 0), ties, odd and even counts, a
    mask covering everything (all the data used), no data (1); the standard-deviation fallback of a zero MAD within
    float32 rounding (numpy sums in float32, the device in FP64);
  * the box dilation of the emission (bb_box_dilate) bit for bit;
  * the lag sums (bb_lag_xy, bb_lag_z) against direct FP64 sums (brute force) to 1e-12 relative: ragged tiles, a
    one-row cube, lags in several bands and a window above 64 KB of LDS, unused channels, runs of channels;
  * noise_correlation against the oracle's FFT restatement to 1e-9 relative in every area (the same sums, rounded
    differently), and exactly in the counts and the branches taken: the reference's known-answer cubes, the emission
    excluded, blanks, every spectral branch, the subsample of more than 2e6 noise voxels, too few voxels."""
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


# ---- the noise rms and the correlation of the noise ------------------------------------------------------------------
from oracle.noise import noise_rms as oracle_rms, noise_correlation as oracle_nc, lag_sums_brute, _box_dilate


def _torch0():
    import torch
    torch.cuda.set_device(0)
    return torch


@pytest.mark.parametrize("case", range(7))
def test_noise_rms_exact(case):
    from bbarolo2_amd.noise import noise_rms
    _torch0()
    rng = np.random.default_rng(100 + case)
    shape = [(5, 33, 47), (1, 1, 7), (16, 64, 64), (3, 10, 10), (8, 40, 40), (2, 3, 5), (6, 50, 70)][case]
    d = (rng.standard_normal(shape) * 0.01 + 0.001).astype(np.float32)
    mask = rng.random(shape) < 0.3
    if case == 1:
        d[0, 0, ::2] = 0.                                   # (blanks)
    elif case == 2:
        d[:, :10] = np.nan
        d[3] = np.round(d[3], 2)                            # (ties)
    elif case == 3:
        mask[:] = True                                      # (all the data used)
    elif case == 5:
        d[:] = 0.                                           # (no data: 1)
    got, want = noise_rms(d, mask), oracle_rms(d, mask)
    print(case, shape, got, want, flush=True)
    assert got == want


def test_noise_rms_zero_mad_falls_back_to_std():
    from bbarolo2_amd.noise import noise_rms
    _torch0()
    rng = np.random.default_rng(7)
    d = np.full((4, 30, 30), 0.5, np.float32)
    pick = rng.random(d.shape) < 0.3
    d[pick] = rng.standard_normal(int(pick.sum())).astype(np.float32)
    got, want = noise_rms(d), oracle_rms(d)
    assert want > 0 and got == pytest.approx(want, rel=2e-6)


@pytest.mark.parametrize("r", [(2, 3, 3), (0, 1, 0), (1, 0, 4), (0, 0, 0), (5, 9, 2)])
def test_box_dilate_exact(r):
    torch = _torch0()
    from bbarolo2_amd.noise import box_dilate_device
    rng = np.random.default_rng(sum(r))
    m = rng.random((9, 37, 41)) < 0.01
    m[0, 0, 0] = m[-1, -1, -1] = True
    got = box_dilate_device(torch.from_numpy(m.view(np.uint8)).cuda(), *r).cpu().numpy().astype(bool)
    assert np.array_equal(got, _box_dilate(m, *r))


@pytest.mark.parametrize("shape,L,K", [((3, 37, 45), 5, 3), ((5, 1, 30), 1, 7), ((4, 64, 64), 12, 0),
                                       ((2, 90, 100), 80, 6), ((9, 33, 33), 32, 5)])
def test_lag_sums_direct(shape, L, K):
    torch = _torch0()
    from bbarolo2_amd import noise as nz_
    rng = np.random.default_rng(L * 10 + K)
    d = rng.standard_normal(shape).astype(np.float32)
    w = rng.random(shape) < 0.85
    med = float(np.median(d[w]))
    use = np.ones(shape[0], bool)
    use[1] = False                                          # (a channel not used)
    dt = torch.from_numpy(d).cuda()
    wt = torch.from_numpy(w.view(np.uint8)).cuda()
    cn, cw = nz_._lag_xy(torch, dt, wt, use, L, med)
    bn, bw, _, _ = lag_sums_brute(d, w & use[:, None, None], med, L, 0)
    scale = np.abs(bn).max()
    assert np.array_equal(cw, bw)
    assert np.abs(cn - bn).max() <= 1e-12 * scale
    # spectral: runs [0, 2) and [3, nz) of the channels, and every pixel / the pixels of a map
    nzc = shape[0]
    runid = np.array([0, 0, -1] + [1] * (nzc - 3), np.int32)[:nzc]
    kk = min(K, nzc - 1)
    pix = rng.random(shape[1:]) < 0.7
    for pu in (None, pix):
        pt = None if pu is None else torch.from_numpy(np.ascontiguousarray(pu).view(np.uint8)).cuda()
        cz, wz = nz_._lag_z(torch, dt, wt, pt, runid, kk, med)
        ww = w if pu is None else w & pu[None]
        ez, ew = np.zeros(kk + 1), np.zeros(kk + 1)
        for a, b in ((0, 2), (3, nzc)):
            if b - a < 1:
                continue
            _, _, sz, sw = lag_sums_brute(d[a:b], ww[a:b], med, 1, min(kk, b - a - 1))
            ez[:len(sz)] += sz
            ew[:len(sw)] += sw
        assert np.array_equal(wz, ew)
        assert np.abs(cz - ez).max() <= 1e-12 * np.abs(ez).max()


def _compare_nc(got, want):
    if want is None:
        assert got is None
        return
    for k in ("nnoise", "maxlag_z", "maxlag_xy", "spatial_from", "spectral_from"):
        assert got[k] == want[k], (k, got[k], want[k])
    for k in ("spatial", "spectral", "volume", "spatial_sq", "spectral_sq", "spatial_core", "spectral_core"):
        assert got[k] == pytest.approx(want[k], rel=1e-9, abs=1e-12), (k, got[k], want[k])
    assert np.allclose(got["rho_spectral"], want["rho_spectral"], rtol=1e-9, atol=1e-12)


def _smoothed(shape, b, seed, scale=1.):
    from scipy.ndimage import gaussian_filter
    s = scale * b / (2 * np.sqrt(2 * np.log(2)))
    return gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma=(0, s, s),
                           mode="wrap").astype(np.float32)


@pytest.mark.parametrize("case", ["beam", "beam_is_corr", "hanning", "emission", "hanning_emission", "blanks",
                                  "all_masked_channels", "outside_mask", "few_voxels", "big_beam"])
def test_noise_correlation_matches_oracle(case):
    from bbarolo2_amd.noise import noise_correlation
    _torch0()
    b = 4.
    n = _smoothed((40, 96, 96), b, 5)
    h = (0.25 * np.roll(n, 1, 0) + 0.5 * n + 0.25 * np.roll(n, -1, 0)).astype(np.float32)
    mask = np.zeros(n.shape, bool)
    mask[10:30, 30:60, 30:60] = True
    kw = dict(beam_fwhm=b)
    d, m = n, None
    if case == "beam_is_corr":
        d = _smoothed((40, 96, 96), b, 6, 1 / np.sqrt(2))
    elif case == "hanning":
        d = h
    elif case == "emission":
        d, m = n.copy(), mask
        d[mask] += 10.
    elif case == "hanning_emission":
        d, m = h.copy(), mask
        d[mask] += 10.
    elif case == "blanks":
        d, m = h.copy(), mask
        d[mask] += 10.
        d[:, :, :16] = 0.
        d[5, 50:55, 70:80] = np.nan
        kw["zero_blank"] = True
    elif case == "all_masked_channels":                  # (no emission-free channel: the free pixels)
        d, m = h.copy(), np.zeros(n.shape, bool)
        m[:, 40:60, 40:60] = True
        d[m] += 5.
    elif case == "outside_mask":                         # (too few free voxels anywhere)
        d, m = h.copy(), mask
        kw["minfree"] = 10 ** 9
    elif case == "few_voxels":
        d = n[:2, :20, :20]
    elif case == "big_beam":
        kw = dict(beam_fwhm=14., maxlag_z=7)
    got, want = noise_correlation(d, m, **kw), oracle_nc(d, m, **kw)
    if want is not None:
        print(case, {k: want[k] for k in ("spatial", "spectral", "spatial_from", "spectral_from")}, flush=True)
    _compare_nc(got, want)


def test_noise_correlation_subsample():
    # more than 2e6 noise voxels: the median of the reference's random subsample (the ranks of numpy's generator)
    from bbarolo2_amd.noise import noise_correlation
    _torch0()
    d = _smoothed((36, 256, 256), 3., 9)
    d += np.float32(0.01)
    got, want = noise_correlation(d, None, beam_fwhm=3.), oracle_nc(d, None, beam_fwhm=3.)
    assert want["nnoise"] > 2000000
    _compare_nc(got, want)


def test_setup_noise_likefactor():
    from bbarolo2_amd.header import emptyheader, beam_area_pix
    from bbarolo2_amd.noise import setup_noise
    _torch0()
    h = emptyheader([96, 96, 40], [-10. / 3600., 10. / 3600., 5.], beam=40. / 3600.)
    d = _smoothed((40, 96, 96), 4., 5)
    mask = np.zeros(d.shape, bool)
    mask[10:30, 30:60, 30:60] = True
    d[mask] += 10.
    d[:, :, :8] = 0.
    s = setup_noise(h, d, mask)
    nc = oracle_nc(np.where(d == 0, np.nan, d), mask, 4.)
    _compare_nc(s["noise_correlation"], nc)
    assert s["noiserms"] == oracle_rms(d, mask)
    ba = beam_area_pix(h)
    sp = max(nc["spatial"], nc["spatial_core"], ba)
    assert s["likefactor"] == pytest.approx(1. / (sp * max(nc["spectral"], nc["spectral_core"], 1.)), rel=1e-9)
    assert s["likefactor2d"] == pytest.approx(1. / sp, rel=1e-9)
    assert setup_noise(h, d, mask, noisecorr="beam")["likefactor"] == 1. / ba


@pytest.mark.parametrize("shape", [(37, 29, 41), (16, 64, 64), (1, 7, 300), (100, 12, 9)])
def test_like_maps_exact(shape):
    # the per-pixel maps of the likelihood voxels bit for bit against the reference's chunked sums, and the data of
    # the likelihood in a box (NaN outside the voxels used and the region)
    torch = _torch0()
    from bbarolo2_amd import _lib
    from oracle.pipeline import like_maps
    rng = np.random.default_rng(shape[0])
    d = (rng.standard_normal(shape) * 3).astype(np.float32)
    d[rng.random(shape) < 0.05] = 0.
    d[rng.random(shape) < 0.02] = np.nan
    lm = rng.random(shape) < 0.7
    reg = rng.random(shape[1:]) < 0.8
    nz, ny, nx = shape
    box = (1, 2, nx - 1, ny)                              # [blo0, blo1, bhi0, bhi1)
    dt = torch.from_numpy(d).cuda()
    lt = torch.from_numpy(lm.view(np.uint8)).cuda()
    rt = torch.from_numpy(reg.view(np.uint8)).cuda()
    by, bx = box[3] - box[1], box[2] - box[0]
    dv = torch.empty((by, bx, nz), dtype=torch.float32, device="cuda")
    dm2 = torch.empty((ny, nx), dtype=torch.float64, device="cuda")
    nv = torch.empty((ny, nx), dtype=torch.int32, device="cuda")
    m0 = torch.empty((ny, nx), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().bb_like_maps(dt.data_ptr(), lt.data_ptr(), rt.data_ptr(), nz, ny, nx, _lib.bb_box(*box),
                                       dv.data_ptr(), dm2.data_ptr(), nv.data_ptr(), m0.data_ptr(), None))
    torch.cuda.synchronize()
    e2, en, e0 = like_maps(d, lm)
    assert np.array_equal(dm2.cpu().numpy(), e2) and np.array_equal(nv.cpu().numpy(), en)
    assert np.array_equal(m0.cpu().numpy(), e0)
    v = lm & np.isfinite(d) & (d != 0) & reg[None]
    want = np.where(v, d, np.nan)[:, box[1]:box[3], box[0]:box[2]].transpose(1, 2, 0)
    assert np.array_equal(dv.cpu().numpy(), want, equal_nan=True)
