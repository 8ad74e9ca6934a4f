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


# ---- the noise rms and the correlation of the noise -------------------------------------------------------------------
# noise_rms restates Bayesian3DFIT._setup_noise (bbarolo/bayesian/likelihood.py:105-111): 1.4826 x numpy.median of
# |x - numpy.median(x)| over the float32 data outside the mask and the blanks (NaN or 0), in float32 as numpy >= 2
# evaluates it. noise_correlation restates noise_correlation (bbarolo/bayesian/noise.py:33-186) with scipy FFTs, as
# the reference computes the autocorrelations. Pinned in tests/test_oracle_noise.py by the reference's own
# known-answer cases (test_bbarolo.py:677-704: twice the beam area for beam-smoothed white noise, one beam area when
# the correlation is the beam, 8/3 channels after Hanning smoothing, the emission excluded, None with too few voxels),
# by direct sums of the lag products on small cubes (brute force, no FFT), and the medians by a sort.

def noise_rms(data, mask=None):
    d = np.asarray(data, np.float32)
    ok = np.isfinite(d) & (d != 0)
    out = d[ok & ~np.asarray(mask, bool)] if mask is not None else d[ok]
    if out.size == 0:
        out = d[ok]
    if out.size == 0:
        return 1.
    med = np.median(out)
    rms = float(np.float32(1.4826) * np.median(np.abs(out - med)))
    if not rms > 0:
        rms = float(np.nanstd(out))
    return rms


def _box_dilate(m, rz, ry, rx):
    from scipy.ndimage import maximum_filter1d
    out = np.ascontiguousarray(m, dtype=bool).view(np.uint8)
    for axis, r in ((0, rz), (1, ry), (2, rx)):
        if r > 0:
            out = maximum_filter1d(out, size=2 * r + 1, axis=axis, mode="constant", cval=0)
    return out.astype(bool)


def noise_flags(data, mask, beam_fwhm, zero_blank=False):
    """(noise voxels, emission-free channels, emission-free pixels) of noise.py:71-89."""
    d = np.asarray(data)
    nz, ny, nx = d.shape
    valid = np.isfinite(d) & ((d != 0) if zero_blank else True)
    bp = float(beam_fwhm) if beam_fwhm is not None and beam_fwhm > 0 else 2.
    if mask is not None and np.any(mask):
        m = np.asarray(mask, bool)
        r = max(1, int(np.ceil(bp)))
        emis = _box_dilate(m, 2, r, r)
        frac = m.reshape(nz, -1).mean(axis=1)
        free = frac <= 0.001
        for k in np.flatnonzero(frac > 0.01):
            free[max(k - 1, 0):k + 2] = False
        return valid & ~emis, free, ~emis.any(axis=0)
    return valid, np.ones(nz, bool), np.ones((ny, nx), bool)


def autocorr_fft_xy(c, L):
    """sum over the channels of the linear 2-D autocorrelation of c (nz, ny, nx) at the lags |dy|, |dx| <= L, index
    [L + dy, L + dx], from zero-padded FFTs."""
    from scipy import fft
    nz, ny, nx = c.shape
    fy, fx = fft.next_fast_len(ny + L), fft.next_fast_len(nx + L)
    p = np.zeros((fy, fx // 2 + 1))
    for k in range(0, nz, 8):
        a = fft.rfft2(c[k:k + 8], s=(fy, fx))
        p += (a.real ** 2 + a.imag ** 2).sum(axis=0)
    r = fft.irfft2(p, s=(fy, fx))
    lags = np.arange(-L, L + 1)
    return r[np.ix_(lags % fy, lags % fx)]


def autocorr_fft_z(c, K):
    """sum over the pixels of the linear autocorrelation of the spectra of c (nz, ny, nx) at the lags 0..K."""
    from scipy import fft
    nz = c.shape[0]
    fz = fft.next_fast_len(nz + K)
    a = fft.rfft(c, n=fz, axis=0)
    return fft.irfft((a.real ** 2 + a.imag ** 2).sum(axis=(1, 2)), n=fz)[:K + 1]


def noise_correlation(data, mask=None, beam_fwhm=None, maxlag_xy=None, maxlag_z=6, minvox=2000, minfree=20000,
                      zero_blank=False):
    d = np.asarray(data)
    nz, ny, nx = d.shape
    bp = float(beam_fwhm) if beam_fwhm is not None and beam_fwhm > 0 else 2.
    noise, free, pixfree = noise_flags(d, mask, beam_fwhm, zero_blank)
    nnoise = int(noise.sum())
    if nnoise < minvox:
        return None
    vals = d[noise]
    if vals.size > 2000000:
        vals = vals[np.random.default_rng(0).choice(vals.size, 2000000, replace=False)]
    med = float(np.median(vals))

    free_ch = int(noise[free].sum()) >= minfree
    use = free if free_ch else np.ones(nz, bool)
    R = float(maxlag_xy) if maxlag_xy else 2.5 * bp
    L = max(1, min(int(np.ceil(R)), ny - 1, nx - 1))
    w = noise & use[:, None, None]
    cn = autocorr_fft_xy(np.where(w, d.astype(np.float64) - med, 0.), L)
    cw = autocorr_fft_xy(w.astype(np.float64), L)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(cw > 0.5, cn / cw, 0.)
    lags = np.arange(-L, L + 1)
    rl = np.sqrt(lags[:, None] ** 2.0 + lags[None, :] ** 2.0)
    inside = rl <= R
    if c[L, L] > 0:
        rho = c / c[L, L]
        spatial = float(c[inside].sum() / c[L, L])
        spatial_sq = float((rho[inside] ** 2).sum())
        rb = np.rint(rl).astype(int)
        prof = [rho[rb == k].mean() if (rb == k).any() else np.nan for k in range(int(np.ceil(R)) + 1)]
        r0 = next((k for k, v in enumerate(prof) if v <= 0), len(prof))
        spatial_core = float(rho[inside & (rb < r0)].sum())
    else:
        spatial = spatial_sq = spatial_core = 1.

    K = max(0, min(int(maxlag_z), nz - 1))
    runs, a = [], 0
    while a < nz:
        if free[a]:
            b = a + 1
            while b < nz and free[b]:
                b += 1
            runs.append((a, b))
            a = b
        else:
            a += 1
    longest = max([b - a for a, b in runs], default=0)
    pairs1 = sum(int((noise[a:b - 1] & noise[a + 1:b]).sum()) for a, b in runs if b - a >= 2)
    free_px = int((np.isfinite(d) & ((d != 0) if zero_blank else True) & pixfree[None]).sum()) >= minfree
    if longest >= 3 and pairs1 >= minfree // 2:
        spectral_from = "free channels"
        K = min(K, longest - 1)
        cz, wz = np.zeros(K + 1), np.zeros(K + 1)
        for a, b in runs:
            ok = noise[a:b]
            seg = np.where(ok, d[a:b].astype(np.float64) - med, 0.)
            for t in range(min(K, b - a - 1) + 1):
                both = ok[:b - a - t] & ok[t:]
                cz[t] += float((seg[:b - a - t] * seg[t:])[both].sum())
                wz[t] += float(both.sum())
    else:
        spectral_from = "free pixels" if free_px else "voxels outside the mask"
        w = (noise & pixfree[None]) if free_px else noise
        cz = autocorr_fft_z(np.where(w, d.astype(np.float64) - med, 0.), K)
        wz = autocorr_fft_z(w.astype(np.float64), K)
    with np.errstate(invalid="ignore", divide="ignore"):
        rz = np.where(wz > 0.5, cz / wz, 0.)
    rz = rz / rz[0] if rz[0] > 0 else np.r_[1., np.zeros(K)]
    spectral = float(1. + 2. * rz[1:].sum())
    spectral_sq = float(1. + 2. * (rz[1:] ** 2).sum())
    k0 = next((k for k, v in enumerate(rz) if v <= 0), len(rz))
    spectral_core = float(1. + 2. * rz[1:k0].sum())
    spatial, spectral = max(spatial, 1.), max(spectral, 1.)
    return dict(spatial=spatial, spectral=spectral, volume=spatial * spectral, rho_spectral=rz, nnoise=nnoise,
                spatial_sq=max(spatial_sq, 1.), spectral_sq=max(spectral_sq, 1.),
                spatial_core=max(spatial_core, 1.), spectral_core=max(spectral_core, 1.), maxlag_xy=R, maxlag_z=K,
                spatial_from="free channels" if free_ch else "voxels outside the mask", spectral_from=spectral_from)


def lag_sums_brute(data, w, med, L, K):
    """Direct sums (no FFT) of the lag products of the centred data on the voxels w: (cn, cw) over the spatial lags
    [L + dy, L + dx] (each channel's pairs) and (cz, wz) over the spectral lags 0..K (each spectrum's pairs)."""
    d = np.asarray(data, np.float64)
    c = np.where(w, d - med, 0.)
    wf = np.asarray(w, np.float64)
    nz, ny, nx = d.shape
    cn = np.zeros((2 * L + 1, 2 * L + 1))
    cw = np.zeros_like(cn)
    for dy in range(-L, L + 1):
        for dx in range(-L, L + 1):
            ya, yb = max(0, -dy), min(ny, ny - dy)
            xa, xb = max(0, -dx), min(nx, nx - dx)
            if ya >= yb or xa >= xb:
                continue
            cn[L + dy, L + dx] = (c[:, ya:yb, xa:xb] * c[:, ya + dy:yb + dy, xa + dx:xb + dx]).sum()
            cw[L + dy, L + dx] = (wf[:, ya:yb, xa:xb] * wf[:, ya + dy:yb + dy, xa + dx:xb + dx]).sum()
    cz = np.array([(c[:nz - t] * c[t:]).sum() for t in range(K + 1)])
    wz = np.array([(wf[:nz - t] * wf[t:]).sum() for t in range(K + 1)])
    return cn, cw, cz, wz
