"""Per-fit set-up of the likelihood on the GPU: the likelihood mask, the noise rms and the correlation of the noise.

dilate_mask follows the reference's (bbarolo/bayesian/noise.py:18-30, used by _setup_likemask,
bbarolo/bayesian/likelihood.py:160-175): the emission mask dilated by a disk of the given radius (pixels) in each
channel, then by +-nchan channels. It runs on the device (bb_dilate_mask: the same set of voxels, formed directly
from the mask pixels within the radius instead of a distance transform).

noise_rms, noise_correlation, set_likefactor and setup_noise follow Bayesian3DFIT._setup_noise, noise_correlation
and _set_likefactor (likelihood.py:92-128, 231-288; noise.py:33-186). The passes over the cube run on the device:
the flags of the voxels used and their counts (bb_noise_flags), the box dilation of the mask (bb_box_dilate), the
medians as exact radix selections (bb_masked_median: numpy.median's float32 result, bit for bit), the subsample of
more than 2e6 noise voxels (bb_flag_subsample, the ranks drawn by numpy's generator as the reference draws them) and
the autocorrelation sums, direct FP64 sums over the lags kept (bb_lag_xy, bb_lag_z) where the reference uses
zero-padded FFTs (the same sums, different rounding: about 1e-12 relative). What is left on the host is arithmetic
on the per-channel counts and on the (2L + 1)^2 + K + 1 sums, as in the reference.
"""

import math

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr

SUBSAMPLE = 2000000          # noise.py:94-95: the median of at most this many noise voxels


def dilate_mask_device(mask, radius, nchan, out=None):
    """The dilated mask of a device tensor mask (nz, ny, nx) (bool or uint8) as a uint8 device tensor (0 / 1)."""
    torch = _torch()
    m = mask.to(torch.uint8).contiguous() if mask.dtype != torch.uint8 or not mask.is_contiguous() else mask
    nz, ny, nx = m.shape
    if out is None:
        out = torch.empty_like(m)
    radius, nchan = float(radius), int(nchan)
    tmp = torch.empty_like(m) if radius >= 1 and nchan > 0 else None
    _lib.check(_lib.lib().bb_dilate_mask(m.data_ptr(), nz, ny, nx, radius, nchan,
                                         None if tmp is None else tmp.data_ptr(), out.data_ptr(),
                                         _stream_ptr(torch)))
    return out


def dilate_mask(mask, radius, nchan):
    """A disk of the given radius (pixels) in each channel, then +-nchan channels: the numpy bool (nz, ny, nx)
    dilation of the numpy mask, computed on the current GPU."""
    torch = _torch()
    m = np.ascontiguousarray(np.asarray(mask, bool)).view(np.uint8)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = dilate_mask_device(torch.from_numpy(m).to(dev), radius, nchan)
    return out.cpu().numpy().astype(bool)


# ---- device helpers ----------------------------------------------------------------------------------------------------
def _dev(torch):
    return torch.device("cuda", torch.cuda.current_device())


def _cube(torch, data):
    """data as a contiguous float32 (nz, ny, nx) tensor on the current GPU (numpy or torch input)."""
    if isinstance(data, torch.Tensor):
        d = data.to(device=_dev(torch), dtype=torch.float32)
    else:
        d = torch.from_numpy(np.ascontiguousarray(np.asarray(data, np.float32))).to(_dev(torch))
    if d.dim() != 3:
        raise ValueError(f"the cube must be (nz, ny, nx), not {tuple(d.shape)}")
    return d.contiguous()


def _flags(torch, mask, shape):
    """mask != 0 as a contiguous uint8 tensor of the cube's shape on the current GPU."""
    if isinstance(mask, torch.Tensor):
        m = (mask.to(_dev(torch)) != 0).to(torch.uint8)
    else:
        m = torch.from_numpy(np.ascontiguousarray(np.asarray(mask) != 0).view(np.uint8)).to(_dev(torch))
    if tuple(m.shape) != tuple(shape):
        raise ValueError(f"the mask {tuple(m.shape)} does not match the cube {tuple(shape)}")
    return m.contiguous()


def _ptr(t):
    return None if t is None else t.data_ptr()


class _Flags:
    """bb_noise_flags of a cube: used (uint8 per voxel), chan_mask, chan_used (int64 per channel, host),
    pix_emis (uint8 (ny, nx), device), pix_used (int32 (ny, nx), device)."""

    def __init__(self, torch, d, mask, emis, zero_blank):
        nz, ny, nx = d.shape
        dev = d.device
        self.used = torch.empty((nz, ny, nx), dtype=torch.uint8, device=dev)
        cm = torch.empty(nz, dtype=torch.int64, device=dev)
        cu = torch.empty(nz, dtype=torch.int64, device=dev)
        self.pix_emis = torch.empty((ny, nx), dtype=torch.uint8, device=dev)
        self.pix_used = torch.empty((ny, nx), dtype=torch.int32, device=dev)
        _lib.check(_lib.lib().bb_noise_flags(d.data_ptr(), _ptr(mask), _ptr(emis), nz, ny, nx, int(bool(zero_blank)),
                                             self.used.data_ptr(), cm.data_ptr(), cu.data_ptr(),
                                             self.pix_emis.data_ptr(), self.pix_used.data_ptr(), _stream_ptr(torch)))
        self.chan_mask = cm.cpu().numpy()
        self.chan_used = cu.cpu().numpy()


def _median(torch, d, sel, center=None):
    """(numpy.median of the float32 values d[sel] or |d[sel] - center|, as a numpy float32; their number)."""
    lib = _lib.lib()
    dev = d.device
    work = torch.empty(int(lib.bb_masked_median_work_size()), dtype=torch.uint8, device=dev)
    out = torch.empty(1, dtype=torch.float32, device=dev)
    cnt = torch.empty(1, dtype=torch.int64, device=dev)
    _lib.check(lib.bb_masked_median(d.data_ptr(), sel.data_ptr(), d.numel(), _ptr(center), work.data_ptr(),
                                    out.data_ptr(), cnt.data_ptr(), _stream_ptr(torch)))
    return out, cnt


def _std(torch, d, sel, cnt):
    lib = _lib.lib()
    work = torch.empty(int(lib.bb_masked_std_work_size()), dtype=torch.uint8, device=d.device)
    out = torch.empty(1, dtype=torch.float32, device=d.device)
    _lib.check(lib.bb_masked_std(d.data_ptr(), sel.data_ptr(), d.numel(), cnt.data_ptr(), work.data_ptr(),
                                 out.data_ptr(), _stream_ptr(torch)))
    return np.float32(out.item())


def box_dilate_device(m, rz, ry, rx):
    """_dilate (noise.py:8-15) of a uint8 device mask (nz, ny, nx): a uint8 device tensor (0 / 1)."""
    torch = _torch()
    nz, ny, nx = m.shape
    out = torch.empty_like(m)
    tmp = torch.empty_like(m) if (rz > 0) + (ry > 0) + (rx > 0) > 1 else None
    _lib.check(_lib.lib().bb_box_dilate(m.data_ptr(), nz, ny, nx, int(rz), int(ry), int(rx), _ptr(tmp),
                                        out.data_ptr(), _stream_ptr(torch)))
    return out


# ---- the noise rms -----------------------------------------------------------------------------------------------------
def noise_rms(data, mask=None):
    """The noise rms of _setup_noise (likelihood.py:105-111): 1.4826 x the median absolute deviation of the data
    outside the mask and the blanks (NaN or 0), of all the non-blank data if the mask leaves none, its standard
    deviation if that is not positive, 1 without data. The float32 product of numpy >= 2 (a Python float times a
    numpy float32). data, mask: numpy arrays or device tensors (nz, ny, nx)."""
    torch = _torch()
    d = _cube(torch, data)
    m = None if mask is None else _flags(torch, mask, d.shape)
    f = _Flags(torch, d, None, m, True)
    sel = f.used
    if f.chan_used.sum() == 0 and m is not None:
        sel = _Flags(torch, d, None, None, True).used
    med, cnt = _median(torch, d, sel)
    mad, _ = _median(torch, d, sel, center=med)
    n = int(cnt.item())
    rms = float(np.float32(1.4826) * np.float32(mad.item()))
    if not rms > 0:
        rms = float(_std(torch, d, sel, cnt)) if n else 1.
    return rms


# ---- the correlation of the noise --------------------------------------------------------------------------------------
def _half_lags(L):
    """(dy, dx) of the half-plane of lags in the order of bb_lag_xy."""
    dy = np.concatenate([np.zeros(L + 1, int), np.repeat(np.arange(1, L + 1), 2 * L + 1)])
    dx = np.concatenate([np.arange(L + 1), np.tile(np.arange(-L, L + 1), L)])
    return dy, dx


def _lag_xy(torch, d, used, chan_use, L, med):
    """The sums of bb_lag_xy spread over the full (2L + 1)^2 grid of lags (index [L + dy, L + dx])."""
    lib = _lib.lib()
    nz, ny, nx = d.shape
    if L > lib.bb_lag_xy_max():
        raise _lib.BBError(f"noise correlation: spatial lags up to {L} pixels exceed the device's window "
                           f"({lib.bb_lag_xy_max()} pixels)")
    nw = int(lib.bb_lag_xy_work_size(nz, ny, nx, L))
    nh = 2 * L * L + 2 * L + 1
    work = torch.empty(nw, dtype=torch.float64, device=d.device)
    c = torch.empty(nh, dtype=torch.float64, device=d.device)
    w = torch.empty(nh, dtype=torch.float64, device=d.device)
    cu = torch.from_numpy(np.ascontiguousarray(chan_use, np.uint8)).to(d.device)
    _lib.check(lib.bb_lag_xy(d.data_ptr(), used.data_ptr(), cu.data_ptr(), nz, ny, nx, L, float(med),
                             work.data_ptr(), nw, c.data_ptr(), w.data_ptr(), _stream_ptr(torch)))
    ch, wh = c.cpu().numpy(), w.cpu().numpy()
    dy, dx = _half_lags(L)
    cn = np.zeros((2 * L + 1, 2 * L + 1))
    cw = np.zeros((2 * L + 1, 2 * L + 1))
    for a, b in ((L + dy, L + dx), (L - dy, L - dx)):
        cn[a, b] = ch
        cw[a, b] = wh
    return cn, cw


def _lag_z(torch, d, used, pix_use, runid, K, med):
    lib = _lib.lib()
    nz, ny, nx = d.shape
    nw = int(lib.bb_lag_z_work_size(ny, nx, K))
    work = torch.empty(nw, dtype=torch.float64, device=d.device)
    c = torch.empty(K + 1, dtype=torch.float64, device=d.device)
    w = torch.empty(K + 1, dtype=torch.float64, device=d.device)
    rid = torch.from_numpy(np.ascontiguousarray(runid, np.int32)).to(d.device)
    _lib.check(lib.bb_lag_z(d.data_ptr(), used.data_ptr(), _ptr(pix_use), rid.data_ptr(), nz, ny, nx, K, float(med),
                            work.data_ptr(), nw, c.data_ptr(), w.data_ptr(), _stream_ptr(torch)))
    return c.cpu().numpy(), w.cpu().numpy()


def _free_runs(chanfree):
    """The runs [a, b) of consecutive emission-free channels, and the run index of every channel (-1 outside)."""
    runs, k, nz = [], 0, len(chanfree)
    runid = np.full(nz, -1, np.int32)
    while k < nz:
        if chanfree[k]:
            b = k
            while b + 1 < nz and chanfree[b + 1]:
                b += 1
            runid[k:b + 1] = len(runs)
            runs.append((k, b + 1))
            k = b + 1
        else:
            k += 1
    return runs, runid


def _areas_spatial(cn, cw, L, R):
    """spatial, spatial_sq, spatial_core from the lag sums (noise.py:120-134)."""
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(cw > 0.5, cn / cw, 0.)
    lags = np.arange(-L, L + 1)
    rlag = np.hypot(lags[:, None], lags[None, :])
    inside = rlag <= R
    c0 = c[L, L]
    if not c0 > 0:
        return 1., 1., 1.
    rho = c / c0
    spatial = float(np.sum(c[inside]) / c0)
    spatial_sq = float(np.sum(rho[inside] ** 2))
    rbin = np.rint(rlag).astype(int)
    with np.errstate(invalid="ignore", divide="ignore"):
        prof = np.array([rho[rbin == k].mean() if np.any(rbin == k) else np.nan
                         for k in range(int(np.ceil(R)) + 1)])
    neg = prof <= 0
    r0 = int(np.argmax(neg)) if np.any(neg) else len(prof)
    spatial_core = float(np.sum(rho[inside & (rbin < r0)]))
    return spatial, spatial_sq, spatial_core


def noise_correlation(data, mask=None, beam_fwhm=None, maxlag_xy=None, maxlag_z=6, minvox=2000, minfree=20000,
                      zero_blank=False):
    """The correlation of the noise of a cube, measured where there is no emission: noise_correlation of the
    reference (noise.py:33-186) with the passes over the cube on the current GPU. data (nz, ny, nx) (numpy or a
    device tensor; float32 on the device), mask the emission (same shape, may be None), beam_fwhm in pixels;
    zero_blank: zeros are blanks as NaN are (the likelihood's call, likelihood.py:266-268). Returns None with fewer
    than minvox noise voxels, else the reference's dict (spatial, spectral, volume, rho_spectral, nnoise, spatial_sq,
    spectral_sq, spatial_core, spectral_core, maxlag_xy, maxlag_z, spatial_from, spectral_from). maxlag_z <= 7."""
    torch = _torch()
    d = _cube(torch, data)
    nz, ny, nx = d.shape
    bp = float(beam_fwhm) if beam_fwhm is not None and beam_fwhm > 0 else 2.
    m = None if mask is None else _flags(torch, mask, d.shape)
    if m is not None and bool(m.any()):
        r = max(1, int(math.ceil(bp)))
        emis = box_dilate_device(m, 2, r, r)
        f = _Flags(torch, d, m, emis, zero_blank)
        del emis
        frac = f.chan_mask / float(ny * nx)
        chanfree = frac <= 0.001
        for k in np.flatnonzero(frac > 0.01):
            chanfree[max(k - 1, 0):k + 2] = False
        pixfree = (f.pix_emis == 0).cpu().numpy()
    else:
        f = _Flags(torch, d, None, None, zero_blank)
        chanfree = np.ones(nz, bool)
        pixfree = None                                   # (every pixel)
    nnoise = int(f.chan_used.sum())
    if nnoise < minvox:
        return None
    sel = f.used
    if nnoise > SUBSAMPLE:
        ranks = np.random.default_rng(0).choice(nnoise, SUBSAMPLE, replace=False)
        lib = _lib.lib()
        nb = int(lib.bb_flag_subsample_work_size(d.numel(), nnoise))
        work = torch.empty(nb, dtype=torch.uint8, device=d.device)
        rk = torch.from_numpy(ranks.astype(np.int64)).to(d.device)
        sel = torch.empty_like(f.used)
        _lib.check(lib.bb_flag_subsample(f.used.data_ptr(), d.numel(), nnoise, rk.data_ptr(), len(ranks),
                                         work.data_ptr(), nb, sel.data_ptr(), _stream_ptr(torch)))
        del work, rk
    medt, _ = _median(torch, d, sel)
    med = float(medt.item())
    del sel

    # spatial: the emission-free channels if they hold enough noise voxels, else every channel
    free_ch = int(f.chan_used[chanfree].sum()) >= minfree
    use = chanfree if free_ch else np.ones(nz, bool)
    R = float(maxlag_xy) if maxlag_xy else 2.5 * bp
    L = max(1, min(int(math.ceil(R)), ny - 1, nx - 1))
    cn, cw = _lag_xy(torch, d, f.used, use, L, med)
    spatial, spatial_sq, spatial_core = _areas_spatial(cn, cw, L, R)

    # spectral: the runs of emission-free channels, else the emission-free pixels, else every noise voxel
    K = max(0, min(int(maxlag_z), nz - 1))
    if K > 7:
        raise ValueError(f"noise correlation: spectral lags up to {K} channels (at most 7 on the device)")
    runs, runid = _free_runs(chanfree)
    longest = max([b - a for a, b in runs], default=0)
    if longest >= 2:
        k1 = min(max(K, 1), nz - 1)
        cz_f, wz_f = _lag_z(torch, d, f.used, None, runid, k1, med)
        pairs1 = int(round(wz_f[1]))
    else:
        pairs1 = 0
    if pixfree is None:
        npx = nnoise
        pix_use = None
    else:
        npx = int(f.pix_used.cpu().numpy()[pixfree].astype(np.int64).sum())
        pix_use = torch.from_numpy(np.ascontiguousarray(pixfree).view(np.uint8)).to(d.device)
    free_px = npx >= minfree
    if longest >= 3 and pairs1 >= minfree // 2:
        spectral_from = "free channels"
        K = min(K, longest - 1)
        cz, wz = cz_f[:K + 1], wz_f[:K + 1]
    else:
        spectral_from = "free pixels" if free_px else "voxels outside the mask"
        cz, wz = _lag_z(torch, d, f.used, pix_use if free_px else None, np.zeros(nz, np.int32), K, med)
    with np.errstate(invalid="ignore", divide="ignore"):
        rz = np.where(wz > 0.5, cz / wz, 0.)
    rz = rz / rz[0] if rz[0] > 0 else np.r_[1., np.zeros(K)]
    spectral = float(1. + 2. * np.sum(rz[1:]))
    spectral_sq = float(1. + 2. * np.sum(rz[1:] ** 2))
    k0 = int(np.argmax(rz <= 0)) if np.any(rz <= 0) else len(rz)
    spectral_core = float(1. + 2. * np.sum(rz[1:k0]))
    spatial, spectral = max(spatial, 1.), max(spectral, 1.)
    return dict(spatial=spatial, spectral=spectral, volume=spatial * spectral, rho_spectral=rz, nnoise=nnoise,
                spatial_sq=max(spatial_sq, 1.), spectral_sq=max(spectral_sq, 1.),
                spatial_core=max(spatial_core, 1.), spectral_core=max(spectral_core, 1.), maxlag_xy=R, maxlag_z=K,
                spatial_from="free channels" if free_ch else "voxels outside the mask", spectral_from=spectral_from)


# ---- the factor of the likelihood ---------------------------------------------------------------------------------------
def set_likefactor(noisecorr, beam_area, nc=None):
    """(likefactor, likefactor2d, note) of _set_likefactor (likelihood.py:231-288) for LIKEOPTS noisecorr ('measured',
    'beam' or a number > 0), the beam area in pixels and the measured correlation nc (noise_correlation's dict, or
    None when it could not be measured)."""
    A = A2 = max(float(beam_area), 1.)
    note = f"1/{A:.2f} (beam area)"
    mode = noisecorr
    if isinstance(mode, (int, float, np.integer, np.floating)) and not isinstance(mode, (bool, np.bool_)):
        if not (np.isfinite(mode) and mode > 0):
            raise ValueError(f"ERROR! noisecorr must be 'beam', 'measured' or a number > 0, not {mode!r}.")
        A = A2 = float(mode)
        note = f"1/{A:g} (given)"
    elif str(mode).lower() == "measured":
        if nc is None:
            note = f"1/{A:.2f} (beam area: noise correlation not measurable)"
        else:
            spm = max(nc["spatial"], nc.get("spatial_core", nc["spatial"]))
            sp = max(spm, float(beam_area))
            sz = max(nc["spectral"], nc.get("spectral_core", nc["spectral"]), 1.)
            A, A2 = sp * sz, sp
            note = f"1/{A:.2f} = 1/({sp:.2f} pix x {sz:.2f} chan), measured noise correlation"
            if spm < beam_area:
                note += f" (spatial: beam area, measured {spm:.2f})"
    elif str(mode).lower() != "beam":
        raise ValueError(f"ERROR! noisecorr must be 'beam', 'measured' or a number, not {mode!r}.")
    return 1. / A, 1. / A2, note


def setup_noise(header, data, mask=None, noisecorr="measured"):
    """The noise of a fit, as Bayesian3DFIT._setup_noise and _set_likefactor set it up (likelihood.py:92-128,
    231-288): a dict with noiserms, beam_area (pixels; 1 if the beam is unknown), noise_correlation (the measured
    dict, or None when not measured or not measurable), likefactor, likefactor2d and note. The zeros of the cube are
    blanks; the correlation uses the beam FWHM in pixels, |BMAJ / CDELT1|."""
    from .header import beam_area_pix
    torch = _torch()
    data = _cube(torch, data)                        # (on the device once, for both passes)
    if mask is not None:
        mask = _flags(torch, mask, data.shape)
    rms = noise_rms(data, mask)
    barea = beam_area_pix(header)
    nc = None
    if isinstance(noisecorr, str) and noisecorr.lower() == "measured":
        bmaj, cd = header.Bmaj(), header.Cdelt(0)
        bpix = abs(bmaj / cd) if bmaj and cd else None
        nc = noise_correlation(data, mask, bpix, zero_blank=True)
    lf, lf2, note = set_likefactor(noisecorr, barea, nc)
    return dict(noiserms=rms, beam_area=barea, noise_correlation=nc, likefactor=lf, likefactor2d=lf2, note=note)
