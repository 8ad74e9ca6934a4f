"""ORACLE -- test infrastructure only (the checker of tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg). The product (bbarolo2_amd) never imports it.

ctypes wrappers of liboracle.so (oracle.c: FP64 C restatement of the reference's numba kernels) and
NumPy restatements of the host-side arithmetic that the reference does with numpy/scipy (the density
profile, bbarolo/profile.py). Pinning of each function (tests/test_oracle.py):
  orc_galmod_analytic   closed forms: total flux = sum dens x ring area, thin edge-on layer vs erf
                        profile, box = crop, single-sample hand computation on a tiny ring
  orc_ring_maps_analytic linearity: sum_r dens_r K_r = total map of orc_galmod_analytic (<= 1e-9)
  orc_conv_same         brute-force direct sum on tiny inputs, FFT (numpy) 'same' convolution,
                        flux conservation
  orc_sumsq_residuals / orc_residuals_fracerr  closed forms on hand-built cubes (maskmode 0/1/2, region)
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")


class orc_grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("radsep", C.c_double), ("cdens", C.c_double), ("pixarea", C.c_double), ("cdelt0", C.c_double),
                ("cdelt1", C.c_double), ("crota2", C.c_double),
                ("blo0", C.c_int64), ("blo1", C.c_int64), ("bhi0", C.c_int64), ("bhi1", C.c_int64),
                ("ltype", C.c_int), ("sig_instr", C.c_double), ("axtyp", C.c_int), ("velsys", C.c_int),
                ("nlines", C.c_int), ("crval3", C.c_double), ("cdelt3", C.c_double), ("crpix3", C.c_double),
                ("chwidth", C.c_double), ("zstep", C.c_double)]


class orc_mcgrid(C.Structure):
    _fields_ = [("nz", C.c_int32), ("ltype", C.c_int32), ("axtyp", C.c_int32), ("velsys", C.c_int32),
                ("nlines", C.c_int32), ("cmode", C.c_int32), ("seed", C.c_uint32), ("pad0", C.c_int32),
                ("radsep", C.c_double), ("cdens", C.c_double), ("pixarea", C.c_double), ("cdelt0", C.c_double),
                ("cdelt1", C.c_double), ("crota2", C.c_double), ("sig_instr", C.c_double), ("crval3", C.c_double),
                ("cdelt3", C.c_double), ("crpix3", C.c_double),
                ("blo0", C.c_int64), ("blo1", C.c_int64), ("bhi0", C.c_int64), ("bhi1", C.c_int64)]


_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        P = C.c_void_p
        L.orc_galmod_analytic.argtypes = [P, P, C.c_int, C.c_int, C.c_int, C.POINTER(orc_grid), P, P, P, P,
                                          C.c_int, P, C.c_int]
        L.orc_galmod_analytic.restype = None
        L.orc_ring_maps_analytic.argtypes = [P, C.c_int, P, C.c_int, P, P, C.POINTER(orc_grid), C.c_int]
        L.orc_ring_maps_analytic.restype = None
        L.orc_zgrid.argtypes = [P, P, C.c_int, C.c_int] + [C.c_double] * 11 + [C.c_int64] * 4 + [C.c_double]
        L.orc_zgrid.restype = C.c_int
        L.orc_conv_same.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_double, C.c_int]
        L.orc_conv_same.restype = None
        L.orc_sumsq_residuals.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_int]
        L.orc_sumsq_residuals.restype = C.c_double
        L.orc_residuals_fracerr.argtypes = [P, P, P, C.c_int, P, C.c_int, C.c_int, C.c_int, C.c_double,
                                            C.c_double, C.POINTER(C.c_double)]
        L.orc_residuals_fracerr.restype = C.c_double
        U4 = C.c_uint32 * 4
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_philox4x32_10.restype = None
        L.orc_galmod_mc.argtypes = [P, P, P, C.c_int, P, C.POINTER(orc_mcgrid), P, P, P]
        L.orc_galmod_mc.restype = C.c_double
        for f in ("orc_mc_log", "orc_mc_sqrt", "orc_mc_rcp"):
            getattr(L, f).argtypes = [C.c_float]
            getattr(L, f).restype = C.c_float
        L.orc_mc_sincos_turn.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_mc_layer_dev.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
        L.orc_mc_layer_dev.restype = C.c_float
        L.orc_mc_box_muller.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_norm.argtypes = [P, C.c_int, C.c_int, C.c_int, P, P, C.c_int, C.c_int, C.c_int64, C.c_int64,
                               C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                               C.c_int, C.c_int, P, C.c_double, C.c_double, C.c_int, P, P, P, C.c_int, P, C.c_int,
                               C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_norm.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def make_grid(setup, blo, bhi):
    """The orc_grid of a GalmodSetup-like object (bbarolo2_amd.galmod.GalmodSetup) and a box."""
    g = orc_grid()
    g.nx, g.ny, g.nz = bhi[0] - blo[0], bhi[1] - blo[1], setup.nsubs
    g.radsep, g.cdens, g.pixarea = setup.radsep, setup.cdens, setup.pixarea
    g.cdelt0, g.cdelt1, g.crota2 = setup.cdelt[0], setup.cdelt[1], setup.crota2
    g.blo0, g.blo1, g.bhi0, g.bhi1 = blo[0], blo[1], bhi[0], bhi[1]
    g.ltype, g.sig_instr, g.axtyp, g.velsys = setup.ltype, setup.sig_instr, setup.axtyp, setup.velsys
    g.nlines = len(setup.lfreq0)
    g.crval3, g.cdelt3, g.crpix3 = setup.crval3, setup.cdelt3, setup.crpix3
    g.chwidth, g.zstep = setup.chmin, setup.zstep
    return g


def galmod_analytic(setup, sr, blo, bhi, nzq, ptab, out=None, ring_lo=0, ring_hi=None):
    """orc_galmod_analytic: the ANALYTIC model (by, bx, nz) float64 of the sub-rings sr (16, nsub)."""
    sr = np.ascontiguousarray(sr, np.float64)
    nsub = sr.shape[1]
    g = make_grid(setup, blo, bhi)
    if out is None:
        out = np.zeros((g.ny, g.nx, g.nz))
    ve = np.ascontiguousarray(setup.vedges_lines, np.float64)
    ri = np.ascontiguousarray(setup.relint, np.float64)
    cd = np.ascontiguousarray(setup.cd2i, np.float64)
    lf = np.ascontiguousarray(setup.lfreq0, np.float64)
    pt = np.ascontiguousarray(ptab, np.float64)
    lib().orc_galmod_analytic(_ptr(out), _ptr(sr), nsub, ring_lo, nsub if ring_hi is None else ring_hi,
                              C.byref(g), _ptr(ve), _ptr(ri), _ptr(cd), _ptr(lf), int(nzq), _ptr(pt), len(pt))
    return out


def ring_maps_analytic(setup, sr, widx, wval, nr, blo, bhi, nzq):
    """orc_ring_maps_analytic: unit-density total-intensity maps (nr, by, bx)."""
    sr = np.ascontiguousarray(sr, np.float64)
    widx = np.ascontiguousarray(widx, np.int64)
    wval = np.ascontiguousarray(wval, np.float64)
    g = make_grid(setup, blo, bhi)
    out = np.zeros((nr, g.ny, g.nx))
    lib().orc_ring_maps_analytic(_ptr(out), nr, _ptr(sr), sr.shape[1], _ptr(widx), _ptr(wval), C.byref(g), int(nzq))
    return out


def zgrid(nmax, ltype, z0, sinc, pixsize, dvdz, sig, R0, radsep, xpos, ypos, cdelt0, cdelt1, blo, bhi, zstep):
    zq = np.zeros(max(nmax, 1))
    wq = np.zeros(max(nmax, 1))
    n = lib().orc_zgrid(_ptr(zq), _ptr(wq), nmax, ltype, z0, sinc, pixsize, dvdz, sig, R0, radsep, xpos, ypos,
                        cdelt0, cdelt1, blo[0], blo[1], bhi[0], bhi[1], zstep)
    return zq[:n], wq[:n]


def conv_same(model_yxz, kern, scale=1.0, thresh=True):
    """orc_conv_same: beam-convolved model (ny, nx, nz) float64, thresholded at 1e-12 (thresh) and scaled."""
    m = np.ascontiguousarray(model_yxz, np.float64)
    k = np.ascontiguousarray(kern, np.float64)
    out = np.empty_like(m)
    ny, nx, nz = m.shape
    lib().orc_conv_same(_ptr(m), _ptr(out), ny, nx, nz, _ptr(k), k.shape[0], k.shape[1], float(scale), int(bool(thresh)))
    return out


def sumsq_residuals(model_zyx, data_zyx, mask=None, maskmode=2, region=None):
    m = np.ascontiguousarray(model_zyx, np.float64)
    d = np.ascontiguousarray(data_zyx, np.float32)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    rg = None if region is None or region.size == 0 else np.ascontiguousarray(region, np.uint8)
    if mk is None:
        maskmode = 0
    nz, ny, nx = m.shape
    return lib().orc_sumsq_residuals(_ptr(m), _ptr(d), _ptr(mk), int(maskmode), _ptr(rg), nz, ny, nx)


def residuals_fracerr(model_zyx, data_zyx, mask, maskmode, region, sigma2, eps2):
    m = np.ascontiguousarray(model_zyx, np.float64)
    d = np.ascontiguousarray(data_zyx, np.float32)
    mk = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    rg = None if region is None or region.size == 0 else np.ascontiguousarray(region, np.uint8)
    if mk is None:
        maskmode = 0
    nz, ny, nx = m.shape
    lg = C.c_double(0.)
    s = lib().orc_residuals_fracerr(_ptr(m), _ptr(d), _ptr(mk), int(maskmode), _ptr(rg), nz, ny, nx,
                                    float(sigma2), float(eps2), C.byref(lg))
    return s, lg.value


def conv_fft(model_yxz, kern, scale=1.0, thresh=True):
    """The FFT 'same' convolution of the reference (FFTConvolver.__call__, bbarolo/smooth.py:107-165: rfft2 of
    the non-empty planes padded to next_fast_len(n + k//2), times the kernel transform, irfft2, crop at k//2),
    then the 1e-12 threshold and the scale (DiskModel.convolve, galmod.py:1273-1283). (ny, nx, nz) -> (ny, nx, nz).
    Used as the CPU baseline's convolution (the reference's own method); equal to conv_same to ~1e-15."""
    import scipy.fft as sfft
    d = np.ascontiguousarray(np.asarray(model_yxz, np.float64).transpose(2, 0, 1))
    nz, ny, nx = d.shape
    full = d.reshape(nz, -1).any(axis=1)
    out = np.zeros_like(d)
    if full.any():
        k = np.asarray(kern, np.float64)
        ky, kx = k.shape
        Ny = sfft.next_fast_len(ny + ky // 2, real=True)
        Nx = sfft.next_fast_len(nx + kx // 2, real=True)
        f = sfft.rfft2(d[full], s=(Ny, Nx))
        f *= sfft.rfft2(k, s=(Ny, Nx))
        out[full] = sfft.irfft2(f, s=(Ny, Nx))[..., ky // 2:ky // 2 + ny, kx // 2:kx // 2 + nx]
    if thresh:
        out[out < 1.E-12] = 0.
    if scale != 1:
        out *= scale
    return np.ascontiguousarray(out.transpose(1, 2, 0))


# ---- counter-based Monte-Carlo GALMOD (mc.c) -----------------------------------------------------------------
def philox(ctr, key):
    """Philox4x32-10 of a counter (4 uint32) and key (2 uint32) -> 4 uint32."""
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def make_mcgrid(setup, blo, bhi, seed):
    g = orc_mcgrid()
    g.nz, g.ltype, g.axtyp, g.velsys = setup.nsubs, setup.ltype, setup.axtyp, setup.velsys
    g.nlines, g.cmode, g.seed = len(setup.lfreq0), int(setup.cmode), int(seed) & 0xFFFFFFFF
    g.radsep, g.cdens, g.pixarea = setup.radsep, setup.cdens, setup.pixarea
    g.cdelt0, g.cdelt1, g.crota2 = setup.cdelt[0], setup.cdelt[1], setup.crota2
    g.sig_instr, g.crval3, g.cdelt3, g.crpix3 = setup.sig_instr, setup.crval3, setup.cdelt3, setup.crpix3
    g.blo0, g.blo1, g.bhi0, g.bhi1 = blo[0], blo[1], bhi[0], bhi[1]
    return g


def galmod_mc(setup, sr, nv, blo, bhi, seed):
    """orc_galmod_mc: (counts (by, bx, nz) uint64, model (by, bx, nz) float64, q) of the counter-based MC model."""
    sr = np.ascontiguousarray(sr, np.float64)
    nv = np.ascontiguousarray(nv, np.int64)
    g = make_mcgrid(setup, blo, bhi, seed)
    counts = np.zeros((bhi[1] - blo[1], bhi[0] - blo[0], setup.nsubs), np.uint64)
    out = np.zeros(counts.shape)
    ri = np.ascontiguousarray(setup.relint, np.float64)
    cd = np.ascontiguousarray(setup.cd2i, np.float64)
    lf = np.ascontiguousarray(setup.lfreq0, np.float64)
    q = lib().orc_galmod_mc(_ptr(counts), _ptr(out), _ptr(sr), sr.shape[1], _ptr(nv), C.byref(g), _ptr(ri),
                            _ptr(cd), _ptr(lf))
    return counts, out, q


# ---- the 3DFIT ring residual (norm.c) ------------------------------------------------------------------------
def norm3dfit(model_yxz, data_yxz, mask_yxz, blo, x0, y0, phi, inc, r1, r2, wpow, side, ftype, chan_noise,
              data_noise, bweight, normtype, pre=None, ilo=None, ihi=None, thr=None, foot=0, hmax=0.):
    """orc_norm: (value, model scaled in place (a copy), nblanks, ntot) of _norm (galfit.py:39-188)."""
    m = np.array(model_yxz, np.float64, order="C", copy=True)
    by, bx, nz = m.shape
    d = np.ascontiguousarray(data_yxz, np.float32)
    mk = np.ascontiguousarray(mask_yxz, np.uint8)
    ny, nx = d.shape[0], d.shape[1]
    cn = np.ascontiguousarray(chan_noise, np.float64)
    if pre is not None:
        pre = np.ascontiguousarray(pre, np.float64)
        ilo = np.ascontiguousarray(ilo, np.int64)
        ihi = np.ascontiguousarray(ihi, np.int64)
        thr = np.ascontiguousarray(thr, np.float64)
        ky = len(ilo)
    else:
        ky = 1
    nb, nt = C.c_int64(), C.c_int64()
    v = lib().orc_norm(_ptr(m), by, bx, nz, _ptr(d), _ptr(mk), ny, nx, int(blo[0]), int(blo[1]), float(x0), float(y0),
                       float(phi), float(inc), float(r1), float(r2), float(wpow), int(side), int(ftype), _ptr(cn),
                       float(data_noise), float(bweight), int(normtype), _ptr(pre), _ptr(ilo), _ptr(ihi), ky,
                       _ptr(thr), int(foot), float(hmax), C.byref(nb), C.byref(nt))
    return v, m, nb.value, nt.value
