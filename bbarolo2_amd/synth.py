"""Synthetic cubes of the benchmark configurations, and the likelihood inputs of a fit on them.

Headers follow emptyfits (bbarolo/utils.py:36-102); truth cubes are GALMOD models with NORM NONE (as
SimulatedGalaxyCube.run, utils.py:519-535); noise follows DiskModel.addnoise (galmod.py:1365-1377): numpy
default_rng(seed) normal noise, convolved with the beam and rescaled to the rms.

Likelihood inputs (the once-per-fit setup of LikelihoodMixin, bbarolo/bayesian/likelihood.py:83-440), with
these simplifications, stated once here:
  * the emission mask is the truth model above 2 noise rms (the reference's default is MASK SEARCH);
  * the likelihood mask dilates it by one beam FWHM in each channel and +-3 channels (noise.dilate_mask on the
    GPU, bayesian/noise.py:18-30, the default LIKEOPTS dilation (1, 3));
  * sigma is 1.4826 x the MAD of the data outside the emission mask (noise.noise_rms on the GPU,
    likelihood.py:105-111);
  * likefactor is 1 / beam area in pixels (LIKEOPTS noisecorr='beam');
  * the region is the box of the model of the initial rings (DiskModel.box, final=True).
"""

import math

import numpy as np

from .header import emptyheader, beam_area_pix
from .rings import Rings
from .galmod import GalmodPar
from .noise import dilate_mask, noise_rms as device_noise_rms


def config(n):
    """(header, truth Rings, GalmodPar, noise fraction of the peak, noise seed, free spec, description) of
    configuration n (1: CLASSIC 128x128x64, 2: GLOBAL 160x160x96, 3: 16-D Bayesian 128x128x64, 5: 13-D Bayesian
    64x64x40; "ngc2403": the geometry of the reference's likelihood benchmark)."""
    if n == 1:
        h = emptyheader([128, 128, 64], [-10 / 3600., 10 / 3600., 5.], beam=30 / 3600., obj="CFG1")
        R = np.arange(15., 346., 30.)
        r = Rings().setRings(radii=R, xpos=64., ypos=64., vsys=0., vrot=150 * (2 / np.pi) * np.arctan(R / 60.),
                             vdisp=10., z0=10., dens=1., inc=60., phi=120.)
        par = GalmodPar(LTYPE=2, GALMODTYPE="MC")
        return h, r, par, 0.05, 42, [("vrot", "ring"), ("vdisp", "ring")], "128x128x64 CLASSIC"
    if n == 2:
        h = emptyheader([160, 160, 96], [-8 / 3600., 8 / 3600., 5.2], beam=30 / 3600., obj="CFG2")
        R = np.arange(15., 1186., 30.)
        r = Rings().setRings(radii=R, xpos=80., ypos=80., vsys=0., vrot=130 * (1 - np.exp(-R / 200.)), vdisp=8.,
                             z0=10., dens=np.exp(-R / 400.), inc=60., phi=124.)
        par = GalmodPar(GALMODTYPE="ANALYTIC")
        return (h, r, par, 0.03, 7, [("vrot", "ring"), ("vdisp", "ring"), ("inc", "single"), ("phi", "single")],
                "160x160x96 GLOBAL")
    if n == 3:
        h = emptyheader([128, 128, 64], [-10 / 3600., 10 / 3600., 10.], beam=40 / 3600., obj="CFG3")
        R = np.arange(20., 301., 40.)
        r = Rings().setRings(radii=R, xpos=64., ypos=64., vsys=0., vrot=200 * (2 / np.pi) * np.arctan(R / 60.),
                             vdisp=10., z0=10., dens=1., inc=60., phi=45.)
        par = GalmodPar(GALMODTYPE="ANALYTIC")
        return h, r, par, 0.05, 3, [("vrot", "ring"), ("vdisp", "ring")], "128x128x64 BAYESIAN 16-D"
    if n == 5:
        h = emptyheader([64, 64, 40], [-10 / 3600., 10 / 3600., 20.], beam=(40 / 3600., 30 / 3600., 30.), obj="CFG5")
        R = np.arange(10., 111., 20.)
        r = Rings().setRings(radii=R, xpos=32., ypos=32., vsys=0., vrot=200 * (2 / np.pi) * np.arctan(R / 30.),
                             vdisp=40., z0=0., dens=1., inc=50., phi=60.)
        par = GalmodPar(GALMODTYPE="ANALYTIC")
        return (h, r, par, 0.05, 5, [("vrot", "ring"), ("vdisp", "ring"), ("inc", "single")],
                "64x64x40 BAYESIAN 13-D surrogate")
    if n == "ngc2403":
        # the geometry of the reference's own likelihood benchmark (performance/bench.py:46-49: the NGC 2403 cube,
        # 150x150x62, 16" pixels, 5.15 km/s channels, 59" beam; rings 60:700:120, free VROT VDISP_SINGLE INC_SINGLE
        # PHI_SINGLE), with a synthetic truth in place of the observed data
        h = emptyheader([150, 150, 62], [-16 / 3600., 16 / 3600., 5.15], beam=59 / 3600., obj="NGC2403LIKE")
        R = np.arange(60., 700., 120.)
        r = Rings().setRings(radii=R, xpos=77., ypos=77., vsys=0., vrot=100., vdisp=10., z0=10., dens=1., inc=60.,
                             phi=123.7)
        par = GalmodPar(GALMODTYPE="ANALYTIC")
        return (h, r, par, 0.05, 1, [("vrot", "ring"), ("vdisp", "single"), ("inc", "single"), ("phi", "single")],
                "150x150x62 NGC 2403-like, the reference's likelihood benchmark")
    raise ValueError(f"no configuration {n}")


MULTI_PIX = 10.      # arcsec: the pixel of every cube of the uneven multi-galaxy batch


def multi_galaxy_configs(n=64, seed=11):
    """The uneven multi-galaxy batch: n galaxies drawn with default_rng(seed), each as config(n) returns one
    (header, truth Rings, GalmodPar, noise fraction, noise seed, free spec, description). Square cubes of side
    48-256 pixels (10") and 32-128 channels; 5-40 rings (at most one per pixel of the disk radius, 0.4 of the
    side: GALMOD needs rings wider than its 0.75-pixel sub-rings); inclination 30-80 deg, position angle 0-360 deg;
    arctan rotation curve of vflat 50-300 km/s turning over at 0.15 of the disk radius; dispersion 5-30 km/s;
    exponential density of scale length half the disk radius; 10" thick layer; beam of 3-6 pixels FWHM; noise
    2-10% of the peak. The channel width puts +-(vflat sin i + 4 vdisp) in 80% of the channels. Every galaxy is
    fitted GLOBAL with free VROT VDISP per ring and one INC and PA."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.integers(48, 257))
        nz = int(rng.integers(32, 129))
        nr = int(rng.integers(5, 41))
        inc, pa = float(rng.uniform(30, 80)), float(rng.uniform(0, 360))
        vflat, vdisp = float(rng.uniform(50, 300)), float(rng.uniform(5, 30))
        beam_pix, noise = float(rng.uniform(3, 6)), float(rng.uniform(0.02, 0.10))
        rmax = 0.4 * side * MULTI_PIX
        nr = min(nr, int(rmax / MULTI_PIX))
        width = rmax / nr
        R = (np.arange(nr) + 0.5) * width
        vext = vflat * np.sin(np.radians(inc)) + 4 * vdisp
        dv = 2 * vext / (0.8 * nz)
        h = emptyheader([side, side, nz], [-MULTI_PIX / 3600., MULTI_PIX / 3600., dv],
                        beam=beam_pix * MULTI_PIX / 3600., obj=f"MULTI{i:02d}")
        r = Rings().setRings(radii=R, xpos=side / 2., ypos=side / 2., vsys=0.,
                             vrot=vflat * (2 / np.pi) * np.arctan(R / (0.15 * rmax)), vdisp=vdisp, z0=10.,
                             dens=np.exp(-R / (0.5 * rmax)), inc=inc, phi=pa)
        par = GalmodPar(GALMODTYPE="ANALYTIC")
        out.append((h, r, par, noise, 1000 + i, [("vrot", "ring"), ("vdisp", "ring"), ("inc", "single"),
                                                 ("phi", "single")],
                    f"galaxy {i}: {side}x{side}x{nz}, {nr} rings, beam {beam_pix:.1f} px"))
    return out


def theta_of(rings, free):
    """The parameter vector of the rings for the free parameters (FreePars)."""
    th = np.zeros(free.ndim)
    for k, m in free.spec:
        v = np.asarray(getattr(rings, k), float)
        th[free.index[k]] = v if m == "ring" else v[0]
    return th


def default_priors(rings, free):
    """(lo, hi) of the default uniform priors (bayesian/fitter.py:355-372): VROT 0 - 2 max VROT, VDISP 0 -
    max(2 max VDISP, 50), INC +-10 deg within 0-90, PA +-15 deg."""
    lo, hi = np.zeros(free.ndim), np.zeros(free.ndim)
    for k, m in free.spec:
        v = np.asarray(getattr(rings, k), float)
        if k == "vrot":
            a, b = 0., 2. * v.max() if v.max() > 0 else 350.
        elif k == "vdisp":
            a, b = 0., max(2. * v.max(), 50.)
        elif k == "inc":
            a, b = max(v.min() - 10., 0.), min(v.max() + 10., 90.)
        elif k == "phi":
            a, b = v.min() - 15., v.max() + 15.
        else:
            raise ValueError(k)
        lo[free.index[k]], hi[free.index[k]] = a, b
    return lo, hi


def beam_smooth(a, kern):
    """'Same' convolution of every plane of a (nz, ny, nx) with the kernel (host, setup only)."""
    from scipy.signal import fftconvolve
    return fftconvolve(a, kern[None], mode="same", axes=(1, 2))


def add_noise(model_zyx, rms, kern, seed):
    """model + Gaussian noise of the given rms, smoothed to the beam and rescaled (galmod.py:1365-1377)."""
    rng = np.random.default_rng(seed)
    noise = rng.normal(0., rms, model_zyx.shape)
    if kern is not None:
        noise = beam_smooth(noise, kern)
        noise *= rms / np.std(noise)
    return (model_zyx + noise).astype(np.float32)


def likelihood_inputs(h, truth_zyx, data, noise_rms, box):
    """(likemask, region, sigma, likefactor) for a synthetic cube (see the module docstring)."""
    mask = truth_zyx > 2. * noise_rms
    bpix = abs(h.Bmaj() / h.Cdelt(0)) if h.Bmaj() else 1.
    likemask = dilate_mask(mask, bpix, 3)          # (on the GPU)
    sigma = device_noise_rms(data, mask)           # (on the GPU)
    nz, ny, nx = data.shape
    region = np.zeros((ny, nx), bool)
    (blo, bhi) = box
    region[blo[1]:bhi[1], blo[0]:bhi[0]] = True
    return likemask, region, sigma, 1. / max(beam_area_pix(h), 1.)
