"""ORACLE -- test infrastructure only (imported by tests/ and bench.py's cpu_baseline leg; never by the product
path). The FP64 log-likelihood of one parameter vector, restated from Bayesian3DFIT._log_likelihood
(bbarolo/bayesian/likelihood.py:567-619) with the default Gaussian residuals:

  * theta -> rings: the free values, per ring or one for all the rings, stored as float32 in the ring table
    (bayesian/fitter.py:1109-1136, bayesian/rings.py:54-76); a free RADMAX, RADMAXA or RADMAXR is the edge of the disk
    or of a half, as a float (fitter.py:1112-1120);
  * zero likelihood where a free positive-only parameter or RADMAX is negative or a free FTHICK is above 1
    (likelihood.py:526-559);
  * DENS not fitted: the deconvolved density profile of the vector (bayesian/density.py:13-88 through
    oracle.profile), the densities float32 in the ring table;
  * the model box of the vector within the likelihood box (likelihood.py:429-435, DiskModel.box galmod.py:1224-1240);
  * the ANALYTIC model of the sub-rings (GalmodSetup.subrings per vector, orc_galmod_analytic; or the counter-based MC
    model, orc_galmod_mc, when the GalmodPar asks for it), the 'same' beam
    convolution with the 1e-12 cut and the flux scale (orc_conv_same, or the FFT form);
  * the masked residual in the box (orc_sumsq_residuals / orc_residuals_fracerr) plus the squared data of the
    region outside the box from a summed-area table (likelihood.py:437-440), and
    log L = -likefactor [S / (2 sigma^2) + nvox ln(2 pi sigma^2) / 2] (+ the fracerr terms).

The per-fit maps follow _setup_likemask / _set_region_sums (likelihood.py:166-229): the likelihood voxels are the
likelihood mask without the blanks (NaN or 0); dm2 = the sum of their squared data per pixel, nvox_map their
count, data_mom0 their sum, map_rms = sigma sqrt(nvox_map).

Every step calls the oracle's own restatements (oracle.c, oracle/profile.py), pinned in tests/test_oracle*.py; the
header set-up, ring conventions and sub-ring tables are the host code of bbarolo2_amd (GalmodSetup.subrings per
vector), pinned by tests/test_host.py against the reference's golden lists.
"""

import math

import numpy as np

import oracle
from oracle import profile as op
from bbarolo2_amd.galmod import GalmodSetup, model_rings, model_box, nzq_for, zstep_for, seed32, PHI_TAB
from bbarolo2_amd.smooth import beam_kernel

# ring parameters that make the likelihood zero when negative (likelihood.py:526-535)
POSITIVE = ('vrot', 'vdisp', 'inc', 'xpos', 'ypos', 'z0', 'dens', 'zcyl', 'z0thick', 'fthick', 'vdispthick')
# free parameters that are not ring values (SIGMA, FRACERR, the edges of the disk: bayesian/fitter.py:1112-1120)
GLOBAL_KEYS = ('sigma', 'fracerr', 'radmax', 'radmaxa', 'radmaxr')
# free parameters that change the deconvolved profile: with none of them free it is that of the initial rings
# (bayesian/fitter.py:606-615)
PROFILE_KEYS = ('inc', 'phi', 'xpos', 'ypos', 'radmax', 'z0', 'fthick')


def like_maps(data, likemask=None):
    """(dm2, nvox_map, mom0) of _setup_likemask (likelihood.py:176-185): per pixel the sum of the squared data of the
    likelihood voxels (likemask, not blank: NaN or 0), their number and the sum of their data, in FP64, summed over
    chunks of 16 channels in order as the reference does."""
    data = np.asarray(data, np.float32)
    nz, ny, nx = data.shape
    dm2, mom0 = np.zeros((ny, nx)), np.zeros((ny, nx))
    nvox = np.zeros((ny, nx), np.int64)
    for k in range(0, nz, 16):
        dk = data[k:k + 16]
        v = ~(~np.isfinite(dk) | (dk == 0))
        if likemask is not None:
            v &= np.asarray(likemask[k:k + 16], bool)
        d = np.where(v, dk, 0.).astype(np.float64)
        dm2 += np.sum(d * d, axis=0)
        nvox += v.sum(axis=0)
        mom0 += d.sum(axis=0)
    return dm2, nvox, mom0


class Likelihood:
    """The likelihood of a cube for the free parameters spec ([(key, 'ring' | 'single'), ...]) around rings0.

    h: header; par: GalmodPar; data (nz, ny, nx) float32; likemask (nz, ny, nx) bool or None (all); region
    (ny, nx) bool; likebox ((blo), (bhi)); sigma, likefactor; fracerr: fixed fractional error (0: none);
    smooth: densprof_smooth; likefactor2d: the profile's noise correlation (default likefactor); fft: beam
    convolutions by FFT (else direct sums); radmax_width: the width of the edge of a free RADMAX (0: sharp, arcsec)."""

    def __init__(self, h, par, rings0, spec, data, likemask, region, likebox, sigma, likefactor, fracerr=0.,
                 maskmode=2, smooth=0.3, likefactor2d=None, ringmode="CENTRED", fft=False, radmax_width=0.):
        self.h = h
        self.radmax_width = float(radmax_width)
        self.st = GalmodSetup(h, par)
        self.seed = seed32(par.ISEED)
        self.ringmode = ringmode
        kb = beam_kernel(h) if h.Bmaj() > 0 else None
        self.kern = None if kb is None else kb[0]
        self.scale = 1.0 if (kb is None or self.st.densIsFlux) else kb[1]
        self.st.zstep = zstep_for(h, self.st, self.kern is not None)
        self.rings0 = rings0.copy()
        for k, v in rings0.as_dict().items():
            setattr(self.rings0, k, np.asarray(v, np.float32).astype(np.float64))
        if rings0.nr > 1:
            self.rings0.radsep = float(np.float32(rings0.radii[-1] - rings0.radii[-2]))
        self.spec = [(k.lower(), m) for k, m in spec]
        self.data = np.asarray(data, np.float32)
        nz, ny, nx = self.data.shape
        self.likebox = (tuple(likebox[0]), tuple(likebox[1]))
        self.region = np.asarray(region, bool)
        self.sigma, self.likefactor, self.fracerr = float(sigma), float(likefactor), float(fracerr or 0.)
        self.maskmode = int(maskmode)
        self.fft = bool(fft)
        self.smooth = smooth
        self.lf2d = self.likefactor if likefactor2d is None else float(likefactor2d)
        lm = np.ones(self.data.shape, bool) if likemask is None else np.asarray(likemask, bool)
        self.likemask = lm
        dm2, self.nvox_map, self.mom0 = like_maps(self.data, lm)
        self.map_rms = self.sigma * np.sqrt(self.nvox_map.astype(np.float64))
        a = np.where(self.region, dm2, 0.)
        self.sat = np.zeros((ny + 1, nx + 1))
        self.sat[1:, 1:] = np.cumsum(np.cumsum(a, axis=0), axis=1)
        self.nvox = int(self.nvox_map[self.region].sum())
        self.use_norm = not any(k == "dens" for k, _ in self.spec)
        self.update_prof = self.use_norm and any(k in PROFILE_KEYS for k, _ in self.spec)
        self.last_dens = None
        self.last_box = None
        self.dens0 = None
        if self.use_norm and not self.update_prof:
            self.dens0 = self.density(self.rings0)
            self.rings0.dens = np.asarray(self.dens0, np.float32).astype(np.float64)

    def rings_of(self, theta):
        """(rings, globals): the float32 ring table with the free values of theta, and the non-ring values."""
        r = self.rings0.copy()
        g = {}
        i = 0
        for k, m in self.spec:
            n = r.nr if m == "ring" else 1
            v = np.asarray(theta[i:i + n], float)
            i += n
            if k in GLOBAL_KEYS:
                g[k] = float(v[0])
                continue
            v = np.asarray(v if m == "ring" else np.repeat(v, r.nr), np.float32).astype(np.float64)
            setattr(r, k, v)
        # the edges, as floats (not float32): RADMAX the edge of the disk, RADMAXA / RADMAXR those of the halves
        if "radmax" in g:
            r.edge, r.edgewidth = g["radmax"], self.radmax_width
        if "radmaxa" in g:
            r.edgeapp = g["radmaxa"]
        if "radmaxr" in g:
            r.edgerec = g["radmaxr"]
        return r, g

    def zero(self, rings, g):
        """True where the likelihood is zero (likelihood.py:537-559)."""
        for k, _ in self.spec:
            if k in POSITIVE and np.any(np.asarray(getattr(rings, k)) < 0):
                return True
            if k in ("sigma", "fracerr", "radmax") and g[k] < 0:
                return True
        if any(k == "fthick" for k, _ in self.spec) and np.any(np.asarray(rings.fthick) > 1):
            return True
        return False

    def _boxsum(self, blo, bhi):
        S = self.sat
        return S[bhi[1], bhi[0]] - S[blo[1], bhi[0]] - S[bhi[1], blo[0]] + S[blo[1], blo[0]]

    def model_box(self, rings):
        """The box of the vector's model within the likelihood box (likelihood.py:429-435)."""
        blo, bhi = model_box(self.h, rings, self.ringmode, final=True)
        Lb, Hb = self.likebox
        blo = [max(blo[0], Lb[0]), max(blo[1], Lb[1])]
        bhi = [max(min(bhi[0], Hb[0]), blo[0]), max(min(bhi[1], Hb[1]), blo[1])]
        return blo, bhi

    def density(self, rings):
        """The deconvolved profile of the rings (float64), as _deconvolved_profile."""
        Lb, Hb = self.likebox
        sl = (slice(Lb[1], Hb[1]), slice(Lb[0], Hb[0]))
        ok = self.region[sl] & (self.nvox_map[sl] > 0)
        bm = self.h.Bmaj()
        return op.deconvolved(self.st, self.kern, self.scale, rings, (Lb, Hb), self.mom0[sl], self.map_rms[sl], ok,
                              smooth=self.smooth, beam_arcsec=abs(bm) * 3600. if bm else None, corr=1. / self.lf2d,
                              mode=self.ringmode, fft=True)

    def model(self, rings, blo, bhi):
        """The convolved ANALYTIC model (by, bx, nz) of the rings in the box."""
        sr, nv = self.st.subrings(model_rings(rings, self.ringmode, self.st.radsep))
        if self.st.galmodtype == "MC":
            m = oracle.galmod_mc(self.st, sr, nv, blo, bhi, self.seed)[1]
        else:
            m = oracle.galmod_analytic(self.st, sr, blo, bhi, nzq_for(sr), PHI_TAB)
        if self.kern is not None:
            conv = oracle.conv_fft if self.fft else oracle.conv_same
            m = conv(m, self.kern, self.scale)
        return m

    def residual(self, m, blo, bhi, sigma, eps):
        """(S, lg): the residual sums of the model m (by, bx, nz) in the box, S including the data outside it."""
        sl = (slice(None), slice(blo[1], bhi[1]), slice(blo[0], bhi[0]))
        mz = np.ascontiguousarray(m.transpose(2, 0, 1))
        d, mk, rg = self.data[sl], self.likemask[sl], self.region[sl[1:]]
        out = max(float(self.sat[-1, -1]) - self._boxsum(blo, bhi), 0.)
        if eps > 0:
            ss, lg = oracle.residuals_fracerr(mz, d, mk, self.maskmode, rg, sigma * sigma, eps * eps)
            return ss + out / (sigma * sigma), lg
        return oracle.sumsq_residuals(mz, d, mk, self.maskmode, rg) + out, 0.

    def loglike(self, theta):
        """The log-likelihood of one parameter vector (likelihood.py:567-619); -inf where it is zero."""
        rings, g = self.rings_of(np.asarray(theta, float))
        if self.zero(rings, g):
            return -np.inf
        if self.update_prof:
            dens = self.density(rings)
            self.last_dens = dens
            rings.dens = np.asarray(dens, np.float32).astype(np.float64)
        blo, bhi = self.model_box(rings)
        self.last_box = (blo, bhi)
        sigma = g.get("sigma", self.sigma)
        eps = g.get("fracerr", self.fracerr)
        s2 = sigma * sigma
        if bhi[0] > blo[0] and bhi[1] > blo[1]:
            S, lg = self.residual(self.model(rings, blo, bhi), blo, bhi, sigma, eps)
        else:
            S = float(self.sat[-1, -1]) / (s2 if eps > 0 else 1.)
            lg = 0.
        if eps > 0:
            nll = 0.5 * S + 0.5 * self.nvox * math.log(2 * math.pi * s2) + 0.5 * lg
        else:
            nll = 0.5 * S / s2 + 0.5 * self.nvox * math.log(2 * math.pi * s2)
        self.last_resid = 0.5 * S / (1. if eps > 0 else s2)
        return -self.likefactor * nll


def of_fit(fit, **kw):
    """The oracle likelihood of a synthetic fit dict (bbarolo2_amd.bayes.fit_of)."""
    return Likelihood(fit["h"], fit["par"], fit["truth"], fit["free"].spec, fit["data"], fit["likemask"],
                      fit["region"], fit["box"], fit["sigma"], fit["likefactor"], **kw)
