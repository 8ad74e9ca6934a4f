"""Bayesian fits of the synthetic configurations on the GPU: the vectorised likelihood (GpuLikelihood.loglike) sampled
by the batched nested sampler (sampler.Sampler), plainly or with its proposals screened by the BF16 surrogate
(surrogate.MLPSurrogate) -- the BAYESIAN task of the reference (Bayesian3DFIT.run with nautilus,
bbarolo/bayesian/fitter.py:687-946, sampling.py:252-360) restricted to the likelihood and the sampler.
"""

import time

import numpy as np

from .engine import DiskModel
from .galmod import GalmodSetup, model_rings, zstep_for, model_box
from .likelihood import FreePars, GpuLikelihood, float32_rings
from .sampler import Sampler, weighted_quantiles
from .smooth import beam_kernel
from .surrogate import MLPSurrogate
from .synth import config, default_priors, add_noise, likelihood_inputs, theta_of


def synthetic_fit(n, chunk=1024, graphs=False):
    """The likelihood of configuration n on the current GPU: its truth cube (ANALYTIC GALMOD of the truth rings on
    the GPU, convolved with the beam), noise, likelihood inputs, and the uniform priors of the free parameters.
    Returns a dict: like (GpuLikelihood), dm (DiskModel), free, lo, hi, theta_true, data, h, truth, par, box, sigma,
    likemask, region, likefactor, kern, desc."""
    return fit_of(config(n), chunk, graphs)


def fit_of(cfg, chunk=1024, graphs=False):
    """synthetic_fit for a configuration tuple as synth.config returns it (header, truth, GalmodPar, noise
    fraction, noise seed, free spec, description)."""
    h, truth, par, noise_frac, noise_seed, spec, desc = cfg
    dm = DiskModel(h, par, galmodtype=par.GALMODTYPE if par.GALMODTYPE else "ANALYTIC")
    st = dm.setup
    r32 = float32_rings(truth)
    sr, _ = st.subrings(model_rings(r32, dm.ringmode, st.radsep))
    nx, ny = h.DimAx(0), h.DimAx(1)
    mod = dm.array_batch(sr[None], (0, 0), (nx, ny), convolve=True)[0].double().cpu().numpy()
    model_zyx = mod.transpose(2, 0, 1)
    rms = noise_frac * model_zyx.max()
    kern, _ = beam_kernel(h)
    data = add_noise(model_zyx, rms, kern, noise_seed)
    box = model_box(h, r32, dm.ringmode, final=True)
    likemask, region, sigma, likefactor = likelihood_inputs(h, model_zyx, data, rms, box)
    free = FreePars(spec, truth.nr)
    like = GpuLikelihood(dm, data, truth, free, likemask=likemask, region=region, likebox=box, sigma=sigma,
                         likefactor=likefactor, chunk=chunk, graphs=graphs)
    lo, hi = default_priors(truth, free)
    return dict(like=like, free=free, lo=lo, hi=hi, theta_true=theta_of(truth, free), data=data, h=h, truth=truth,
                par=par, box=box, sigma=sigma, desc=desc, dm=dm, likemask=likemask, region=region,
                likefactor=likefactor, kern=kern)


def nested_fit(fit, n_live=1000, n_batch=1000, seed=1, surrogate=False, f_live=0.01, n_like_max=np.inf,
               verbose=False, **kw):
    """Run the nested sampler on the likelihood of a synthetic_fit dict; surrogate: screen the proposals with the
    BF16 network (trained on the device on the sampler's own evaluations). Returns the Sampler."""
    lo, hi = fit["lo"], fit["hi"]
    like = fit["like"]

    def prior(u):
        return lo + np.asarray(u, float) * (hi - lo)

    net = MLPSurrogate(len(lo), nmax=kw.pop("nmax", 1 << 15), seed=seed) if surrogate else None
    s = Sampler(prior, like.loglike, len(lo), n_live=n_live, n_batch=n_batch, seed=seed, surrogate=net, **kw)
    return s.run(f_live=f_live, n_like_max=n_like_max, verbose=verbose)


def summary(s):
    """(median, 16%, 84%) of every parameter of the weighted posterior of a Sampler: (ndim, 3)."""
    pts, logw, _ = s.posterior()
    w = np.exp(logw)
    return np.array([weighted_quantiles(pts[:, k], w, [0.5, 0.16, 0.84]) for k in range(pts.shape[1])])


def compare(plain, screened):
    """How far the screened posterior is from the plain one: the largest |median shift| in units of the plain
    68% half-width, and the largest relative change of the 68% width."""
    a, b = summary(plain), summary(screened)
    half = 0.5 * (a[:, 2] - a[:, 1])
    shift = np.abs(b[:, 0] - a[:, 0]) / np.maximum(half, 1e-300)
    width = np.abs((b[:, 2] - b[:, 1]) / np.maximum(a[:, 2] - a[:, 1], 1e-300) - 1)
    return float(shift.max()), float(width.max())
