"""The batched nested sampler (host): evidence and posterior of a Gaussian likelihood, whose values are known in
closed form, and screening by a surrogate that keeps the posterior while evaluating fewer proposals."""
import math

import numpy as np

from bbarolo2_amd.sampler import Sampler, Ellipsoid, weighted_quantiles

MU, SIG, D = 0.5, 0.05, 4


def gauss_logl(theta):
    """A normalised Gaussian density in theta: its integral over the unit cube (the evidence under the uniform prior)
    is 1 to 1e-40."""
    th = np.atleast_2d(theta)
    return -0.5 * np.sum(((th - MU) / SIG) ** 2, axis=1) - th.shape[1] * math.log(SIG * math.sqrt(2 * math.pi))


def identity(u):
    return np.asarray(u, float)


def test_ellipsoid_contains_points_and_draws_inside():
    rng = np.random.default_rng(0)
    p = 0.5 + 0.05 * rng.normal(size=(500, 3)) @ np.array([[1, 0.5, 0], [0, 1, 0.3], [0, 0, 1]])
    e = Ellipsoid(p, enlarge=1.1)
    assert np.all(e.contains(p))
    x = e.sample(2000, rng)
    assert x.shape == (2000, 3) and np.all(e.contains(x)) and np.all((x >= 0) & (x <= 1))
    # uniform: the mean of the draws is the centre
    assert np.allclose(x.mean(axis=0), e.c, atol=0.01)


def test_gaussian_evidence_and_posterior():
    s = Sampler(identity, gauss_logl, D, n_live=400, n_batch=200, seed=1).run(f_live=0.01)
    # log Z = 0 with the sampler's own error sqrt(H / n_live); H = -D/2 ln(2 pi e SIG^2) = 6.31 nats here
    assert abs(s.log_z) < 0.4, s.log_z
    assert abs(s.information - (-0.5 * D * math.log(2 * math.pi * math.e * SIG ** 2))) < 0.25, s.information
    pts, logw, _ = s.posterior()
    w = np.exp(logw)
    mean = (w[:, None] * pts).sum(axis=0)
    std = np.sqrt((w[:, None] * (pts - mean) ** 2).sum(axis=0))
    assert np.all(np.abs(mean - MU) < 0.2 * SIG), mean
    assert np.all(np.abs(std / SIG - 1) < 0.15), std
    q = np.array([weighted_quantiles(pts[:, k], w, [0.16, 0.5, 0.84]) for k in range(D)])
    assert np.all(np.abs((q[:, 2] - q[:, 0]) / (2 * SIG) - 1) < 0.15)
    assert s.n_like >= 400 and s.n_screened == 0


class NoisyExact:
    """A stand-in surrogate on the CPU: the exact log-likelihood of the unit-cube points plus noise of 0.2 nats."""

    def __init__(self, seed=0):
        self.rng = np.random.default_rng(seed)
        self.calls = 0

    def fit(self, u, y, steps=0):
        self.calls += 1
        return 0.0

    def predict(self, u):
        return gauss_logl(u) + self.rng.normal(0, 0.2, len(u))


def test_screening_keeps_the_posterior_with_fewer_evaluations():
    d = 8          # (the enlarged ellipsoid of 8 dimensions has room for proposals below the threshold)
    plain = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=2).run(f_live=0.01)
    scr = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=2, surrogate=NoisyExact(),
                  retrain_every=200).run(f_live=0.01)
    assert scr.surrogate.calls > 0 and scr.n_screened > 0
    assert scr.n_like < 0.8 * plain.n_like, (scr.n_like, plain.n_like)
    for s in (plain, scr):
        assert abs(s.log_z) < 0.6
    pp, lwp, _ = plain.posterior()
    ps, lws, _ = scr.posterior()
    for k in range(d):
        qp = weighted_quantiles(pp[:, k], np.exp(lwp), [0.16, 0.5, 0.84])
        qs = weighted_quantiles(ps[:, k], np.exp(lws), [0.16, 0.5, 0.84])
        width = qp[2] - qp[0]
        assert abs(qs[1] - qp[1]) < 0.1 * width
        assert abs((qs[2] - qs[0]) / width - 1) < 0.1


def banana_logl(theta):
    """A curved ridge (x1 ~ x0^2) in 6 dimensions: a single ellipsoid bounds it poorly."""
    th = np.atleast_2d(theta)
    x0, x1 = 4 * (th[:, 0] - 0.5), 4 * (th[:, 1] - 0.5)
    r = -0.5 * (x0 / 0.5) ** 2 - 0.5 * ((x1 - x0 ** 2 + 0.5) / 0.05) ** 2
    return r - 0.5 * np.sum(((th[:, 2:] - 0.5) / 0.05) ** 2, axis=1)


def test_random_walk_proposals_sample_the_gaussian_and_the_ridge():
    s = Sampler(identity, gauss_logl, D, n_live=400, n_batch=200, seed=4, bound="rwalk", walks=20).run(f_live=0.01)
    assert abs(s.log_z) < 0.4, s.log_z
    pts, logw, _ = s.posterior()
    w = np.exp(logw)
    mean = (w[:, None] * pts).sum(axis=0)
    std = np.sqrt((w[:, None] * (pts - mean) ** 2).sum(axis=0))
    assert np.all(np.abs(mean - MU) < 0.2 * SIG) and np.all(np.abs(std / SIG - 1) < 0.15), (mean, std)
    # the ridge: the random walks need fewer evaluations than the ellipsoid and agree on the evidence
    e = Sampler(identity, banana_logl, 6, n_live=300, n_batch=300, seed=5).run(f_live=0.05)
    r = Sampler(identity, banana_logl, 6, n_live=300, n_batch=300, seed=5, bound="rwalk", walks=20).run(f_live=0.05)
    assert abs(e.log_z - r.log_z) < 0.6, (e.log_z, r.log_z)


def test_surrogate_walks_keep_the_posterior_with_fewer_evaluations():
    d = 8
    plain = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=6, bound="rwalk", walks=20).run(f_live=0.01)
    scr = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=6, bound="rwalk", walks=20,
                  surrogate=NoisyExact(1), retrain_every=1000).run(f_live=0.01)
    assert scr.n_walk_judged > 0 and scr.n_walk_passed > 0.5 * scr.n_walk_judged
    assert scr.n_like < 0.5 * plain.n_like, (scr.n_like, plain.n_like)
    assert abs(scr.log_z - plain.log_z) < 0.5
    pp, lwp, _ = plain.posterior()
    ps, lws, _ = scr.posterior()
    for k in range(d):
        qp = weighted_quantiles(pp[:, k], np.exp(lwp), [0.16, 0.5, 0.84])
        qs = weighted_quantiles(ps[:, k], np.exp(lws), [0.16, 0.5, 0.84])
        width = qp[2] - qp[0]
        assert abs(qs[1] - qp[1]) < 0.1 * width
        assert abs((qs[2] - qs[0]) / width - 1) < 0.1


class Improving(NoisyExact):
    """A stand-in surrogate that is poor (noise 4 nats, biased down by 3) for its first three trainings, then good."""

    def predict(self, u):
        if self.calls <= 3:
            return gauss_logl(u) - 3.0 + self.rng.normal(0, 4.0, len(u))
        return super().predict(u)


def test_low_pass_retrain_keeps_walking_on_the_surrogate():
    import pytest
    d = 8
    with pytest.raises(ValueError):
        Sampler(identity, gauss_logl, d, low_pass="never")
    plain = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=7, bound="rwalk", walks=20).run(f_live=0.01)
    runs = {}
    for mode in ("exact", "retrain"):
        runs[mode] = Sampler(identity, gauss_logl, d, n_live=400, n_batch=200, seed=7, bound="rwalk", walks=20,
                             surrogate=Improving(2), retrain_every=1000, walk_min_pass=0.3,
                             low_pass=mode).run(f_live=0.01)
    ex, rt = runs["exact"], runs["retrain"]
    # a low pass rate trains the network again at once instead of a batch of exact walks
    assert rt.n_train > ex.n_train and rt.n_like < ex.n_like, (rt.n_train, ex.n_train, rt.n_like, ex.n_like)
    assert abs(rt.log_z - plain.log_z) < 0.5
    pp, lwp, _ = plain.posterior()
    ps, lws, _ = rt.posterior()
    for k in range(d):
        qp = weighted_quantiles(pp[:, k], np.exp(lwp), [0.16, 0.5, 0.84])
        qs = weighted_quantiles(ps[:, k], np.exp(lws), [0.16, 0.5, 0.84])
        width = qp[2] - qp[0]
        assert abs(qs[1] - qp[1]) < 0.1 * width
        assert abs((qs[2] - qs[0]) / width - 1) < 0.1
