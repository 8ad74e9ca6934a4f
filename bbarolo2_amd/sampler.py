"""A batched nested sampler for the vectorised likelihood, with the option to screen its proposals by a learned
surrogate of the log-likelihood.

The reference samples Bayesian3DFIT with nautilus (bbarolo/bayesian/sampling.py:100-110, 252-360; the vendored
bbarolo/bayesian/nautilus/sampler.py), an importance nested sampler whose bounds come from neural networks fitted
to the likelihood of the live points (nautilus/neural.py:60-127, bounds/neural.py). Its arguments are mirrored
where they mean the same here: n_live, n_batch (the points evaluated per likelihood call), enlarge_per_dim, f_live
(stop once the live set holds less than this fraction of the evidence), n_like_max, seed; run() / posterior() /
log_z as Sampler.run (sampler.py:388), Sampler.posterior (:693) and Sampler.log_z (:834). The likelihood is called
with (n_batch, n_dim) parameter arrays, as nautilus calls a vectorized likelihood (sampler.py:985-1063).

The algorithm is Skilling's nested sampling with a batch of replacements per likelihood call: the worst live point
dies with prior volume X_i = exp(-i / n_live); its replacement is the earliest evaluated proposal whose likelihood
exceeds the dead one's; proposals are drawn uniformly from the live points' bounding ellipsoid (enlarged by
enlarge_per_dim per dimension) within the unit cube. Proposals above the current threshold wait in a pool for the
next deaths; those below it are discarded (the thresholds only rise).

Proposals come from the live points' bounding ellipsoid (bound="ellipsoid": uniform draws, enlarged by
enlarge_per_dim per dimension, within the unit cube) or from constrained random walks started at live points
(bound="rwalk": `walks` Gaussian steps with the live points' covariance, all chains stepping together, one
likelihood call per step), which follow curved degeneracies (VROT x sin INC) that a single ellipsoid bounds poorly.

Screening (surrogate=...): the surrogate (fit(x, y), predict(x) on unit-cube points) is trained once at least
retrain_every evaluations and a factor retrain_growth more evaluations than at its last training have been made,
on every exactly evaluated proposal, its targets clipped one live-set range below the
threshold. Its margin is screen_sigma x its RMS error on the newest evaluations it was not trained on; while that
error is larger than half the live set's range of log-likelihood, or the training set is small, nothing is
screened (the plain sampler). With the ellipsoid, proposals it predicts below threshold - margin are not
evaluated. With random walks, the chains walk on the surrogate's region {prediction >= threshold - margin} (on the
device when the surrogate offers walk(): one captured graph per walk) and only their end points are evaluated; those
walks take surrogate_walks steps (default 4 x walks): a walk that has not mixed over the region leaves the end points
that pass too far inside it (the margin shell takes mass from near the boundary), which biases log Z upwards. Every
accepted point is judged by the exact likelihood.
"""

import math
import time

import numpy as np


def _logsumexp(a):
    a = np.asarray(a, float)
    m = np.max(a) if a.size else -np.inf
    if not np.isfinite(m):
        return m
    return float(m + np.log(np.sum(np.exp(a - m))))


class Ellipsoid:
    """The bounding ellipsoid of points in the unit cube: {x : (x - c)^T A (x - c) <= 1}, enlarged by `enlarge` along
    every axis, and uniform draws from its intersection with the unit cube."""

    def __init__(self, points, enlarge=1.1):
        p = np.asarray(points, float)
        self.d = p.shape[1]
        self.c = p.mean(axis=0)
        cov = np.cov(p, rowvar=False) + 1e-12 * np.eye(self.d)
        inv = np.linalg.inv(cov)
        dx = p - self.c
        k = float(np.max(np.einsum("ij,jk,ik->i", dx, inv, dx)))
        self.A = inv / (k * enlarge ** 2)
        self.L = np.linalg.cholesky(np.linalg.inv(self.A))
        # log volume: unit ball times det(L)
        self.log_vol = (self.d / 2) * math.log(math.pi) - math.lgamma(self.d / 2 + 1) + float(
            np.sum(np.log(np.diag(self.L))))

    def contains(self, x):
        dx = np.atleast_2d(x) - self.c
        return np.einsum("ij,jk,ik->i", dx, self.A, dx) <= 1.0

    def sample(self, n, rng):
        """n uniform draws from the ellipsoid within the unit cube (the cube itself when it is the smaller)."""
        if self.log_vol >= 0.0:
            out = []
            while sum(len(o) for o in out) < n:
                u = rng.random((2 * n, self.d))
                out.append(u[self.contains(u)])
            return np.concatenate(out)[:n]
        out, got = [], 0
        while got < n:
            m = max(2 * (n - got), 64)
            z = rng.normal(size=(m, self.d))
            z /= np.linalg.norm(z, axis=1, keepdims=True)
            r = rng.random(m) ** (1.0 / self.d)
            x = self.c + (z * r[:, None]) @ self.L.T
            x = x[np.all((x >= 0) & (x <= 1), axis=1)]
            out.append(x)
            got += len(x)
        return np.concatenate(out)[:n]


class Sampler:
    """Nested sampling of likelihood(prior(u)) over the unit cube u in [0, 1]^n_dim.

    prior: vectorised map of unit-cube points (m, n_dim) to parameters; likelihood: vectorised log-likelihood of
    parameters (m, n_dim) -> (m,). surrogate: None (plain) or an object with fit(u, y) and predict(u)."""

    def __init__(self, prior, likelihood, n_dim, n_live=2000, n_batch=None, enlarge_per_dim=1.1, n_update=None,
                 seed=None, surrogate=None, screen_sigma=3.0, train_steps=300, retrain_every=None, train_max=16384,
                 bound="ellipsoid", walks=25, retrain_growth=1.0, walk_min_pass=0.2, low_pass="exact",
                 surrogate_walks=None):
        self.prior, self.likelihood = prior, likelihood
        self.n_dim, self.n_live = int(n_dim), int(n_live)
        self.n_batch = int(n_batch) if n_batch else max(100, self.n_live // 4)
        self.enlarge = float(enlarge_per_dim)
        self.n_update = int(n_update) if n_update else max(self.n_live // 10, 1)
        self.rng = np.random.default_rng(seed)
        self.surrogate = surrogate
        self.screen_sigma = float(screen_sigma)
        self.train_steps = int(train_steps)
        self.retrain_every = int(retrain_every) if retrain_every else 10 * self.n_batch
        self.train_max = int(train_max)
        self.retrain_growth = float(retrain_growth)
        if bound not in ("ellipsoid", "rwalk"):
            raise ValueError(f"bound {bound!r}: 'ellipsoid' or 'rwalk'")
        self.bound, self.walks, self.scale = bound, int(walks), 1.0
        self.surrogate_walks = int(surrogate_walks) if surrogate_walks else 4 * self.walks
        self._last_keep = None
        self._walk_pass = 1.0        # fraction of the last surrogate walks' end points above the threshold
        self.walk_min_pass = float(walk_min_pass)
        if low_pass not in ("exact", "retrain"):
            raise ValueError(f"low_pass {low_pass!r}: 'exact' or 'retrain'")
        self.low_pass = low_pass
        self._retrain_now = False    # (low_pass "retrain": a low pass rate asks for a training at the next batch)
        self._low_streak = 0         # surrogate walks in a row with a low pass rate
        self.n_train = 0             # trainings of the surrogate
        self.train_log = []          # (evaluations, threshold, live range, error on fresh points, training rows)
        self.n_walk_judged = 0       # end points of surrogate walks judged by the exact likelihood
        self.n_walk_passed = 0
        self.n_like = 0            # exact likelihood evaluations
        self.n_screened = 0        # proposals the surrogate dropped
        self.n_calls = 0           # likelihood calls (batches)
        self.t_like = 0.0          # seconds in the likelihood
        self.t_surrogate = 0.0     # seconds training and predicting
        self.screen_err = None
        self._dead_u, self._dead_l, self._dead_logx = [], [], []
        self._eval_u, self._eval_l = [], []      # every exact evaluation (the surrogate's training set)
        self._trained_upto = 0

    # ------------------------------------------------------------------ likelihood
    def _evaluate(self, u):
        t0 = time.perf_counter()
        theta = self.prior(u)
        logl = np.asarray(self.likelihood(theta), float).ravel()
        self.t_like += time.perf_counter() - t0
        self.n_like += len(u)
        self.n_calls += 1
        logl = np.where(np.isnan(logl), -np.inf, logl)
        self._eval_u.append(np.asarray(u, np.float32))
        self._eval_l.append(logl)
        return logl

    # ------------------------------------------------------------------ surrogate
    def _train(self, lmin, lrange):
        """Fit the surrogate on the evaluations so far (targets clipped at lmin - lrange); sets screen_err from the
        evaluations it had not been trained on yet."""
        t0 = time.perf_counter()
        U = np.concatenate(self._eval_u)
        Y = np.concatenate(self._eval_l)
        floor = lmin - lrange
        new = slice(self._trained_upto, len(Y))
        if self._trained_upto > 0:
            un, yn = U[new], Y[new]
            keep = yn >= lmin - 0.5 * lrange
            if np.count_nonzero(keep) >= 20:
                pred = self.surrogate.predict(un[keep])
                self.screen_err = float(np.sqrt(np.mean((pred - yn[keep]) ** 2)))
        ok = np.isfinite(Y)
        U, Y = U[ok], np.maximum(Y[ok], floor)
        if len(Y) > self.train_max:
            # the most recent points, which are the ones near the threshold
            U, Y = U[-self.train_max:], Y[-self.train_max:]
        self.surrogate.fit(U, Y, steps=self.train_steps)
        self.n_train += 1
        self.train_log.append((self.n_like, float(lmin), float(lrange), self.screen_err, len(Y)))
        self._trained_upto = sum(len(a) for a in self._eval_l)
        self.t_surrogate += time.perf_counter() - t0

    def _screen(self, u, lmin, lrange):
        """The proposals to evaluate: all, or those the surrogate does not put below lmin - margin."""
        if not self._screen_ready(lrange):
            return u
        t0 = time.perf_counter()
        pred = self.surrogate.predict(u)
        self.t_surrogate += time.perf_counter() - t0
        keep = pred >= lmin - self.screen_sigma * self.screen_err
        self.n_screened += int(len(u) - np.count_nonzero(keep))
        self._last_keep = np.flatnonzero(keep)
        return u[keep]

    # ------------------------------------------------------------------ run
    def run(self, f_live=0.01, n_like_max=np.inf, verbose=False):
        """Sample until the live set holds less than f_live of the evidence (or n_like_max evaluations)."""
        t_start = time.perf_counter()
        u = self.rng.random((self.n_live, self.n_dim))
        logl = np.concatenate([self._evaluate(u[a:a + self.n_batch]) for a in range(0, self.n_live, self.n_batch)])
        live_u, live_l = u.copy(), logl.copy()
        pool_u, pool_l = np.zeros((0, self.n_dim)), np.zeros(0)
        bound = None
        it = len(self._dead_l)
        lz_dead = _logsumexp(np.asarray(self._dead_l) + self._log_dx()) if it else -np.inf
        logdx1 = math.log(-math.expm1(-1.0 / self.n_live))      # log(1 - exp(-1 / n_live))
        since_bound = self.n_update
        last_train = 0
        while True:
            i = int(np.argmin(live_l))
            lmin = live_l[i]
            # stop: the live set's share of the evidence
            logx = -it / self.n_live
            if it > 0 and np.isfinite(lmin):
                lz_live = _logsumexp(live_l) - math.log(self.n_live) + logx
                if lz_live - np.logaddexp(lz_dead, lz_live) < math.log(f_live):
                    break
            if self.n_like >= n_like_max:
                break
            # the replacement: the earliest pooled proposal above the threshold
            pool_ok = pool_l > lmin
            if not np.any(pool_ok):
                pool_u, pool_l = pool_u[pool_ok], pool_l[pool_ok]
                fin = live_l[np.isfinite(live_l)]
                lrange = float(fin.max() - lmin) if fin.size and np.isfinite(lmin) else np.inf
                due = (self.n_like - last_train >= self.retrain_every and
                       self.n_like >= self.retrain_growth * last_train)
                if (self.surrogate is not None and np.isfinite(lmin) and
                        (due or (self._retrain_now and self.n_like > last_train))):
                    self._train(lmin, lrange)
                    last_train = self.n_like
                    self._walk_pass = 1.0        # (a new network gets its chance)
                    self._retrain_now = False
                    if due:
                        self._low_streak = 0
                if since_bound >= self.n_update or bound is None:
                    bound = Ellipsoid(live_u, self.enlarge) if self.bound == "ellipsoid" else np.linalg.cholesky(
                        np.cov(live_u, rowvar=False) + 1e-12 * np.eye(self.n_dim))
                    since_bound = 0
                if self.bound == "rwalk" and np.isfinite(lmin):
                    x, lx = self._rwalk(live_u, live_l, lmin, lrange, bound)
                else:
                    if self.bound != "ellipsoid":
                        bound = Ellipsoid(live_u, self.enlarge)
                        since_bound = self.n_update
                    # a full batch of proposals that pass the screen (the surrogate is cheap beside the likelihood)
                    cand = []
                    got = 0
                    for _ in range(64):
                        x = bound.sample(self.n_batch, self.rng)
                        x = self._screen(x, lmin, lrange) if np.isfinite(lmin) else x
                        cand.append(x)
                        got += len(x)
                        if got >= self.n_batch:
                            break
                    x = np.concatenate(cand)[:self.n_batch]
                    if len(x) == 0:
                        continue
                    lx = self._evaluate(x)
                pool_u = np.concatenate([pool_u, x])
                pool_l = np.concatenate([pool_l, lx])
                continue
            j = int(np.argmax(pool_ok))
            self._dead_u.append(live_u[i].copy())
            self._dead_l.append(float(lmin))
            self._dead_logx.append(-(it + 1) / self.n_live)
            # running evidence of the dead points: L_i (X_{i-1} - X_i), X_i = exp(-i / n_live)
            lz_dead = np.logaddexp(lz_dead, lmin + logdx1 - it / self.n_live)
            live_u[i], live_l[i] = pool_u[j], pool_l[j]
            keep = np.ones(len(pool_l), bool)
            keep[j] = False
            keep &= pool_l > lmin
            pool_u, pool_l = pool_u[keep], pool_l[keep]
            it += 1
            since_bound += 1
            if verbose and it % (5 * self.n_live) == 0:
                print(f"[sampler] it {it} logL_min {lmin:.3f} n_like {self.n_like} screened {self.n_screened}",
                      flush=True)
        self._live_u, self._live_l = live_u, live_l
        self.wall = time.perf_counter() - t_start
        return self

    def _screen_ready(self, lrange):
        """The surrogate is trained on enough points and its error on fresh ones is below half the live range."""
        return (self.surrogate is not None and self.screen_err is not None and self._trained_upto >= 4 * self.n_live
                and self.screen_err < 0.5 * lrange)

    def _rwalk(self, live_u, live_l, lmin, lrange, chol):
        """n_batch new points above lmin by constrained random walks: chains start at live points above lmin and take
        `walks` Gaussian steps (covariance of the live points times scale^2, scale adapted to an acceptance of
        0.3); a step is kept if it stays in the unit cube and satisfies the constraint. Returns the chains that
        moved, with their exact log-likelihoods.

        Plain: the constraint is the exact likelihood > lmin, one likelihood call per step for all chains. Screened
        (the surrogate is ready): the chains walk on the surrogate's region {prediction >= lmin - margin} with no
        likelihood call, then their end points are judged by the exact likelihood (one call) and those at or below
        lmin are dropped; while fewer than walk_min_pass of the end points pass, the walks go back to the exact
        likelihood until the next training (low_pass "exact"), or the network is trained again at once on the
        points just judged and the walks go back to the exact likelihood only if the next pass rate is low too
        (low_pass "retrain")."""
        above = np.flatnonzero(live_l > lmin)
        if above.size == 0:
            above = np.arange(len(live_l))
        start = self.rng.choice(above, self.n_batch)
        x, lx = live_u[start].copy(), live_l[start].copy()
        moved = np.zeros(self.n_batch, bool)
        screened = self._screen_ready(lrange) and self._walk_pass > self.walk_min_pass
        margin = self.screen_sigma * self.screen_err if screened else 0.0
        t_sur = 0.0
        if screened and hasattr(self.surrogate, "walk"):
            # the surrogate walks the chains on its own device (no host round trip per step)
            t0 = time.perf_counter()
            x, moved, self.scale = self.surrogate.walk(x, chol, self.scale, lmin - margin, self.surrogate_walks,
                                                       int(self.rng.integers(1 << 62)))
            self.t_surrogate += time.perf_counter() - t0
            self.n_screened += self.n_batch * self.surrogate_walks
            return self._judge(x[moved], lmin)
        for _ in range(self.walks):
            prop = x + self.scale * (self.rng.normal(size=x.shape) @ chol.T)
            idx = np.flatnonzero(np.all((prop >= 0) & (prop <= 1), axis=1))
            acc = np.zeros(self.n_batch, bool)
            if idx.size:
                if screened:
                    t0 = time.perf_counter()
                    ok = self.surrogate.predict(prop[idx]) >= lmin - margin
                    t_sur += time.perf_counter() - t0
                    self.n_screened += int(idx.size)
                    acc[idx[ok]] = True
                    x[idx[ok]] = prop[idx[ok]]
                else:
                    lp = self._evaluate(prop[idx])
                    ok = lp > lmin
                    acc[idx[ok]] = True
                    x[idx[ok]], lx[idx[ok]] = prop[idx[ok]], lp[ok]
            moved |= acc
            self.scale *= math.exp(float(acc.mean()) - 0.3)
            self.scale = min(max(self.scale, 1e-4), 3.0)
        self.t_surrogate += t_sur
        if not screened:
            return x[moved], lx[moved]
        return self._judge(x[moved], lmin)

    def _judge(self, xm, lmin):
        """The exact likelihood judges the end points xm of surrogate walks: those above lmin are returned (with
        their log-likelihoods); the pass rate decides whether the next walks use the surrogate (low_pass)."""
        if len(xm) == 0:
            self._walk_pass = 0.0
            return xm, np.zeros(0)
        lm = self._evaluate(xm)
        good = lm > lmin
        self._walk_pass = float(good.mean())
        if self._walk_pass > self.walk_min_pass:
            self._low_streak = 0
        else:
            self._low_streak += 1
            if self.low_pass == "retrain" and self._low_streak == 1:
                # train again on the points just judged and keep walking on the surrogate; a second low pass rate
                # in a row (the new network no better) sends the walks back to the exact likelihood
                self._retrain_now = True
                self._walk_pass = 1.0
        self.n_walk_judged += len(xm)
        self.n_walk_passed += int(good.sum())
        return xm[good], lm[good]

    def _log_dx(self):
        """log(X_{i-1} - X_i) of the dead points."""
        lx = np.asarray(self._dead_logx)
        prev = np.concatenate([[0.0], lx[:-1]])
        return prev + np.log1p(-np.exp(lx - prev))

    def posterior(self, equal_weight=False):
        """(points (parameters), log_w (normalised), log_l) of the dead and the final live points; equal_weight:
        a resampled set of equal weights."""
        lx_end = self._dead_logx[-1] if self._dead_logx else 0.0
        u = np.concatenate([np.asarray(self._dead_u).reshape(-1, self.n_dim), self._live_u])
        logl = np.concatenate([np.asarray(self._dead_l), self._live_l])
        logw = np.concatenate([self._log_dx(), np.full(self.n_live, lx_end - math.log(self.n_live))]) + logl
        logw = np.where(np.isfinite(logw), logw, -np.inf)
        logw -= _logsumexp(logw)
        pts = self.prior(u)
        if equal_weight:
            w = np.exp(logw)
            n = int(np.ceil(1.0 / np.sum(w ** 2)))
            idx = self.rng.choice(len(w), n, p=w / w.sum())
            return pts[idx], np.full(n, -math.log(n)), logl[idx]
        return pts, logw, logl

    @property
    def log_z(self):
        lx_end = self._dead_logx[-1] if self._dead_logx else 0.0
        lz_dead = _logsumexp(np.asarray(self._dead_l) + self._log_dx()) if self._dead_l else -np.inf
        lz_live = _logsumexp(self._live_l) - math.log(self.n_live) + lx_end
        return float(np.logaddexp(lz_dead, lz_live))

    @property
    def information(self):
        """H = the posterior mean of ln(L / Z), in nats: the log Z of a run scatters by about sqrt(H / n_live)
        (Skilling 2006)."""
        _, logw, logl = self.posterior()
        w = np.exp(logw)
        ok = w > 0
        return float(np.sum(w[ok] * logl[ok]) - self.log_z)

    @property
    def n_eff(self):
        _, logw, _ = self.posterior()
        w = np.exp(logw)
        return float(1.0 / np.sum(w ** 2))


def weighted_quantiles(x, w, q):
    """Quantiles q of the samples x (m,) with weights w (m,)."""
    o = np.argsort(x)
    c = np.cumsum(w[o])
    c /= c[-1]
    return np.interp(q, c, x[o])
