"""The likelihood surrogate on the GPU: a small network din -> 128 (SiLU) x 4 -> 1 with BF16 weights, trained on the
device (bb_mlp_train: hand-written matrix-core forward and backward passes, full-batch Adam) on the
(parameters, log-likelihood) pairs that a sampler has evaluated, and used to screen its proposals.

The reference's learned component is the regression network of nautilus's neural bounds
(bbarolo/bayesian/nautilus/mlp.py:17-222, trained by neural.py:60-127 on the live points and used by
bounds/neural.py to decide where to draw). This one models the log-likelihood of the unit-cube coordinates; the
inputs are the unit-cube points, the targets are standardised on the host.
"""

import ctypes as C

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr

DIN_MAX = 32


def glorot_params(din, hidden, nhidden, seed=0):
    """Initial parameters (float32, bb_mlp layout): Glorot-uniform weights with the padded input columns zero,
    zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for l in range(nhidden):
        K = DIN_MAX if l == 0 else hidden
        fan_in = din if l == 0 else hidden
        a = np.sqrt(6.0 / (fan_in + hidden))
        W = rng.uniform(-a, a, (hidden, K)).astype(np.float32)
        if l == 0:
            W[:, din:] = 0
        parts += [W.ravel(), np.zeros(hidden, np.float32)]
    a = np.sqrt(6.0 / (hidden + 1))
    parts += [rng.uniform(-a, a, hidden).astype(np.float32), np.zeros(1, np.float32)]
    return np.concatenate(parts)


class MLPSurrogate:
    """A bb_mlp network on the current device. fit(x, y) trains it on unit-cube points x (n, din) and values y (n,)
    (standardised internally); predict(x) returns values in the units of y."""

    def __init__(self, din, hidden=128, nhidden=4, nmax=1 << 15, seed=0):
        torch = _torch()
        self.torch = torch
        self.din, self.hidden, self.nhidden, self.nmax = int(din), int(hidden), int(nhidden), int(nmax)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.bb_mlp_create(self.din, self.hidden, self.nhidden, self.nmax, C.byref(h)))
        self.handle = h
        self.nparams = int(L.bb_mlp_param_count(h))
        self.seed = seed
        self.reset()

    def __del__(self):
        try:
            _lib.lib().bb_mlp_destroy(self.handle)
        except Exception:
            pass

    def reset(self):
        """Fresh Glorot parameters (and Adam state)."""
        self.set_params(glorot_params(self.din, self.hidden, self.nhidden, self.seed))
        self.mu, self.sd = 0.0, 1.0
        self.xc, self.xs = np.zeros(self.din), np.ones(self.din)

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.nparams:
            raise ValueError(f"{p.size} parameters, the network has {self.nparams}")
        _lib.check(_lib.lib().bb_mlp_set_params(self.handle, p.ctypes.data))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        _lib.check(_lib.lib().bb_mlp_get_params(self.handle, p.ctypes.data))
        return p

    def _dev(self, a):
        t = self.torch
        return t.as_tensor(np.ascontiguousarray(a, np.float32), device=self.dev)

    def _rows(self, x):
        x = np.atleast_2d(np.asarray(x, np.float32))
        if x.shape[1] != self.din:
            raise ValueError(f"inputs of width {x.shape[1]}, the network takes {self.din}")
        if x.shape[0] > self.nmax:
            raise ValueError(f"{x.shape[0]} rows, at most {self.nmax}")
        return x

    def grad(self, x, t):
        """(gradient (nparams,), loss) of 0.5 mean (net(x) - t)^2 (raw targets, no standardisation)."""
        x = self._rows(x)
        dx, dt = self._dev(x), self._dev(np.asarray(t, np.float32).ravel())
        g = self.torch.empty(self.nparams, dtype=self.torch.float32, device=self.dev)
        ls = self.torch.empty(1, dtype=self.torch.float32, device=self.dev)
        _lib.check(_lib.lib().bb_mlp_grad(self.handle, dx.data_ptr(), dt.data_ptr(), x.shape[0], g.data_ptr(),
                                          ls.data_ptr(), _stream_ptr(self.torch)))
        return g.cpu().numpy(), float(ls.cpu().numpy()[0])

    def raw_forward(self, x):
        """net(x) without the target scaling."""
        x = self._rows(x)
        dx = self._dev(x)
        y = self.torch.empty(x.shape[0], dtype=self.torch.float32, device=self.dev)
        _lib.check(_lib.lib().bb_mlp_forward(self.handle, dx.data_ptr(), x.shape[0], y.data_ptr(),
                                             _stream_ptr(self.torch)))
        return y.cpu().numpy().astype(np.float64)

    def train_raw(self, x, t, steps, lr):
        """steps Adam steps on (x, t) as given; returns the loss before each step (numpy)."""
        x = self._rows(x)
        dx, dt = self._dev(x), self._dev(np.asarray(t, np.float32).ravel())
        ls = self.torch.empty(max(steps, 1), dtype=self.torch.float32, device=self.dev)
        _lib.check(_lib.lib().bb_mlp_train(self.handle, dx.data_ptr(), dt.data_ptr(), x.shape[0], int(steps),
                                           float(lr), ls.data_ptr(), _stream_ptr(self.torch)))
        return ls[:steps].cpu().numpy()

    def fit(self, x, y, steps=400, lr=3e-3, restart=False):
        """Train on unit-cube points x and values y; both are standardised (the inputs per axis to the mean and
        spread of x, the targets to mean 0 and deviation 1), so that the network sees the region the points fill
        (a posterior can fill 1e-3 of the cube along an axis) at unit scale. Returns the last loss."""
        x = np.asarray(x, np.float64)
        y = np.asarray(y, np.float64).ravel()
        if restart:
            self.reset()
        self.mu, self.sd = float(y.mean()), float(y.std()) or 1.0
        self.xc = x.mean(axis=0)
        self.xs = np.maximum(x.std(axis=0), 1e-12)
        ls = self.train_raw(self._norm(x), (y - self.mu) / self.sd, steps, lr)
        return float(ls[-1]) if len(ls) else float("nan")

    def _norm(self, x):
        return ((np.asarray(x, np.float64) - self.xc) / self.xs).astype(np.float32)

    def walk(self, x0, chol, scale, thr, steps, seed):
        """Random walks on the region {prediction >= thr} on the device (bb_mlp_walk, one captured graph): from the
        unit-cube points x0 (n, din), `steps` Gaussian steps x + scale z chol^T (z standard normal, torch's generator
        seeded with seed, drawn before the walk); a step is kept if it stays in the unit cube and its prediction is
        >= thr; after each step the scale is multiplied by exp(acceptance - 0.3) (clamped to [1e-4, 3]). Returns
        (end points (n, din) float64, moved (n,) bool, final scale); one synchronisation at the end."""
        t = self.torch
        x0 = self._rows(x0)
        n, d = x0.shape
        steps = int(steps)
        if n == 0:
            return np.zeros((0, d)), np.zeros(0, bool), float(scale)
        key = (n, steps)
        if getattr(self, "_wbuf_key", None) != key:
            f32 = dict(dtype=t.float32, device=self.dev)
            self._wbuf = dict(x=t.empty((n, d), **f32), zl=t.empty((steps * n, d), **f32), sc=t.empty(steps + 1, **f32),
                              thr=t.empty(1, **f32), xc=t.empty(d, **f32), xs=t.empty(d, **f32),
                              moved=t.empty(n, dtype=t.uint8, device=self.dev))
            self._wbuf_key = key
        b = self._wbuf
        g = t.Generator(device=self.dev)
        g.manual_seed(int(seed))
        L = self._dev(np.asarray(chol, np.float64).T)
        t.matmul(t.randn((steps * n, d), generator=g, device=self.dev), L, out=b["zl"])
        b["x"].copy_(t.from_numpy(x0))
        b["xc"].copy_(t.from_numpy(np.asarray(self.xc, np.float32)))
        b["xs"].copy_(t.from_numpy(np.asarray(self.xs, np.float32)))
        b["sc"][:1].fill_(float(scale))
        b["thr"].fill_(float((thr - self.mu) / self.sd))
        _lib.check(_lib.lib().bb_mlp_walk(self.handle, b["x"].data_ptr(), b["zl"].data_ptr(), n, steps,
                                          b["thr"].data_ptr(), b["xc"].data_ptr(), b["xs"].data_ptr(),
                                          b["sc"].data_ptr(), b["moved"].data_ptr(), _stream_ptr(t)))
        return b["x"].double().cpu().numpy(), b["moved"].cpu().numpy().astype(bool), float(b["sc"][steps].item())

    def predict(self, x):
        x = self._norm(np.atleast_2d(x))
        out = np.empty(x.shape[0])
        for a in range(0, x.shape[0], self.nmax):
            out[a:a + self.nmax] = self.raw_forward(x[a:a + self.nmax])
        return out * self.sd + self.mu
