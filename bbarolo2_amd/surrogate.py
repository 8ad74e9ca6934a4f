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
        """Train on unit-cube points x and values y (standardised: mean 0, deviation 1); returns the last loss."""
        y = np.asarray(y, np.float64).ravel()
        if restart:
            self.reset()
        self.mu, self.sd = float(y.mean()), float(y.std()) or 1.0
        ls = self.train_raw(x, (y - self.mu) / self.sd, steps, lr)
        return float(ls[-1]) if len(ls) else float("nan")

    def predict(self, x):
        x = np.atleast_2d(np.asarray(x, np.float32))
        out = np.empty(x.shape[0])
        for a in range(0, x.shape[0], self.nmax):
            out[a:a + self.nmax] = self.raw_forward(x[a:a + self.nmax])
        return out * self.sd + self.mu
