"""ORACLE -- test infrastructure only (imported by tests/; never by the product path). NumPy restatement of the
likelihood surrogate network of bb_mlp (din -> H (SiLU) x nh -> 1) and of its gradient, with the BF16 roundings of
the device at the same points:

  * inputs, weights and the activations fed to a matrix product are rounded to BF16 (round to nearest even);
  * products are summed in FP64 here (FP32 on the matrix cores), pre-activations and the output in FP64;
  * the layer gradients fed to a matrix product are rounded to BF16 after the SiLU derivative is applied;
  * loss = 0.5 mean (y - t)^2; the gradient of the FP32 master parameters is that of their BF16 copies.

The reference's learned component is nautilus's regression network (bbarolo/bayesian/nautilus/mlp.py:17-222), a
NumPy ReLU network trained with Adam on the live points' likelihood ranks; this surrogate regresses the
log-likelihood instead, so the reference holds no vectors for it. Pinned in tests/test_oracle_mlp.py by
mathematics: with the roundings off the analytic gradient equals central finite differences of the FP64 loss
(every parameter block, to 1e-7 relative), the BF16 rounding matches hand-worked bit patterns (ties to even), SiLU
and its derivative match their closed forms, and Adam's first step moves every parameter by lr against the sign
of its gradient.
"""

import numpy as np

DP = 32     # padded input width


def bf16(x):
    """Round float32 values to BF16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffffffff
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def layout(din, H, nh):
    """[(W offset, W shape, b offset)] of the hidden layers, then (w offset, b offset) of the output, and the count."""
    out, o = [], 0
    for l in range(nh):
        K = DP if l == 0 else H
        out.append((o, (H, K), o + H * K))
        o += H * K + H
    return out, (o, o + H), o + H + 1


def forward_backward(p, x, t, din, H, nh, rounding=True):
    """(y, loss, grad) of the network with parameters p on inputs x (n, din) and targets t (n,); rounding=False
    is the same network in plain FP64 (the finite-difference check)."""
    rd = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    p = np.asarray(p, np.float64)
    n = x.shape[0]
    lay, (wo, bo), cnt = layout(din, H, nh)
    xin = np.zeros((n, DP))
    xin[:, :din] = x
    a = rd(xin).astype(np.float64)
    acts, zs, Wb = [a], [], []
    for (o, (h, k), b) in lay:
        W = rd(p[o:o + h * k].reshape(h, k)).astype(np.float64)
        Wb.append(W)
        z = a @ W.T + p[b:b + h]
        zs.append(z)
        a = rd(silu(z)).astype(np.float64)
        acts.append(a)
    w = rd(p[wo:wo + H]).astype(np.float64)
    y = a @ w + p[bo]
    r = y - t
    loss = 0.5 * np.mean(r * r)
    g = np.zeros(cnt)
    dy = r / n
    g[wo:wo + H] = acts[-1].T @ dy
    g[bo] = dy.sum()
    d = rd(dy[:, None] * w[None, :] * dsilu(zs[-1])).astype(np.float64)
    for l in range(nh - 1, -1, -1):
        o, (h, k), b = lay[l]
        g[o:o + h * k] = (d.T @ acts[l]).ravel()
        g[b:b + h] = d.sum(axis=0)
        if l > 0:
            d = rd((d @ Wb[l]) * dsilu(zs[l - 1])).astype(np.float64)
    return y, loss, g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (step counts from 1): (p, m, v) updated copies."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
