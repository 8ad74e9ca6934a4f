"""The surrogate network's oracle (oracle/mlp.py), pinned by mathematics: finite differences of the FP64 loss, the
BF16 rounding on hand-worked bit patterns, SiLU's closed forms, Adam's first step."""
import numpy as np

from oracle import mlp as om
from bbarolo2_amd.surrogate import glorot_params


def test_bf16_rounding_hand_worked():
    one = np.float32(1.0)
    e7, e8 = np.float32(2.0 ** -7), np.float32(2.0 ** -8)
    x = np.array([one, one + e8, one + 3 * e8, one + e8 + np.float32(2.0 ** -20), -(one + 3 * e8), 0.0, 3.140625],
                 np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 * e7, 1.0 + e7, -(1.0 + 2 * e7), 0.0, 3.140625], np.float32)
    assert np.array_equal(om.bf16(x), want)
    # BF16 values are fixed points of the rounding; the relative error is at most the unit roundoff 2^-8
    r = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    b = om.bf16(r)
    assert np.array_equal(om.bf16(b), b)
    assert np.max(np.abs(b - r) / np.abs(r)) <= 2.0 ** -8


def test_silu_closed_forms():
    assert om.silu(0.0) == 0.0 and om.dsilu(0.0) == 0.5
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    assert np.allclose(om.dsilu(z), (om.silu(z + h) - om.silu(z - h)) / (2 * h), rtol=1e-7, atol=1e-9)
    assert np.allclose(om.silu(z), z / (1 + np.exp(-z)))


def test_layout_matches_the_product_parameters():
    lay, (wo, bo), n = om.layout(13, 128, 4)
    assert glorot_params(13, 128, 4).size == n == 32 * 128 + 3 * 128 * 128 + 4 * 128 + 129
    assert lay[0][1] == (128, 32) and lay[1][0] == 128 * 32 + 128 and bo == wo + 128


def test_gradient_equals_finite_differences_fp64():
    din, H, nh = 3, 32, 2
    rng = np.random.default_rng(1)
    p = glorot_params(din, H, nh, seed=3).astype(np.float64)
    lay, (wo, bo), n = om.layout(din, H, nh)
    p[lay[0][2]:lay[0][2] + H] = rng.normal(0, 0.3, H)          # non-zero biases
    x = rng.random((40, din))
    t = rng.normal(size=40)
    _, _, g = om.forward_backward(p, x, t, din, H, nh, rounding=False)
    idx = []
    for o, (h, k), b in lay:
        idx += list(rng.choice(np.arange(o, o + h * k).reshape(h, k)[:, :din].ravel(), 8, replace=False))
        idx += list(rng.choice(np.arange(b, b + h), 4, replace=False))
    idx += list(rng.choice(np.arange(wo, wo + H), 6, replace=False)) + [bo]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (om.forward_backward(pp, x, t, din, H, nh, rounding=False)[1] -
              om.forward_backward(pm, x, t, din, H, nh, rounding=False)[1]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 * max(abs(fd), 1e-3), (i, fd, g[i])


def test_rounded_gradient_close_to_fp64():
    din, H, nh = 13, 128, 4
    rng = np.random.default_rng(2)
    p = glorot_params(din, H, nh, seed=5)
    x = rng.random((256, din)).astype(np.float32)
    t = rng.normal(size=256)
    _, l0, g0 = om.forward_backward(p, x, t, din, H, nh, rounding=False)
    _, l1, g1 = om.forward_backward(p, x, t, din, H, nh, rounding=True)
    assert abs(l1 - l0) < 1e-2 * l0
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)


def test_adam_first_step():
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p, m, v = om.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert np.allclose(p[:3], -0.01 * np.sign(g[:3]), rtol=1e-4) and p[3] == 0.0
