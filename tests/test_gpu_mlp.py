"""GPU parity (MI355X) of the likelihood surrogate: the matrix-core forward and backward passes of bb_mlp against the
NumPy oracle with the same BF16 roundings (oracle/mlp.py), Adam on the device against the oracle's step, and the
training itself (deterministic, the loss falls).

Tolerances: outputs |gpu - oracle| <= 1e-4 x max |y| for 99% of the rows and 1e-2 x max |y| for all (FP32 against
FP64 sums; an activation that rounds to the other BF16 neighbour moves its row); gradients: relative L2 error of
every parameter block <= 2e-2; loss 1e-3 relative."""
import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu

DIN, H, NH = 13, 128, 4


def _net(seed=0, nmax=4096):
    from bbarolo2_amd.surrogate import MLPSurrogate
    return MLPSurrogate(DIN, H, NH, nmax=nmax, seed=seed)


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.random((n, DIN)).astype(np.float32)
    t = (-8.0 * np.sum((x - 0.4) ** 2 * np.linspace(0.5, 2.0, DIN), axis=1) + 0.3 * np.sin(6 * x[:, 0]))
    return x, t.astype(np.float32)


@pytest.mark.parametrize("n", [1, 37, 300, 2048])
def test_forward_matches_oracle(n):
    net = _net(seed=2)
    p = net.get_params()
    rng = np.random.default_rng(3)
    lay, _, _ = om.layout(DIN, H, NH)
    for o, (h, k), b in lay:           # non-zero biases
        p[b:b + h] = rng.normal(0, 0.2, h)
    net.set_params(p)
    x, t = _data(n, seed=n)
    y = net.raw_forward(x)
    yo, _, _ = om.forward_backward(p, x, t, DIN, H, NH, rounding=True)
    err = np.abs(y - yo) / max(np.max(np.abs(yo)), 1e-3)
    # (an activation whose FP32 value rounds to the other BF16 neighbour than its FP64 value moves its row by up
    # to ~2^-8 of that activation's share: rare, and only those rows)
    assert np.max(err) <= 1e-2 and np.mean(err <= 1e-4) >= 0.99, (np.max(err), np.mean(err <= 1e-4))


@pytest.mark.parametrize("n", [5, 300, 1000, 4000])
def test_gradient_matches_oracle(n):
    net = _net(seed=4)
    p = net.get_params()
    x, t = _data(n, seed=10 + n)
    g, loss = net.grad(x, t)
    _, lo, go = om.forward_backward(p, x, t, DIN, H, NH, rounding=True)
    assert abs(loss - lo) <= 1e-3 * lo
    lay, (wo, bo), cnt = om.layout(DIN, H, NH)
    blocks = [(o, o + h * k) for o, (h, k), b in lay] + [(b, b + h) for o, (h, k), b in lay] + [(wo, wo + H),
                                                                                                   (bo, bo + 1)]
    for a, b in blocks:
        err = np.linalg.norm(g[a:b] - go[a:b])
        assert err <= 2e-2 * np.linalg.norm(go[a:b]) + 1e-7, (a, b, err, np.linalg.norm(go[a:b]))
    # the padded input columns get no gradient
    o, (h, k), _ = lay[0]
    assert np.all(g[o:o + h * k].reshape(h, k)[:, DIN:] == 0)


def test_adam_step_matches_oracle_and_training_is_deterministic():
    x, t = _data(512, seed=7)
    net = _net(seed=5)
    p0 = net.get_params().astype(np.float64)
    g, _ = net.grad(x, t)
    net.train_raw(x, t, 1, 1e-3)
    p1 = net.get_params().astype(np.float64)
    want, _, _ = om.adam_step(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    assert np.max(np.abs(p1 - want)) <= 1e-6
    # two trainings from the same start: identical parameters
    ls_a = net.train_raw(x, t, 50, 1e-3)
    pa = net.get_params()
    net.set_params(p0.astype(np.float32))
    net.train_raw(x, t, 1, 1e-3)
    ls_b = net.train_raw(x, t, 50, 1e-3)
    assert np.array_equal(pa, net.get_params()) and np.array_equal(ls_a, ls_b)


def test_training_fits_a_smooth_function():
    x, t = _data(4000, seed=8)
    xv, tv = _data(1000, seed=9)
    net = _net(seed=6)
    loss0 = net.grad(x, (t - t.mean()) / t.std())[1]
    last = net.fit(x, t, steps=1500, lr=3e-3)
    assert last < 0.01 * loss0
    pred = net.predict(xv)
    rms = np.sqrt(np.mean((pred - tv) ** 2))
    assert rms < 0.1 * tv.std(), (rms, tv.std())
