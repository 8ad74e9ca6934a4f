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


@pytest.mark.parametrize("n,steps", [(37, 5), (600, 40)])
def test_walk_matches_its_steps(n, steps):
    """bb_mlp_walk (the whole walk as one captured graph) against its steps replayed from the host with the same
    Gaussian steps and the scales the walk recorded: the same proposals, the network's forward pass on each, the same
    acceptances, end points and moved flags, bit for bit; a second call with other values replays the graph."""
    import torch
    from bbarolo2_amd import _lib
    net = _net(seed=4)
    x, t = _data(2048, seed=5)
    net.fit(x, t, steps=200)
    rng = np.random.default_rng(n)
    dev = torch.device("cuda", torch.cuda.current_device())
    for call in range(2):
        x0 = (0.3 + 0.4 * rng.random((n, DIN))).astype(np.float32)
        zl = (rng.normal(size=(steps, n, DIN)) * 0.05).astype(np.float32)
        thr = np.float32(np.median(net.raw_forward(x0)) - 0.2)
        xc, xs = net.xc.astype(np.float32), net.xs.astype(np.float32)
        dx = torch.from_numpy(x0.copy()).to(dev)
        dz = torch.from_numpy(zl).to(dev)
        dthr = torch.tensor([thr], dtype=torch.float32, device=dev)
        dxc, dxs = torch.from_numpy(xc).to(dev), torch.from_numpy(xs).to(dev)
        dsc = torch.zeros(steps + 1, dtype=torch.float32, device=dev)
        dsc[0] = 0.7 + 0.2 * call
        dmv = torch.empty(n, dtype=torch.uint8, device=dev)
        if call == 1:                       # the same buffers as a replay would use: the walk writes in place
            dx, dz, dthr, dxc, dxs, dsc, dmv = keep
            dx.copy_(torch.from_numpy(x0))
            dz.copy_(torch.from_numpy(zl))
            dthr.fill_(float(thr))
            dsc.zero_()
            dsc[0] = 0.9
        keep = (dx, dz, dthr, dxc, dxs, dsc, dmv)
        _lib.check(_lib.lib().bb_mlp_walk(net.handle, dx.data_ptr(), dz.data_ptr(), n, steps, dthr.data_ptr(),
                                          dxc.data_ptr(), dxs.data_ptr(), dsc.data_ptr(), dmv.data_ptr(), None))
        torch.cuda.synchronize()
        sc = dsc.cpu().numpy()
        xh = x0.copy()
        mv = np.zeros(n, bool)
        nacc = 0
        for s in range(steps):
            prop = (xh + zl[s] * sc[s]).astype(np.float32)
            pn = ((prop - xc) / xs).astype(np.float32)
            y = net.raw_forward(pn).astype(np.float32)
            acc = np.all((prop >= 0) & (prop <= 1), axis=1) & (y >= thr)
            xh[acc] = prop[acc]
            mv |= acc
            nacc += int(acc.sum())
            want = np.float32(min(max(np.float32(sc[s]) * np.exp(np.float32(acc.sum() / n) - np.float32(0.3)),
                                      1e-4), 3.0))
            assert abs(sc[s + 1] - want) <= 4e-7 * want, (s, sc[s + 1], want)
        print(f"[walk n={n} steps={steps} call {call}] moved {mv.sum()}, {nacc} of {n * steps} steps accepted, "
              f"final scale {sc[-1]:.4f}", flush=True)
        assert 0 < nacc < n * steps
        assert np.array_equal(dx.cpu().numpy(), xh) and np.array_equal(dmv.cpu().numpy().astype(bool), mv)
