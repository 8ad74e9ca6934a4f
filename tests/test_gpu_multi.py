"""GPU (MI355X): the uneven multi-galaxy batch and the tile tuner.

- every tile shape that the deposit accepts gives the same log-likelihoods, bit for bit (the deposit's sums are
  exact integers in any order), and the tuner keeps one of them;
- a GalaxyBatch step (all galaxies launched on four streams, one synchronisation) gives each galaxy's own
  loglike, bit for bit, including rows with a negative positive-only parameter (-inf) and a galaxy whose batch is
  all such rows.
The galaxies are the smallest of the 64 of the batch (seed 11), so that the set-up stays short."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small(n):
    from bbarolo2_amd.synth import multi_galaxy_configs
    cfgs = multi_galaxy_configs()
    vox = [c[0].DimAx(0) * c[0].DimAx(1) * c[0].DimAx(2) for c in cfgs]
    return [cfgs[i] for i in np.argsort(vox)[:n]]


def test_tile_shapes_same_result_and_tuner():
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.multi import build_fits, draw_thetas
    from bbarolo2_amd.tuner import tile_candidates, tune_tiles, get_tile
    torch.cuda.set_device(0)
    fits, _ = build_fits(_small(1), chunk=32)
    f = fits[0]
    like = f["like"]
    theta = draw_thetas(fits, 32, seed=2)[0]
    ref = like.loglike(theta)
    assert np.all(np.isfinite(ref))
    h = like.handle
    shapes = tile_candidates(h)
    print("candidates", shapes, flush=True)
    assert len(shapes) >= 3
    for s in shapes:
        _lib.check(_lib.lib().bb_model_set_tile(h, s[0], s[1]))
        v = like.loglike(theta)
        print(s, float(np.max(np.abs(v - ref))), flush=True)
        assert np.array_equal(v, ref), s
    best, times = tune_tiles(like, theta, reps=1, cache={})
    print("tuned", best, times, flush=True)
    assert best in shapes and set(times) == set(shapes)
    assert tuple(sorted(get_tile(h), reverse=True)) == best
    assert np.array_equal(like.loglike(theta), ref)


def test_galaxy_batch_matches_each_galaxy():
    import torch
    from bbarolo2_amd.multi import build_fits, draw_thetas, GalaxyBatch
    torch.cuda.set_device(0)
    fits, secs = build_fits(_small(5), chunk=24)
    thetas = draw_thetas(fits, 24, seed=3)
    thetas[1] = thetas[1][:7].copy()
    thetas[2][3, fits[2]["free"].index["vdisp"][0]] = -1.0          # (one row -inf)
    thetas[3][:, fits[3]["free"].index["inc"][0]] = -5.0             # (the whole batch -inf)
    gb = GalaxyBatch([f["like"] for f in fits], n_streams=4)
    got = gb.loglike(thetas)
    for g, (f, th) in enumerate(zip(fits, thetas)):
        want = f["like"].loglike(th)
        fin = np.isfinite(want)
        diff = float(np.max(np.abs(got[g][fin] - want[fin]))) if fin.any() else 0.0
        print(f"galaxy {g} {f['desc']}: {len(th)} rows, max diff {diff:.3e}", flush=True)
        assert got[g].shape == want.shape
        assert np.array_equal(got[g], want), g
    assert np.isneginf(got[2][3]) and np.all(np.isneginf(got[3]))
    out, ms, done, host = gb.timed_step(thetas)
    assert all(np.array_equal(a, b) for a, b in zip(out, got))
    assert ms > 0 and host > 0 and np.all(done > 0) and np.all(done <= ms * 1.001)
    alone = gb.alone_ms(thetas)
    assert np.isnan(alone[3]) and np.all(alone[[0, 1, 2, 4]] > 0)
    tiles = gb.tune(thetas, reps=1)
    assert tiles[3] is None and all(t is not None for i, t in enumerate(tiles) if i != 3)
    again = gb.loglike(thetas)
    assert all(np.array_equal(a, b) for a, b in zip(again, got))


def test_galaxy_batch_graph_replays_match_eager():
    import torch
    from bbarolo2_amd.multi import build_fits, draw_thetas, GalaxyBatch
    torch.cuda.set_device(0)
    fits, _ = build_fits(_small(3), chunk=16)
    ref, _ = build_fits(_small(3), chunk=16)          # (the same fits, eager)
    gb = GalaxyBatch([f["like"] for f in fits], n_streams=4, graphs=True)
    for rep in range(4):
        thetas = draw_thetas(fits, 16, seed=10 + rep)
        got = gb.graph_step(thetas)
        # other work in between: every galaxy with another batch size
        for f, th in zip(fits, thetas):
            f["like"].loglike(th[:5])
        want = [f["like"].loglike(th) for f, th in zip(ref, thetas)]
        for g in range(3):
            print(f"replay {rep} galaxy {g}: max diff {np.max(np.abs(got[g] - want[g])):.3e}", flush=True)
            assert np.array_equal(got[g], want[g]), (rep, g)
        torch.ones(1 << 20, device="cuda").sum().item()
    assert all(f["like"].graph_replays() >= 3 for f in fits)
    # a row with a negative positive-only parameter: fewer vectors, a batch size of its own (eager the first time)
    before = fits[0]["like"].graph_replays()
    thetas[0][2, fits[0]["free"].index["vdisp"][0]] = -1.0
    got = gb.graph_step(thetas)
    assert np.isneginf(got[0][2]) and fits[0]["like"].graph_replays() == before
    want = ref[0]["like"].loglike(thetas[0])
    assert np.array_equal(got[0], want)
