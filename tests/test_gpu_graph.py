"""GPU (MI355X): likelihood batches replayed as a captured HIP graph (bb_ctx_set_graphs: the first batch of a size
eager, the second captured, later ones replayed) give the eager results, bit for bit, on every replay, with other
eager work and synchronisations between the replays (the 64x64x40 configuration: device tables, ring maps, density
profile, ANALYTIC deposit, beam, residual; two halves on two streams)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager():
    import torch
    from bbarolo2_amd.bayes import synthetic_fit
    torch.cuda.set_device(0)
    eager = synthetic_fit(5, chunk=64)
    graph = synthetic_fit(5, chunk=64, graphs=True)
    assert graph["like"].use_graphs and not eager["like"].use_graphs
    lo, hi = eager["lo"], eager["hi"]
    rng = np.random.default_rng(4)
    for rep in range(4):
        theta = lo + rng.random((64, len(lo))) * (hi - lo)
        if rep == 2:
            theta[:, :] = eager["theta_true"]          # (the same vector in every row)
        le = eager["like"].loglike(theta)
        lg = graph["like"].loglike(theta)
        print(f"replay {rep}: max |eager - graph| = {np.max(np.abs(le - lg)):.3e}", flush=True)
        assert np.array_equal(le, lg), rep
        # eager work in between
        torch.ones(1 << 20, device="cuda").sum().item()
        torch.cuda.synchronize()
    assert graph["like"].graph_replays() == 3


def test_graph_replays_with_alternating_batch_sizes():
    """A graph per batch size: replays of one size after batches of another still match the eager results (the
    buffers a graph writes into are held while it lives)."""
    import torch
    from bbarolo2_amd.bayes import synthetic_fit
    torch.cuda.set_device(0)
    eager = synthetic_fit(5, chunk=64)
    graph = synthetic_fit(5, chunk=64, graphs=True)
    lo, hi = eager["lo"], eager["hi"]
    rng = np.random.default_rng(5)
    for rep, n in enumerate((64, 32, 64, 32, 64)):
        theta = lo + rng.random((n, len(lo))) * (hi - lo)
        le = eager["like"].loglike(theta)
        lg = graph["like"].loglike(theta)
        print(f"batch {rep} ({n} rows): max |eager - graph| = {np.max(np.abs(le - lg)):.3e}", flush=True)
        assert np.array_equal(le, lg), rep
    assert graph["like"].graph_replays() == 3


def test_graph_replays_mc():
    """The MC model's chain (its buffers cleared by kernels, not memsets) replayed as a graph, bit for bit."""
    import torch
    from bbarolo2_amd.bayes import fit_of
    from bbarolo2_amd.synth import config
    torch.cuda.set_device(0)
    h, truth, par, noise, seed, spec, desc = config(5)
    par.GALMODTYPE = "MC"
    cfg = (h, truth, par, noise, seed, [("vrot", "ring"), ("vdisp", "ring")], desc)
    eager = fit_of(cfg, chunk=32)
    graph = fit_of(cfg, chunk=32, graphs=True)
    lo, hi = eager["lo"], eager["hi"]
    rng = np.random.default_rng(6)
    for rep in range(4):
        theta = lo + rng.random((32, len(lo))) * (hi - lo)
        le = eager["like"].loglike(theta)
        lg = graph["like"].loglike(theta)
        print(f"MC replay {rep}: max |eager - graph| = {np.max(np.abs(le - lg)):.3e}", flush=True)
        assert np.array_equal(le, lg), rep
    assert graph["like"].graph_replays() == 3
