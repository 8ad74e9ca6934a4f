"""GPU (MI355X): likelihood batches replayed as a captured HIP graph give the eager results, bit for bit, on every
replay, with other eager work and synchronisations between the replays (the 64x64x40 configuration: ring maps,
density profile, ANALYTIC deposit, beam, residual; two parts on two streams)."""
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
    assert len(graph["like"]._graphs) == 1


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
    assert len(graph["like"]._graphs) == 2
