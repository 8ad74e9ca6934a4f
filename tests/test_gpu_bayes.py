"""GPU (MI355X): what the trained network gains over the sampler without it. The same nested sampling run (random
walks, n_live = n_batch, the same seed) plainly and with its walks screened by the BF16 surrogate trained on the
device (bb_mlp, on the sampler's own evaluations; the walks run on the device): the screened run needs far fewer
exact likelihood evaluations and gives the same posterior (medians within 0.1 of the 68% half-width, widths within
10%, log Z within three times the scatter of two runs, sqrt((H1 + H2) / n_live) with H the information of each).

  * the 64x64x40 configuration (6 rings, free VROT VDISP INC: 13-D), 1000 live points: fewer than a third of the
    exact evaluations; wall times printed, not asserted (at ~3 us per exact evaluation the network's training and
    walks cost about as much as the evaluations they save);
  * the 128x128x64 configuration (8 rings, free VROT VDISP: 16-D, batches as captured HIP graphs), 1000 live points:
    fewer than a fifth of the exact evaluations, and at least twice as fast in wall time."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCREEN = dict(low_pass="retrain", walk_min_pass=0.1, surrogate_walks=100)


def _pair(cfg, n_live, train_max, train_steps, graphs=False):
    import torch
    from bbarolo2_amd import bayes
    torch.cuda.set_device(0)
    fit = bayes.synthetic_fit(cfg, chunk=n_live, graphs=graphs)
    runs, secs = {}, {}
    for scr in (False, True):
        kw = dict(SCREEN, train_max=train_max, train_steps=train_steps) if scr else {}
        t0 = time.perf_counter()
        s = bayes.nested_fit(fit, n_live=n_live, n_batch=n_live, seed=3, surrogate=scr, bound="rwalk", walks=25, **kw)
        secs[scr] = time.perf_counter() - t0
        runs[scr] = s
        print(f"[config {cfg}] surrogate={scr}: {secs[scr]:.2f} s, {s.n_like} exact evaluations, {s.n_screened} "
              f"surrogate steps, {s.n_train} trainings, walk end points {s.n_walk_passed}/{s.n_walk_judged} passed, "
              f"log Z {s.log_z:.3f}", flush=True)
    plain, scr = runs[False], runs[True]
    shift, width = bayes.compare(plain, scr)
    print(f"[config {cfg}] median shift {shift:.3f} of the 68% half-width, 68% width change {width:.3f}, "
          f"speedup {secs[False] / secs[True]:.2f}", flush=True)
    assert scr.n_train > 0 and scr.n_walk_judged > 0
    assert shift < 0.1 and width < 0.1
    sd = np.sqrt((plain.information + scr.information) / n_live)
    print(f"[config {cfg}] log Z {plain.log_z:.3f} / {scr.log_z:.3f}, information {plain.information:.1f} / "
          f"{scr.information:.1f} nats, scatter {sd:.3f}", flush=True)
    assert abs(scr.log_z - plain.log_z) < 3 * sd
    return plain, scr, secs


def test_surrogate_screen_needs_fewer_exact_evaluations():
    plain, scr, _ = _pair(5, 1000, 8192, 1000)
    assert scr.n_like < plain.n_like / 3, (scr.n_like, plain.n_like)


def test_surrogate_screen_faster_on_the_128_cube():
    plain, scr, secs = _pair(3, 1000, 16384, 300, graphs=True)
    assert scr.n_like < plain.n_like / 5, (scr.n_like, plain.n_like)
    assert secs[False] > 2.0 * secs[True], secs
