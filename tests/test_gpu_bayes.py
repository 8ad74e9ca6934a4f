"""GPU (MI355X): what the trained network gains over the sampler without it, on the 64x64x40 configuration (6 rings,
free VROT VDISP INC: 13-D, DENS from the deconvolved profile of every vector). The same nested sampling run (random
walks, n_live = n_batch = 500, the same seed) plainly and with its walks screened by the BF16 surrogate trained on
the device (bb_mlp, on the sampler's own evaluations): the screened run needs far fewer exact likelihood
evaluations and gives the same posterior (medians within 0.1 of the 68% half-width, widths within 10%, log Z within
0.5). Wall times are printed, not asserted: at 4-7 us per exact evaluation the network's training costs about as
much as the evaluations it saves."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_screen_needs_fewer_exact_evaluations():
    import torch
    from bbarolo2_amd import bayes
    torch.cuda.set_device(0)
    fit = bayes.synthetic_fit(5, chunk=500)
    runs = {}
    for scr in (False, True):
        kw = dict(low_pass="retrain", walk_min_pass=0.1, train_max=8192) if scr else {}
        t0 = time.perf_counter()
        s = bayes.nested_fit(fit, n_live=500, n_batch=500, seed=3, surrogate=scr, bound="rwalk", walks=25, **kw)
        runs[scr] = s
        print(f"surrogate={scr}: {time.perf_counter() - t0:.2f} s, {s.n_like} exact evaluations, "
              f"{s.n_screened} surrogate queries, {s.n_train} trainings, walk end points "
              f"{s.n_walk_passed}/{s.n_walk_judged} passed, log Z {s.log_z:.3f}", flush=True)
    plain, scr = runs[False], runs[True]
    assert scr.n_train > 0 and scr.n_walk_judged > 0
    assert scr.n_like < 0.6 * plain.n_like, (scr.n_like, plain.n_like)
    shift, width = bayes.compare(plain, scr)
    print(f"median shift {shift:.3f} of the 68% half-width, 68% width change {width:.3f}", flush=True)
    assert shift < 0.1 and width < 0.1
    assert abs(scr.log_z - plain.log_z) < 0.5
