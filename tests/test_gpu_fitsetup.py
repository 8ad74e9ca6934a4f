"""GPU (MI355X): the per-fit set-up of a fit from a cube, its mask, the initial rings and the priors
(fitsetup.setup_likelihood): the noise rms bit for bit and the measured correlation of the noise (1e-9) against the
oracle, the likelihood mask bit for bit, and the log-likelihoods of the resulting context against the FP64 oracle
pipeline given the same set-up (tolerance 1e-5 of the residual term: FP32 models against FP64)."""
import numpy as np
import pytest

from oracle import pipeline
from oracle.noise import noise_rms as oracle_rms, noise_correlation as oracle_nc, dilate_mask as oracle_dilate
from bbarolo2_amd.galmod import GalmodPar

pytestmark = pytest.mark.gpu


def test_setup_likelihood_vs_oracle():
    import torch
    from test_gpu_ctx import _disk, _data, _compare
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.fitsetup import setup_likelihood
    from bbarolo2_amd.likelihood import FreePars
    torch.cuda.set_device(0)
    h, r = _disk()
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    data, sig = _data(h, r, par)
    mask = data > 4 * sig
    data[:, :, :3] = 0.                                  # (blanks)
    spec = [("vrot", "ring"), ("vdisp", "single"), ("inc", "single"), ("xpos", "single")]
    free = FreePars(spec, r.nr)
    lo = np.r_[np.zeros(r.nr), 0., 50., 29.]
    hi = np.r_[np.full(r.nr, 300.), 50., 70., 34.]
    dm = DiskModel(h, par=par)
    like, s = setup_likelihood(dm, data, mask, r, free, lo, hi, chunk=8)
    bpix = abs(h.Bmaj() / h.Cdelt(0))
    assert s.noise["noiserms"] == oracle_rms(data, mask) == s.sigma
    nc = oracle_nc(np.where(data == 0, np.nan, data), mask, bpix)
    got = s.noise["noise_correlation"]
    for k in ("spatial", "spectral", "spatial_core", "spectral_core"):
        assert got[k] == pytest.approx(nc[k], rel=1e-9), k
    assert got["spatial_from"] == nc["spatial_from"] and got["spectral_from"] == nc["spectral_from"]
    print(f"sigma {s.sigma:.6g} (true {sig:.6g}); {s.noise['note']}; box {s.likebox}, region {int(s.region.sum())} px",
          flush=True)
    lm = oracle_dilate(mask, bpix, 3)
    assert np.array_equal(s.likemask.cpu().numpy().astype(bool), lm)
    orc = pipeline.Likelihood(h, par, r, spec, data, lm, s.region, s.likebox, s.sigma, s.noise["likefactor"],
                              likefactor2d=s.noise["likefactor2d"])
    rng = np.random.default_rng(4)
    theta = np.column_stack([rng.uniform(100, 200, (6, r.nr)), rng.uniform(5, 20, (6, 1)), rng.uniform(52, 68, (6, 1)),
                             rng.uniform(30, 33, (6, 1))])
    ll = _compare(like, orc, theta, "set-up")
    assert np.all(np.isfinite(ll))


def test_reference_default_free_set_vs_oracle():
    """The reference's default FREE of METHOD GLOBAL / BAYESIAN (keywords.py:700: VROT VDISP XPOS_SINGLE YPOS_SINGLE
    VSYS_SINGLE INC_SINGLE PA_SINGLE Z0_SINGLE RADMAX FRACERR), read from a parameter file without FREE: the set-up
    and the log-likelihoods against the oracle at prior draws (RADMAX from inside the disk to beyond it)."""
    import torch
    from test_gpu_ctx import _disk, _data, _compare
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.fitsetup import setup_likelihood
    from bbarolo2_amd.likelihood import FreePars
    from bbarolo2_amd.params import ParFile
    torch.cuda.set_device(0)
    h, r = _disk()
    par = GalmodPar(GALMODTYPE="ANALYTIC")
    data, sig = _data(h, r, par)
    mask = data > 4 * sig
    spec = ParFile({"METHOD": "GLOBAL"}).free_spec(r.nr)
    assert [k for k, _ in spec] == ["vrot", "vdisp", "xpos", "ypos", "vsys", "inc", "phi", "z0", "radmax", "fracerr"]
    free = FreePars(spec, r.nr)
    nr = r.nr
    lo = np.r_[np.zeros(nr), np.zeros(nr), 29., 30., -20., 50., 35., 0., 100., 0.]
    hi = np.r_[np.full(nr, 300.), np.full(nr, 50.), 34., 35., 30., 70., 55., 20., 200., 0.2]
    like, s = setup_likelihood(DiskModel(h, par=par), data, mask, r, free, lo, hi, chunk=8)
    assert like.update_prof and like.box_free and like.ann_cap
    orc = pipeline.Likelihood(h, par, r, spec, data, s.likemask.cpu().numpy().astype(bool), s.region, s.likebox,
                              s.sigma, s.noise["likefactor"], likefactor2d=s.noise["likefactor2d"])
    rng = np.random.default_rng(8)
    theta = lo + rng.random((5, len(lo))) * (hi - lo)
    theta[:, :nr] = rng.uniform(120., 180., (5, nr))
    theta[:, nr:2 * nr] = rng.uniform(5., 15., (5, nr))
    ll = _compare(like, orc, theta, "default FREE")
    assert np.all(np.isfinite(ll))
