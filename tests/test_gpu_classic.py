"""GPU (MI355X): 3DFIT CLASSIC (bbarolo2_amd.classic, RingFit.fit).

  * the stage's rounds (every ring's simplex one step per round, the trials of all the rings in one deposit launch,
    each in its own box) give every ring the fit it gets alone, bit for bit, with the ANALYTIC and the MC model;
  * the 128x128x64 CLASSIC configuration end to end (MC truth of 12 rings, noise 5% of the peak, FREE VROT VDISP,
    FTYPE 2, NORM LOCAL, WFUNC 2): VROT and VDISP recovered, and the LOCAL final model equal to the oracle's
    normalisation (galfit_out.py _local) of the same model;
  * the reference's own quick fit of its NGC 2403 cube (tests/test_bbarolo.py:306-317: rings 105:400:90", free VROT
    VDISP, WFUNC 2, no second stage, the other keywords at their defaults), read from a parameter file, on the
    committed crop of that cube (tests/golden/ngc2403_crop.npz) with the stand-in mask of the source finder: the
    reference's ranges (VROT rising, 60 < VROT[0] < 100, 110 < VROT[-1] < 135).
"""
import json
import os

import numpy as np
import pytest

from bbarolo2_amd.galmod import GalmodPar
from conftest import small_rings

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _ringfit(small_header, gtype):
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import RingFit, Fit3DPar
    torch.cuda.set_device(0)
    dm = DiskModel(small_header, par=GalmodPar(GALMODTYPE=gtype), densflux=True)
    truth = small_rings(8, z0=5.)
    sr, _ = dm.subrings(truth)
    m = dm.array_batch(sr[None], (0, 0), (64, 64))[0].double().cpu().numpy()
    data = m.transpose(2, 0, 1).astype(np.float32)
    mask = data > 0.02 * data.max()
    inr = small_rings(8, z0=5., vrot=130., vdisp=20.)
    return RingFit(dm, data, mask, inr, par=Fit3DPar(FREE="VROT VDISP", NORM="LOCAL"), data_noise=0.01 * float(data.max()))


@pytest.mark.parametrize("gtype", ["ANALYTIC", "MC"])
def test_stage_rounds_equal_rings_alone(small_header, gtype):
    fit = _ringfit(small_header, gtype)
    out, res = fit.fit(stage=1)
    n_batched = fit.nevals
    for ir in range(8):
        one = _ringfit(small_header, gtype)
        r = one.fit_ring(ir, stage=1)
        a = res[ir]
        print(f"[{gtype} ring {ir}] alone {r[:2] if r else None} rounds {a[:2] if a else None}", flush=True)
        assert (r is None) == (a is None)
        if r is not None:
            assert r[0] == a[0] and r[1] == a[1] and np.array_equal(r[2], a[2])
    print(f"[{gtype}] {n_batched} trial models in the stage", flush=True)


def test_config1_classic_end_to_end():
    import torch
    import oracle  # noqa: F401  (the test's checker)
    from bbarolo2_amd.classic import ClassicFit
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import Fit3DPar
    from bbarolo2_amd.synth import config, add_noise
    torch.cuda.set_device(0)
    h, truth, par, noise, seed, _, _ = config(1)
    dm = DiskModel(h, par=par, densflux=True)
    sr, _ = dm.subrings(truth)
    mod = dm.array_batch_mc(sr[None], (0, 0), (128, 128))[0].double().cpu().numpy().transpose(2, 0, 1)
    data = add_noise(mod, noise * mod.max(), dm.cfield, seed)
    inr = truth.copy()
    inr.vrot = np.full(truth.nr, 120.)
    inr.vdisp = np.full(truth.nr, 20.)
    fp = Fit3DPar(FREE="VROT VDISP", FTYPE=2, NORM="LOCAL", WFUNC=2)
    cf = ClassicFit(h, data, inr, par=fp, gpar=par).run()
    r = cf.rings
    print("vrot", np.round(r.vrot, 2), "\ntruth", np.round(truth.vrot, 2), "\nvdisp", np.round(r.vdisp, 2),
          "\nok", cf.fitok, f"\n{cf.nevals} trial models", flush=True)
    inner = slice(1, truth.nr - 1)
    assert all(cf.fitok[inner])
    assert np.abs(r.vrot[inner] - truth.vrot[inner]).max() < 8.
    assert np.abs(r.vdisp[inner] - 10.).max() < 6.
    # the LOCAL final model: the model of the fitted rings normalised by the data in the mask, as the oracle does it
    raw = cf.model(norm="NONE")
    loc = cf.model(norm="LOCAL")
    d = cf.data.astype(np.float64)
    ms = raw.astype(np.float64).sum(axis=0)
    os_ = (d * cf.mask).sum(axis=0)
    f = np.where(ms != 0, os_ / np.where(ms != 0, ms, 1), 0.)
    ref = (raw.astype(np.float64) * f[None]).astype(np.float32)
    assert np.abs(loc - ref).max() <= 1e-6 * np.abs(ref).max()


def _ngc2403():
    from bbarolo2_amd.header import Header
    z = np.load(os.path.join(HERE, "golden", "ngc2403_crop.npz"), allow_pickle=False)
    cards = json.loads(str(z["cards"]))
    for i in range(1, 4):
        cards[f"NAXIS{i}"] = int(cards[f"NAXIS{i}"])
    data = (z["data"].astype(np.float64) * float(z["scale"])).astype(np.float32)
    return Header(cards), data, int(z["offset"])


def test_ngc2403_quick_ranges(tmp_path):
    """From a parameter file with the reference's defaults (NORM AZIM, MC models, FTYPE 2): the quick fit's rings,
    centre in the coordinates of the crop."""
    import torch
    from bbarolo2_amd.params import ParFile
    torch.cuda.set_device(0)
    h, data, off = _ngc2403()
    par = tmp_path / "ngc2403_quick.par"
    par.write_text(f"""FITSFILE  ngc2403_crop.fits
3DFIT     true
RADII     range(105,400,90)
XPOS      {77 - off}
YPOS      {77 - off}
VSYS      132.8
VROT      120
VDISP     8
VRAD      0
Z0        10
INC       60
PA        123.7
FREE      VROT VDISP
WFUNC     2
TWOSTAGE  false
""")
    cf = ParFile.read(str(par)).classic_fit(h, data).run()
    v = np.asarray(cf.rings.vrot)
    print("NGC 2403 vrot", np.round(v, 2), "vdisp", np.round(cf.rings.vdisp, 2), "ok", cf.fitok, flush=True)
    assert cf.par.NORM == "AZIM" and cf.gpar.GALMODTYPE == "MC"
    assert all(cf.fitok)
    assert np.all(np.diff(v) > 0) and 60 < v[0] < 100 and 110 < v[-1] < 135


def test_config1_two_stages_free_inclination():
    # FREE VROT VDISP INC with TWOSTAGE (the reference's default second stage): the first stage's inclinations
    # regularised (REGTYPE auto: their median when they scatter by less than 3 degrees) and fixed, the rings fitted
    # again; the second stage's inclination near the truth and VROT recovered on the inner rings
    import torch
    from bbarolo2_amd.classic import ClassicFit
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import Fit3DPar
    from bbarolo2_amd.regularize import regtypes, regularize
    from bbarolo2_amd.synth import config, add_noise
    torch.cuda.set_device(0)
    h, truth, par, noise, seed, _, _ = config(1)
    dm = DiskModel(h, par=par, densflux=True)
    sr, _ = dm.subrings(truth)
    mod = dm.array_batch_mc(sr[None], (0, 0), (128, 128))[0].double().cpu().numpy().transpose(2, 0, 1)
    data = add_noise(mod, noise * mod.max(), dm.cfield, seed)
    inr = truth.copy()
    inr.vrot = np.full(truth.nr, 120.)
    inr.vdisp = np.full(truth.nr, 20.)
    inr.inc = np.full(truth.nr, 56.)
    fp = Fit3DPar(FREE="VROT VDISP INC", FTYPE=2, NORM="LOCAL", WFUNC=2)
    cf = ClassicFit(h, data, inr, par=fp, gpar=par).run()
    r1, ok1, _ = cf.stage1
    r = cf.rings
    print("stage 1 inc", np.round(r1.inc, 2), "ok", ok1, "\nregularised", cf.regularised, "\nstage 2 inc",
          np.round(r.inc, 2), "\nvrot", np.round(r.vrot, 2), "\ntruth", np.round(truth.vrot, 2), "\nok", cf.fitok,
          f"\n{cf.nevals} trial models", flush=True)
    assert cf.regularised is not None and cf.regularised[0][0] == "inc"
    rtype = regtypes("auto", r1.inc, r1.phi, ok1)[0][0]
    good = np.asarray(ok1)
    want = regularize(r1.radii[good], r1.inc[good], rtype, xout=r1.radii)[1]
    assert np.allclose(r.inc, want)                      # (fixed in the second stage at the regularised values)
    assert np.all(np.abs(r.inc - 60.) < 3.)
    inner = slice(1, truth.nr - 1)
    assert all(cf.fitok[inner])
    assert np.all(np.abs(r.vrot[inner] - truth.vrot[inner]) < 10.)


def test_azim_final_model(small_header):
    """The AZIM final model (finalModel, galfit_out.py:86-159): the model of the fitted rings whose densities are the
    face-on surface densities of the data's intensity map in their annuli (scaled by a power of ten and divided by it
    again: the NONE model of those densities to float32 rounding), whose own intensity map then has about the shape
    of the data's radial profile in the inner rings (the model's units are the deposit's, not the map's)."""
    import torch
    from bbarolo2_amd.classic import ClassicFit, face_on_density
    from bbarolo2_amd.header import delta_vel
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import Fit3DPar
    torch.cuda.set_device(0)
    dm = DiskModel(small_header, par=GalmodPar(GALMODTYPE="ANALYTIC"), densflux=True)
    truth = small_rings(8, z0=5.)
    sr, _ = dm.subrings(truth)
    data = dm.array_batch(sr[None], (0, 0), (64, 64))[0].double().cpu().numpy().transpose(2, 0, 1).astype(np.float32)
    mask = data > 0.01 * data.max()
    inr = small_rings(8, z0=5., vrot=130., vdisp=20.)
    cf = ClassicFit(small_header, data, inr, par=Fit3DPar(FREE="VROT VDISP", NORM="AZIM"),
                    gpar=GalmodPar(GALMODTYPE="ANALYTIC"), mask=mask, twostage=False).run()
    az = cf.model()
    surf = face_on_density(cf.intensity_map(), cf.rings, small_header)
    r = cf.rings.copy()
    on = np.asarray(r.dens) > 0
    r.dens = np.where(on, np.abs(surf), 0.)
    plain = cf.model(norm="NONE", rings=r)
    print("face-on densities of the data", np.round(surf, 4), "ok", cf.fitok, flush=True)
    assert np.all(surf[on] > 0)
    assert np.abs(az - plain).max() <= 1e-5 * np.abs(plain).max()
    mom = (az.astype(np.float64) * mask).sum(axis=0) * abs(float(delta_vel(small_header)))
    surf_model = face_on_density(np.where(mom == 0, np.nan, mom).astype(np.float32), cf.rings, small_header)
    ratio = surf_model[1:6] / surf[1:6]
    print("model/data face-on densities, rings 2-6", np.round(ratio, 3), flush=True)
    assert np.all(np.abs(ratio / np.median(ratio) - 1.) < 0.15)
