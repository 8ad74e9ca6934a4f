"""GPU (MI355X): the REVERSE passes of 3DFIT CLASSIC (ClassicFit with REVERSE, galfit.py:1263-1294).

  * the model of the rings around a ring (RingFit.context: the segments of the rings inside and outside it, built on
    the device and convolved once) against the FP64 oracle's model of the same segments, to 1e-5 of its peak;
  * the residual of a ring with its context (the context added to the trial's model in the ring's box grown by the
    kernel and the strip, the context's sub-clouds added to the ANALYTIC support, the strip region of RINGREGION
    STRIP) against the oracle's residual (orc_norm with foot 1) of the same float32 models, to 1e-9;
  * an inclined disk (INC 80: REVERSE auto) of 8 rings as wide as the beam fitted end to end: one pass of REVERSE
    after the first, VROT recovered (the regridding of narrower rings is tested on the CPU, test_galfit_host.py).
"""
import numpy as np
import pytest
from scipy.signal import convolve2d

import oracle
from bbarolo2_amd.galmod import GalmodPar, PHI_TAB
from conftest import small_rings

pytestmark = pytest.mark.gpu


def _fit(small_header, inc=60., gtype="ANALYTIC", norm="LOCAL"):
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import RingFit, Fit3DPar
    torch.cuda.set_device(0)
    dm = DiskModel(small_header, par=GalmodPar(GALMODTYPE=gtype), densflux=True)
    truth = small_rings(8, z0=5., inc=inc, vrot=np.linspace(60., 150., 8))
    sr, _ = dm.subrings(truth)
    m = dm.array_batch(sr[None], (0, 0), (64, 64))[0].double().cpu().numpy()
    torch.cuda.synchronize()
    data = m.transpose(2, 0, 1).astype(np.float32)
    mask = data > 0.02 * data.max()
    inr = small_rings(8, z0=5., inc=inc, vrot=110., vdisp=20.)
    fit = RingFit(dm, data, mask, inr, par=Fit3DPar(FREE="VROT VDISP", NORM=norm), data_noise=0.01 * float(data.max()))
    return dm, fit, data, mask, truth


def _segments(fit, ir, rings):
    from bbarolo2_amd.galfit import CS_OK, context_dens
    cstat = np.full(rings.nr, CS_OK)
    dens = context_dens(np.linspace(1., 0.5, rings.nr), ir, 1.)
    sub = fit.dm.setup.radsep * 60.
    segs = []
    if ir > 0:
        segs.append(fit.segment(0, ir - 1, dens, rings, cstat, sub))
    if ir < rings.nr - 1:
        segs.append(fit.segment(ir + 1, rings.nr - 1, dens, rings, cstat, sub))
    return segs


@pytest.mark.parametrize("ir", [0, 3, 7])
def test_context_model_vs_oracle(small_header, ir):
    dm, fit, data, mask, truth = _fit(small_header)
    st = dm.setup
    ctx = fit.context(_segments(fit, ir, truth), truth)
    conv = ctx.conv.double().cpu().numpy()
    ref = None
    for seg in _segments(fit, ir, truth):
        sr, _ = st.subrings(seg)
        m = oracle.galmod_analytic(st, sr, ctx.blo, ctx.bhi, 1024, PHI_TAB)
        ref = m if ref is None else ref + m
    ref = oracle.conv_same(ref, dm.cfield, dm.scalefactor)
    err = float(np.abs(conv - ref).max() / np.abs(ref).max())
    print(f"[context of ring {ir}] box {ctx.blo}-{ctx.bhi} max rel err {err:.2e}", flush=True)
    assert ref.max() > 0 and err <= 1e-5
    assert ctx.counts is not None and ctx.grow == (dm.cfield.shape[1] // 2, dm.cfield.shape[0] // 2)


@pytest.mark.parametrize("strip", [True, False])
def test_ring_values_with_context_vs_oracle(small_header, strip):
    import torch
    from bbarolo2_amd.galfit import MC_SUPPORT_NMIN
    dm, fit, data, mask, truth = _fit(small_header, inc=80.)
    st = dm.setup
    beam = abs(float(small_header.Bmaj())) * 3600.
    for ir in (2, 5):
        ctx = fit.context(_segments(fit, ir, truth), truth)
        d = fit.make_dring(ir)
        if strip:
            ctx.hmax = fit.strip_hmax(ir, d.radii[-1], beam)
        trials = []
        for vrot, vdisp in ((100., 10.), (130., 25.)):
            r = d.copy()
            r.vrot[:], r.vdisp[:] = vrot, vdisp
            trials.append(r)
        fit.ctx = ctx
        try:
            vals = fit.values(trials)
        finally:
            fit.ctx = None
        for r, v in zip(trials, vals):
            blo, bhi = dm.box(r, final=False)
            if strip:
                blo, bhi = fit.strip_box(r, blo, bhi, ctx.hmax)
            blo, bhi = ctx.ringbox((blo, bhi))
            sr, nv = st.subrings(r)
            dep = dm.array_batch(sr[None], blo, bhi, convolve=False)[0].cpu().numpy()
            mod = dm.array_batch(sr[None], blo, bhi)[0].cpu().numpy()
            torch.cuda.synchronize()
            cy, cx = ctx.slices(blo, bhi)
            # (the sums of the device, in float32)
            mod = (mod + ctx.conv[cy, cx].cpu().numpy()).astype(np.float64)
            unit = st.subcloud_flux(sr, nv)
            dep = (dep + np.float32(unit) * ctx.counts[cy, cx].cpu().numpy()).astype(np.float64)
            ind = (dm.cfield > 1e-12 / (unit * float(np.max(st.cd2i)) * float(dm.scalefactor))).astype(np.float64)
            lam = np.stack([convolve2d(dep[:, :, z], ind, mode="same") for z in range(dep.shape[2])], axis=2)
            ps = fit.pixscale
            ref, _, nb, nt = oracle.norm3dfit(mod, data.transpose(1, 2, 0), mask.transpose(1, 2, 0).astype(np.uint8),
                                              blo, r.xpos[-1] - blo[0], r.ypos[-1] - blo[1], r.phi[-1], r.inc[-1],
                                              r.radii[0] / ps, r.radii[-1] / ps, fit.par.WFUNC, 0, fit.par.FTYPE,
                                              np.ones(data.shape[0]), fit.data_noise, fit.par.BWEIGHT, 0, pre=lam,
                                              ilo=[0], ihi=[0], thr=MC_SUPPORT_NMIN * unit * np.asarray(st.cd2i, float),
                                              foot=int(strip), hmax=ctx.hmax if strip else 0.)
            print(f"[ring {ir} strip {strip}] gpu {v:.12e} oracle {ref:.12e} blanks {nb}/{nt}", flush=True)
            assert abs(v - ref) <= 1e-9 * abs(ref)


def test_reverse_fit_inclined_disk():
    import torch
    from bbarolo2_amd.classic import ClassicFit
    from bbarolo2_amd.galfit import Fit3DPar
    from bbarolo2_amd.header import emptyheader
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.rings import Rings
    torch.cuda.set_device(0)
    h = emptyheader([64, 64, 48], [-10. / 3600., 10. / 3600., 8.], beam=30. / 3600., obj="EDGE")
    R = np.arange(8) * 30. + 15.
    truth = Rings().setRings(radii=R, xpos=31.5, ypos=31.5, vsys=0., vrot=150. * (2 / np.pi) * np.arctan(R / 60.),
                             vdisp=10., z0=5., dens=1., inc=80., phi=45.)
    dm = DiskModel(h, par=GalmodPar(GALMODTYPE="ANALYTIC"), densflux=True)
    sr, _ = dm.subrings(truth)
    data = dm.array_batch(sr[None], (0, 0), (64, 64))[0].double().cpu().numpy().transpose(2, 0, 1).astype(np.float32)
    mask = data > 0.02 * data.max()
    inr = truth.copy()
    inr.vrot[:], inr.vdisp[:] = 120., 20.
    par = Fit3DPar(FREE="VROT VDISP", NORM="LOCAL", REVERSE="auto", SWEEPS=1)
    cf = ClassicFit(h, data, inr, par=par, gpar=GalmodPar(GALMODTYPE="ANALYTIC"), mask=mask, twostage=False).run()
    assert cf.reverse and cf.sweeps == 1 and cf.strip
    out = cf.rings
    want = np.interp(out.radii, R, truth.vrot)
    print("radii", out.radii, "\nvrot", np.round(out.vrot, 1), "\nwant", np.round(want, 1), "\nok", cf.fitok, flush=True)
    good = np.asarray(cf.fitok, bool)
    assert good.sum() >= out.nr - 1
    assert np.all(np.abs(out.vrot[good][1:] - want[good][1:]) < 15.)
