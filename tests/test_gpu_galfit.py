"""GPU (MI355X) 3DFIT ring fits: RingFit.values (the model of a ring in its box on the device, the beam, and
bb_norm3dfit with the ANALYTIC support) against the oracle's ring residual on the same float32 models, with the
support of each voxel computed independently as a direct convolution of the deposit with the indicator of the
kernel cells above the cut; the MC model; and a first-stage fit that recovers VROT and VDISP of a noiseless disk.

Tolerance of the residuals: 1e-9 relative (the same model values, FP64 sums in another order); blank counts equal."""
import numpy as np
import pytest
from scipy.signal import convolve2d

import oracle
from bbarolo2_amd.galmod import GalmodPar
from conftest import small_rings

pytestmark = pytest.mark.gpu


def _setup(small_header, galmodtype="ANALYTIC"):
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galfit import RingFit, Fit3DPar
    dm = DiskModel(small_header, par=GalmodPar(GALMODTYPE=galmodtype), densflux=True)
    truth = small_rings(8, z0=5.)
    sr, _ = dm.subrings(truth)
    m = dm.array_batch(sr[None], (0, 0), (64, 64))[0].double().cpu().numpy()       # (y, x, z)
    torch.cuda.synchronize()
    data = m.transpose(2, 0, 1).astype(np.float32)
    mask = data > 0.02 * data.max()
    inr = small_rings(8, z0=5., vrot=130., vdisp=20.)
    fit = RingFit(dm, data, mask, inr, par=Fit3DPar(FREE="VROT VDISP", NORM="LOCAL"), data_noise=0.01 * float(data.max()))
    return dm, fit, data, mask


def _expected(dm, fit, r, data, mask, support):
    from bbarolo2_amd.galfit import MC_SUPPORT_NMIN
    st = dm.setup
    blo, bhi = dm.box(r, final=False)
    sr, nv = st.subrings(r)
    dep = dm.array_batch(sr[None], blo, bhi, convolve=False)[0].double().cpu().numpy() if support else None
    if st.galmodtype == "MC":
        mod = dm.array_batch_mc(sr[None], blo, bhi)[0].double().cpu().numpy()
    else:
        mod = dm.array_batch(sr[None], blo, bhi)[0].double().cpu().numpy()
    kw = {}
    if support:
        unit = st.subcloud_flux(sr, nv)
        K = dm.cfield
        assert K.shape[0] % 2 == 1 and K.shape[1] % 2 == 1
        ind = (K > 1e-12 / (unit * float(np.max(st.cd2i)) * float(dm.scalefactor))).astype(np.float64)
        lam = np.stack([convolve2d(dep[:, :, z], ind, mode="same") for z in range(dep.shape[2])], axis=2)
        kw = dict(pre=lam, ilo=[0], ihi=[0], thr=MC_SUPPORT_NMIN * unit * np.asarray(st.cd2i, float))
    ps = fit.pixscale
    return oracle.norm3dfit(mod, data.transpose(1, 2, 0), mask.transpose(1, 2, 0).astype(np.uint8), blo,
                            r.xpos[-1] - blo[0], r.ypos[-1] - blo[1], r.phi[-1], r.inc[-1], r.radii[0] / ps,
                            r.radii[-1] / ps, fit.par.WFUNC, 0, fit.par.FTYPE, np.ones(data.shape[0]), fit.data_noise,
                            fit.par.BWEIGHT, 0, **kw)


@pytest.mark.parametrize("gtype", ["ANALYTIC", "MC"])
def test_ring_values_parity(small_header, gtype):
    dm, fit, data, mask = _setup(small_header, gtype)
    for ir in (1, 4, 7):
        d = fit.make_dring(ir)
        trials = []
        for vrot, vdisp in ((150., 10.), (120., 25.), (175., 4.)):
            r = d.copy()
            r.vrot[:], r.vdisp[:] = vrot, vdisp
            trials.append(r)
        vals = fit.values(trials)
        for r, v in zip(trials, vals):
            ref, _, nb, nt = _expected(dm, fit, r, data, mask, gtype == "ANALYTIC")
            print(f"[ring {ir} {gtype}] vrot {r.vrot[0]} vdisp {r.vdisp[0]}: gpu {v:.12e} oracle {ref:.12e} "
                  f"blanks {nb}/{nt}", flush=True)
            assert abs(v - ref) <= 1e-9 * abs(ref)
            if gtype == "ANALYTIC":
                assert 0 < nb < nt


def test_ring_fit_recovers_disk(small_header):
    """Noiseless data of a disk with VROT 150 and VDISP 10; the fit starts from 130 and 20."""
    dm, fit, data, mask = _setup(small_header)
    for ir in (1, 3, 5):
        fit.rng = np.random.default_rng([12345, 1, ir])
        ok, fmin, p = fit.minimize(fit.make_dring(ir))
        print(f"[fit ring {ir}] ok {ok} min {fmin:.6e} vrot {p[0]:.3f} vdisp {p[1]:.3f} evals {fit.nevals}", flush=True)
        assert ok and abs(p[0] - 150.) < 4. and abs(p[1] - 10.) < 4.
