"""GPU parity (MI355X) of the 3DFIT ring residual: bb_norm3dfit against the FP64 oracle (oracle/norm.c) on the same
float32 models, for every ftype and normalisation, sides, weights, the strip footprint and the ANALYTIC support.

Tolerance: |gpu - oracle| <= 1e-9 x |oracle| (the same model values; FP64 sums in a different order), blank and
voxel counts equal."""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _gpu_norm(models, data, mask, blo, rings, chan_noise, par, pre=None, ilo=None, ihi=None, thr=None):
    import torch
    from bbarolo2_amd import _lib
    B, by, bx, nz = models.shape
    ny, nx = data.shape[:2]
    dev = "cuda"
    dm = torch.from_numpy(np.ascontiguousarray(models, np.float32)).to(dev)
    dd = torch.from_numpy(np.ascontiguousarray(data, np.float32)).to(dev)
    dk = torch.from_numpy(np.ascontiguousarray(mask, np.uint8)).to(dev)
    dr = torch.from_numpy(np.ascontiguousarray(rings, np.float64)).to(dev)
    dn = torch.from_numpy(np.ascontiguousarray(chan_noise, np.float64)).to(dev)
    sup = pre is not None
    dp = torch.from_numpy(np.ascontiguousarray(pre, np.float32)).to(dev) if sup else None
    di = torch.from_numpy(np.ascontiguousarray(ilo, np.int32)).to(dev) if sup else None
    dh = torch.from_numpy(np.ascontiguousarray(ihi, np.int32)).to(dev) if sup else None
    dt = torch.from_numpy(np.ascontiguousarray(thr, np.float64)).to(dev) if sup else None
    L = _lib.lib()
    wl = int(L.bb_norm_work_size(B, by, bx, nz, int(sup)))
    work = torch.empty(wl, dtype=torch.float64, device=dev)
    val = torch.empty(B, dtype=torch.float64, device=dev)
    nb = torch.empty(B, dtype=torch.int64, device=dev)
    nt = torch.empty(B, dtype=torch.int64, device=dev)
    p = _lib.bb_norm_par(**par)
    _lib.check(L.bb_norm3dfit(dm.data_ptr(), B, by, bx, nz, dd.data_ptr(), dk.data_ptr(), ny, nx, blo[0], blo[1],
                              dr.data_ptr(), dn.data_ptr(), C.byref(p), dp.data_ptr() if sup else None,
                              di.data_ptr() if sup else None, dh.data_ptr() if sup else None, len(ilo) if sup else 0,
                              dt.data_ptr() if sup else None, work.data_ptr(), wl, val.data_ptr(), nb.data_ptr(),
                              nt.data_ptr(), None))
    torch.cuda.synchronize()
    return val.cpu().numpy(), nb.cpu().numpy(), nt.cpu().numpy()


def _case(seed=0, by=37, bx=41, nz=23):
    rng = np.random.default_rng(seed)
    ny, nx = 60, 64
    data = (rng.random((ny, nx, nz)) * 2 - 0.3).astype(np.float32)
    mask = (rng.random((ny, nx, nz)) < 0.6).astype(np.uint8)
    models = (rng.random((3, by, bx, nz)) * (rng.random((3, by, bx, nz)) < 0.8)).astype(np.float32)
    rings = np.array([[18.3, 17.6, 35., 55., 4.2, 11.7], [20.1, 16.2, 120., 70., 0., 9.3],
                      [17.7, 18.9, 250., 20., 6.5, 15.2]])
    return data, mask, models, rings, rng.random(nz) + 0.5


CASES = [dict(ftype=f, normtype=n, side=s, wpow=w, foot=0, hmax=0.)
         for f, n, s, w in ((1, 0, 0, 1.), (2, 0, 0, 2.), (3, 1, 1, -1.), (4, 2, 2, 0.), (2, 1, 0, 1.),
                            (4, 0, 1, 2.))] + [dict(ftype=2, normtype=0, side=0, wpow=1., foot=1, hmax=3.5)]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_norm_parity(k):
    data, mask, models, rings, cn = _case(k)
    c = CASES[k]
    par = dict(side=c["side"], ftype=c["ftype"], normtype=c["normtype"], foot=c["foot"], wpow=c["wpow"],
               data_noise=0.37, bweight=1.5, hmax=c["hmax"])
    blo = (7, 9)
    val, nb, nt = _gpu_norm(models, data, mask, blo, rings, cn, par)
    for b in range(3):
        r = rings[b]
        ref, _, rnb, rnt = oracle.norm3dfit(models[b].astype(np.float64), data, mask, blo, r[0], r[1], r[2], r[3], r[4],
                                            r[5], c["wpow"], c["side"], c["ftype"], cn, 0.37, 1.5, c["normtype"],
                                            foot=c["foot"], hmax=c["hmax"])
        print(f"[norm {k}] trial {b}: gpu {val[b]:.12e} oracle {ref:.12e} blanks {nb[b]}/{rnb} voxels {nt[b]}/{rnt}",
              flush=True)
        assert nb[b] == rnb and nt[b] == rnt
        assert abs(val[b] - ref) <= 1e-9 * abs(ref)


def test_norm_support_parity():
    data, mask, models, rings, cn = _case(11)
    rng = np.random.default_rng(12)
    pre = (rng.random(models.shape) * 2).astype(np.float32)
    ilo = np.array([-1, -2, -2, -2, -1])
    ihi = np.array([1, 2, 2, 2, 1])
    thr = rng.random(models.shape[3]) * 20 + 5
    par = dict(side=0, ftype=2, normtype=0, foot=0, wpow=1., data_noise=0.37, bweight=1., hmax=0.)
    val, nb, nt = _gpu_norm(models, data, mask, (7, 9), rings, cn, par, pre, ilo, ihi, thr)
    for b in range(3):
        r = rings[b]
        ref, _, rnb, rnt = oracle.norm3dfit(models[b].astype(np.float64), data, mask, (7, 9), *r, 1., 0, 2, cn, 0.37,
                                            1., 0, pre=pre[b].astype(np.float64), ilo=ilo, ihi=ihi, thr=thr)
        print(f"[norm support] trial {b}: blanks {nb[b]}/{rnb} voxels {nt[b]}/{rnt}", flush=True)
        assert nb[b] == rnb and nt[b] == rnt and 0 < rnb < rnt
        assert abs(val[b] - ref) <= 1e-9 * abs(ref)


@pytest.mark.parametrize("k", [0, 2, 3, 4])
def test_norm_same_in_larger_box(k):
    """The same trials modelled in a larger box (zeros around them) give the same values bit for bit: the pixel sums
    are added in an order fixed by their coordinates in the cube, not by the box."""
    data, mask, models, _, cn = _case(k)
    rings = np.array([[18.3, 17.6, 35., 50., 4.2, 11.7], [20.1, 16.2, 120., 40., 0., 9.3],
                      [17.7, 18.9, 250., 20., 6.5, 13.0]])
    c = CASES[k]
    par = dict(side=c["side"], ftype=c["ftype"], normtype=c["normtype"], foot=0, wpow=c["wpow"], data_noise=0.37,
               bweight=1.5, hmax=0.)
    blo, dx, dy = (7, 9), 3, 3
    B, by, bx, nz = models.shape
    big = np.zeros((B, by + 8, bx + 10, nz), np.float32)
    big[:, dy:dy + by, dx:dx + bx] = models
    rb = rings.copy()
    rb[:, 0] += dx
    rb[:, 1] += dy
    v1, nb1, nt1 = _gpu_norm(models, data, mask, blo, rings, cn, par)
    v2, nb2, nt2 = _gpu_norm(big, data, mask, (blo[0] - dx, blo[1] - dy), rb, cn, par)
    print(f"[norm box {k}] {v1} {v2}", flush=True)
    assert np.array_equal(v1, v2) and np.array_equal(nb1, nb2) and np.array_equal(nt1, nt2)
