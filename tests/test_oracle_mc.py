"""Pins of the oracle's counter-based Monte-Carlo GALMOD (oracle/mc.c, CPU): the Philox4x32-10 known-answer
vectors, the FP32 functions against libm, the deviates' distributions, and the reference's own checks of the
MC model: flux and spectrum against the ANALYTIC model (tests/test_bbarolo.py:78-87), box = crop."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from bbarolo2_amd.galmod import GalmodSetup, GalmodPar, model_rings, nzq_for, PHI_TAB
from conftest import small_rings


def test_philox_known_answers():
    """Random123's known-answer vectors of philox4x32_10 (Salmon et al., SC'11)."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_fp32_functions_against_libm():
    L = oracle.lib()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(1e-7, 1, 3000), rng.uniform(1, 4, 1000), [2 ** -23, 0.5, 1.0, 2.0]])
    x = x.astype(np.float32)
    lg = np.array([L.orc_mc_log(float(v)) for v in x])
    assert np.abs(lg - np.log(x.astype(np.float64))).max() < 3e-7 * np.maximum(1, np.abs(np.log(x))).max()
    sq = np.array([L.orc_mc_sqrt(float(v)) for v in x], np.float32)
    assert np.all(np.abs(sq.astype(np.float64) - np.sqrt(x.astype(np.float64))) <= np.spacing(sq) * 1.0)
    rc = np.array([L.orc_mc_rcp(float(v)) for v in x], np.float32)
    assert np.all(np.abs(rc.astype(np.float64) * x - 1) < 3e-7)
    t = rng.uniform(0, 1, 3000).astype(np.float32)
    cs = []
    for v in t:
        c, s = C.c_float(), C.c_float()
        L.orc_mc_sincos_turn(float(v), C.byref(c), C.byref(s))
        cs.append((c.value, s.value))
    cs = np.array(cs)
    ang = 2 * np.pi * t.astype(np.float64)
    assert np.abs(cs[:, 0] - np.cos(ang)).max() < 3e-7 and np.abs(cs[:, 1] - np.sin(ang)).max() < 3e-7


def _stream(n, seed=5):
    out = np.array([oracle.philox([i, 7, 0, 0], [seed, 3]) for i in range(n)], np.uint32)
    return out


def test_deviates_distributions():
    L = oracle.lib()
    x = _stream(40000)
    n0, n1 = [], []
    for a, b in x[:, :2]:
        p, q = C.c_float(), C.c_float()
        L.orc_mc_box_muller(int(a), int(b), C.byref(p), C.byref(q))
        n0.append(p.value)
        n1.append(q.value)
    n = np.array(n0 + n1)
    assert abs(n.mean()) < 0.01 and abs(n.std() - 1) < 0.01 and abs(np.mean(n ** 4) - 3) < 0.1
    assert abs(np.corrcoef(n0, n1)[0, 1]) < 0.02
    devs = {lt: np.array([L.orc_mc_layer_dev(lt, int(a), int(b)) for a, b in x[:20000, 2:]]) for lt in (2, 3, 4, 5)}
    assert abs(devs[2].std() - math.pi / (2 * math.sqrt(3))) < 0.02         # sech2: logistic / 2
    assert abs(devs[3].std() - math.sqrt(2)) < 0.03                          # exponential: Laplace(1)
    assert abs(np.median(np.abs(devs[4])) - 1) < 0.03                        # Lorentzian: Cauchy
    assert abs(devs[5].std() - 1 / math.sqrt(3)) < 0.01 and np.abs(devs[5]).max() < 1


def _mc(st, r, blo, bhi, seed=1234):
    sr, nv = st.subrings(model_rings(r, "CENTRED", st.radsep))
    return oracle.galmod_mc(st, sr, nv, blo, bhi, seed)


def test_mc_vs_analytic(small_header):
    """The MC model has the flux of the ANALYTIC one within 2% and the same spectrum (test_bbarolo.py:78-87)."""
    st = GalmodSetup(small_header, GalmodPar(), densflux=True)
    r = small_rings()
    counts, mc, q = _mc(st, r, (0, 0), (64, 64))
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    an = oracle.galmod_analytic(st, sr, (0, 0), (64, 64), nzq_for(sr), PHI_TAB)
    assert mc.shape == an.shape and counts.sum() > 0
    assert abs(mc.sum() / an.sum() - 1) < 0.02
    assert np.corrcoef(mc.sum(axis=(0, 1)), an.sum(axis=(0, 1)))[0, 1] > 0.99


@pytest.mark.parametrize("ltype", [1, 2, 3, 4])
def test_mc_box_is_crop_and_layers(small_header, ltype):
    """A box model is the crop of the full model, exactly (the draws do not depend on the box), for every
    layer type; a thick layer keeps the flux of the ANALYTIC model within 3%."""
    st = GalmodSetup(small_header, GalmodPar(LTYPE=ltype), densflux=True)
    r = small_rings(z0=10.)
    full, mfull, _ = _mc(st, r, (0, 0), (64, 64))
    box, _, _ = _mc(st, r, (5, 8), (60, 58))
    assert np.array_equal(box, full[8:58, 5:60])
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    an = oracle.galmod_analytic(st, sr, (0, 0), (64, 64), nzq_for(sr), PHI_TAB)
    assert abs(mfull.sum() / an.sum() - 1) < 0.03


def test_mc_flux_per_subcloud(small_header):
    """Every sub-cloud of a sub-ring carries fluxsc = dens x area / (nc x nv) (galmod.py:103-117): with unit
    weights and a disk inside the cube, the total flux is sum dens x ring area x cd2i, up to the sub-clouds
    whose channel falls outside the band."""
    st = GalmodSetup(small_header, GalmodPar(), densflux=True)
    r = small_rings(vrot=50., vdisp=5.)
    counts, mc, q = _mc(st, r, (0, 0), (64, 64))
    sr, _ = st.subrings(model_rings(r, "CENTRED", st.radsep))
    expect = np.sum(sr[5] * 2 * np.pi * sr[0] * st.radsep) * st.cd2i[0]
    assert abs(mc.sum() / expect - 1) < 1e-4


def test_mc_log_coefficients_from_their_fit():
    """The coefficients of mc_log in the kernel and in the oracle are the float32-rounded least-squares fit that
    bbarolo2_amd/csrc/mc_log_fit.py states, and the evaluation stays within 2e-6 of log."""
    import importlib.util
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("mc_log_fit", os.path.join(root, "bbarolo2_amd", "csrc",
                                                                             "mc_log_fit.py"))
    fit = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fit)
    c = fit.coefficients()
    want = [np.float32(c[k]) for k in range(9, 0, -1)]
    for path, fn in ((("bbarolo2_amd", "csrc", "bb_mc.hip"), "mc_log"), (("oracle", "mc.c"), "log_f")):
        src = open(os.path.join(root, *path)).read()
        body = src[src.index(f"float {fn}(float x)"):]
        body = body[:body.index("return")]
        got = [np.float32(float(v)) for v in re.findall(r"p = (?:fmaf\(p, y, )?(-?[0-9.]+e[-+][0-9]+)f", body)]
        assert got == want, (path, got)
    assert c[0] == np.float32(1.0)
    x = np.exp(np.linspace(-20., 20., 200001)).astype(np.float32)
    err = np.abs(fit.evaluate(x, c).astype(np.float64) - np.log(x.astype(np.float64)))
    assert err.max() < 2e-6
    o = np.array([oracle.lib().orc_mc_log(float(v)) for v in x[::997]], np.float64)
    assert np.abs(o - np.log(x[::997].astype(np.float64))).max() < 2e-6
