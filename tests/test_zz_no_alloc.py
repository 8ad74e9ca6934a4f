"""GPU (MI355X): bb_conv_residual_gauss_noalloc (the fused convolution-residual with its partial sums in the library's
static device array) gives the results of bb_conv_residual_gauss (a caller's work buffer) bit for bit, on two calls in
a row, the second with a smaller batch and box; with and without the fractional model error; and a call that needs more
than the static array holds is rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(dm, B, by, bx, seed):
    import torch
    nz = dm.setup.nsubs
    g = torch.Generator(device="cuda").manual_seed(seed)
    dep = torch.rand((B, by, bx, nz), generator=g, device="cuda", dtype=torch.float32)
    dep[dep < 0.3] = 0.
    dv = torch.rand((by, bx, nz), generator=g, device="cuda", dtype=torch.float32)
    dv[dv < 0.1] = float("nan")
    return dep, dv


def _both(dm, dep, dv, eps2):
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.engine import _stream_ptr
    L = _lib.lib()
    B, by, bx, _ = dep.shape
    h = dm.handle(1)
    wl = int(L.bb_conv_residual_work_size(h, B, by, bx))
    work = torch.empty(wl, dtype=torch.float64, device="cuda")
    out = [torch.empty(B, dtype=torch.float64, device="cuda") for _ in range(4)]
    sp = _stream_ptr(torch)
    _lib.check(L.bb_conv_residual_gauss(h, dep.data_ptr(), dv.data_ptr(), B, by, bx, 0.25, eps2, work.data_ptr(), wl,
                                        out[0].data_ptr(), out[1].data_ptr(), sp))
    _lib.check(L.bb_conv_residual_gauss_noalloc(h, dep.data_ptr(), dv.data_ptr(), B, by, bx, 0.25, eps2,
                                                out[2].data_ptr(), out[3].data_ptr(), sp))
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("eps2", [0.0, 0.04])
def test_noalloc_matches_work_buffer(eps2):
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.header import emptyheader
    torch.cuda.set_device(0)
    h = emptyheader([96, 96, 40], [-10. / 3600., 10. / 3600., 10.], beam=(40 / 3600., 30 / 3600., 30.))
    dm = DiskModel(h)
    for call, (B, by, bx) in enumerate(((24, 80, 72), (5, 41, 37))):
        dep, dv = _inputs(dm, B, by, bx, 100 + call)
        ss, lg, ss_na, lg_na = _both(dm, dep, dv, eps2)
        print(f"[noalloc call {call}] B {B} box {by}x{bx} eps2 {eps2}: ss {ss[:2]} max |d| "
              f"{np.abs(ss - ss_na).max():.3e}", flush=True)
        assert np.all(np.isfinite(ss)) and np.all(ss > 0)
        assert np.array_equal(ss, ss_na) and (eps2 == 0 or np.array_equal(lg, lg_na))


def test_noalloc_rejects_beyond_its_array():
    import torch
    from bbarolo2_amd import _lib
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.header import emptyheader
    torch.cuda.set_device(0)
    L = _lib.lib()
    h = emptyheader([256, 256, 64], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600.)
    dm = DiskModel(h)
    hd = dm.handle(1)
    B = 4096
    need = int(L.bb_conv_residual_work_size(hd, B, 256, 256))
    assert need > int(L.bb_conv_residual_noalloc_max())
    v = C.c_void_p(256)             # (never dereferenced: the size check comes first)
    rc = L.bb_conv_residual_gauss_noalloc(hd, v, v, B, 256, 256, 1.0, 0.0, v, None, None)
    assert rc != 0 and b"scratch" in L.bb_last_error()
