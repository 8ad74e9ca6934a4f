import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


def _have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture
def small_header():
    """The reference's synthetic_cube fixture (tests/test_bbarolo.py:36-43): 64x64x40, 10" pixels,
    10 km/s channels, 30" beam."""
    from bbarolo2_amd.header import emptyheader
    return emptyheader([64, 64, 40], [-10. / 3600., 10. / 3600., 10.], beam=30. / 3600., obj="SYNTH")


def small_rings(nr=8, **kw):
    """The reference's _rings() (tests/test_bbarolo.py:46-48)."""
    from bbarolo2_amd.rings import Rings
    d = dict(radii=np.arange(nr) * 20. + 10., xpos=31.5, ypos=31.5, vsys=0., vrot=150., vdisp=10., z0=0., dens=1.,
             inc=60., phi=45.)
    d.update(kw)
    return Rings().setRings(**d)
