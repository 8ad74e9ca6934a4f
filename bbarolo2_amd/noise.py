"""Per-fit set-up of the likelihood on the GPU: the likelihood mask.

dilate_mask follows the reference's (bbarolo/bayesian/noise.py:18-30, used by _setup_likemask,
bbarolo/bayesian/likelihood.py:160-175): the emission mask dilated by a disk of the given radius (pixels) in each
channel, then by +-nchan channels. It runs on the device (bb_dilate_mask: the same set of voxels, formed directly
from the mask pixels within the radius instead of a distance transform).
"""

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr


def dilate_mask_device(mask, radius, nchan, out=None):
    """The dilated mask of a device tensor mask (nz, ny, nx) (bool or uint8) as a uint8 device tensor (0 / 1)."""
    torch = _torch()
    m = mask.to(torch.uint8).contiguous() if mask.dtype != torch.uint8 or not mask.is_contiguous() else mask
    nz, ny, nx = m.shape
    if out is None:
        out = torch.empty_like(m)
    radius, nchan = float(radius), int(nchan)
    tmp = torch.empty_like(m) if radius >= 1 and nchan > 0 else None
    _lib.check(_lib.lib().bb_dilate_mask(m.data_ptr(), nz, ny, nx, radius, nchan,
                                         None if tmp is None else tmp.data_ptr(), out.data_ptr(),
                                         _stream_ptr(torch)))
    return out


def dilate_mask(mask, radius, nchan):
    """A disk of the given radius (pixels) in each channel, then +-nchan channels: the numpy bool (nz, ny, nx)
    dilation of the numpy mask, computed on the current GPU."""
    torch = _torch()
    m = np.ascontiguousarray(np.asarray(mask, bool)).view(np.uint8)
    dev = torch.device("cuda", torch.cuda.current_device())
    out = dilate_mask_device(torch.from_numpy(m).to(dev), radius, nchan)
    return out.cpu().numpy().astype(bool)
