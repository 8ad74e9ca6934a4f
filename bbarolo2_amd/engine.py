"""DiskModel on the GPU: the reference's model engine API (bbarolo/galmod.py:1147-1377) over libbbgpu.

DiskModel(header, par, ringmode, densflux, galmodtype, smooth) computes the header-derived setup and the
beam kernel once, copies them to the device (bb_model_create), and builds models of rings:
  array(rings, blo, bhi, final, layout)   float64 numpy array (nz, ny, nx) or (ny, nx, nz), as the reference;
  array_batch(tables, box)                a (B, by, bx, nz) float32 device tensor (yxz) of B sub-ring tables
                                          (ANALYTIC); array_batch_mc the same with the Monte-Carlo model;
  ring_maps(rings, box)                   the unit-density total-intensity maps of the rings.
Device memory and streams come from torch; the arithmetic runs in the HIP kernels. There is no CPU path:
without a GPU or without libbbgpu.so these methods raise.
"""

import ctypes as C
import math
from collections import OrderedDict

import numpy as np

from . import _lib
from .galmod import (GalmodSetup, GalmodPar, model_rings, model_box, ringmode_name, zstep_for, nzq_for,
                     ZSAMPLES_MAX, seed32)
from .smooth import beam_kernel

AN_THREADS = 256
PLAN_CACHE = 32          # azimuth plans kept per DiskModel


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.BBError("no GPU visible: the model-cube likelihood runs only on the GPU")
    return torch


def _stream_ptr(torch, stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def naz_of(R0, pixarea):
    """Azimuthal samples of an ANALYTIC sub-ring (galmod.py:341-343)."""
    circ = 2 * math.pi * np.asarray(R0, float) / math.sqrt(pixarea)
    return np.maximum((3 * circ + 0.5).astype(np.int64), 16)


class DiskModel:
    """Models of a rotating disk on the grid of a cube, built on the GPU (see the module docstring)."""

    def __init__(self, header, par=None, ringmode=None, densflux=None, galmodtype=None, smooth=None, device=None):
        self.header = header
        self.par = GalmodPar() if par is None else par
        self.setup = GalmodSetup(header, self.par, galmodtype=galmodtype,
                                 densflux=self.par.DENSFLUX if densflux is None else densflux)
        self.ringmode = ringmode_name(self.par.RINGMODE if ringmode is None else ringmode)
        self.smooth = bool(self.par.SM if smooth is None else smooth)
        self.cfield, self.scalefactor = None, 1.
        if self.smooth:
            kb = beam_kernel(header) if header.Bmaj() > 0 else None
            if kb is None:
                self.smooth = False
            else:
                self.cfield, sc = kb
                self.scalefactor = 1.0 if self.setup.densIsFlux else sc
        self.setup.zstep = zstep_for(header, self.setup, self.cfield is not None)
        self._handle = None
        self._plans = None
        self._device = device
        self._bin = None
        self._deposit = None

    def set_bin(self, per_chunk):
        """The capacity of the deposits' per-tile chunk lists (bb_model_set_bin): 0..8 entries per chunk, 0 for
        none (every tile tests every chunk box). The models do not depend on it."""
        self._bin = int(per_chunk)
        for h in (self._handle or {}).values():
            _lib.check(_lib.lib().bb_model_set_bin(h, self._bin))

    def set_deposit(self, variant):
        """The ANALYTIC deposit kernel (bb_model_set_deposit): 0 one lane per sample (default), 1 the wavefront
        deposit. The models are bit-identical."""
        self._deposit = int(variant)
        for h in (self._handle or {}).values():
            _lib.check(_lib.lib().bb_model_set_deposit(h, self._deposit))

    # ------------------------------------------------------------------ device objects
    def grid(self, nzq):
        st = self.setup
        g = _lib.bb_grid()
        g.nz, g.ltype, g.axtyp, g.velsys = st.nsubs, st.ltype, st.axtyp, st.velsys
        g.nlines, g.nzq, g.cmode = len(st.lfreq0), int(nzq), st.cmode
        if g.nlines > _lib.MAX_LINES:
            raise _lib.BBError(f"at most {_lib.MAX_LINES} lines")
        g.radsep, g.cdens, g.pixarea = st.radsep, st.cdens, st.pixarea
        g.cdelt0, g.cdelt1, g.crota2 = st.cdelt[0], st.cdelt[1], st.crota2
        g.sig_instr, g.crval3, g.cdelt3, g.crpix3 = st.sig_instr, st.crval3, st.cdelt3, st.crpix3
        g.chmin, g.zstep = st.chmin, st.zstep
        for i in range(g.nlines):
            g.relint[i] = st.relint[i]
            g.lfreq0[i] = st.lfreq0[i]
        return g

    def handle(self, nzq=ZSAMPLES_MAX):
        """The bb_model of this DiskModel on the current device (created on first use; nzq fixes the
        vertical-sample capacity: the reference uses 1 for thin disks and 1024 otherwise)."""
        torch = _torch()
        key = (torch.cuda.current_device(), int(nzq))
        if self._handle is None:
            self._handle = {}
        h = self._handle.get(key)
        if h is None:
            L = _lib.lib()
            st = self.setup
            g = self.grid(nzq)
            cd = np.ascontiguousarray(st.cd2i, np.float64)
            ve = np.ascontiguousarray(st.vedges_lines, np.float64)
            if self.cfield is not None:
                k = np.ascontiguousarray(self.cfield, np.float64)
                kp, ky, kx = k.ctypes.data, k.shape[0], k.shape[1]
            else:
                k, kp, ky, kx = None, None, 0, 0
            out = C.c_void_p()
            _lib.check(L.bb_model_create(C.byref(g), cd.ctypes.data, ve.ctypes.data, kp, ky, kx,
                                         float(self.scalefactor), C.byref(out)))
            h = out
            if self._bin is not None:
                _lib.check(L.bb_model_set_bin(h, int(self._bin)))
            if self._deposit is not None:
                _lib.check(L.bb_model_set_deposit(h, int(self._deposit)))
            self._handle[key] = h
        return h

    def azplan(self, R0, nzq=ZSAMPLES_MAX):
        """The bb_azplan of the sub-ring radii R0 (arcmin) on the current device, cached by radii: the PLAN_CACHE most
        recently used plans are kept, the least recently used one is destroyed when a new one would exceed them. The
        plans serve eager launches only (the likelihood context and its captured graphs own plans of their own); a
        destroyed plan's device memory is freed after the launches queued before it (hipFree waits for them)."""
        torch = _torch()
        R0 = np.ascontiguousarray(R0, np.float64)
        key = (torch.cuda.current_device(), int(nzq), R0.tobytes())
        if self._plans is None:
            self._plans = OrderedDict()
        p = self._plans.get(key)
        if p is not None:
            self._plans.move_to_end(key)
            return p
        h = self.handle(nzq)
        out = C.c_void_p()
        _lib.check(_lib.lib().bb_azplan_create(h, R0.ctypes.data, len(R0), C.byref(out)))
        while len(self._plans) >= PLAN_CACHE:
            _, old = self._plans.popitem(last=False)
            _lib.lib().bb_azplan_destroy(old)
        self._plans[key] = out
        return out

    def __del__(self):
        try:
            L = _lib.lib()
            for p in (self._plans or {}).values():
                L.bb_azplan_destroy(p)
            for h in (self._handle or {}).values():
                L.bb_model_destroy(h)
        except Exception:
            pass

    # ------------------------------------------------------------------ reference API
    def box(self, rings, final=False):
        return model_box(self.header, rings, self.ringmode, final)

    def subrings(self, rings, final=True, ringmode=None):
        if final:
            rings = model_rings(rings, self.ringmode if ringmode is None else ringmode, self.setup.radsep)
        return self.setup.subrings(rings)

    def array(self, rings, blo=None, bhi=None, final=True, ringmode=None, galmodtype=None, layout="zyx", seed=None):
        """The model of the rings (convolved with the beam if smoothing is on) in the box [blo, bhi), as a
        float64 array (nz, ny, nx), or (ny, nx, nz) with layout='yxz' (DiskModel.array, galmod.py:1242-1271).
        final=False (models of single rings in the 3DFIT ring fits) uses the rings as nodes. seed: of the MC
        model (default ISEED)."""
        gtype = (galmodtype or self.setup.galmodtype).upper()
        if blo is None or bhi is None:
            blo, bhi = (0, 0), (self.setup.dimx, self.setup.dimy)
        sr, _ = self.subrings(rings, final, ringmode)
        if gtype == "ANALYTIC":
            m = self.array_batch(sr[None], blo, bhi, convolve=self.cfield is not None)
        else:
            m = self.array_batch_mc(sr[None], blo, bhi, convolve=self.cfield is not None, seed=seed)
        a = m[0].double().cpu().numpy()
        return np.ascontiguousarray(a.transpose(2, 0, 1)) if layout == "zyx" else a

    def ring_maps(self, rings, blo=None, bhi=None, ringmode=None):
        """Total-intensity maps (nr, by, bx) float64 of the rings with unit density, beam-convolved when smoothing
        is on (DiskModel.ring_maps, galmod.py:1317-1345): sum_r dens_r maps[r] is the total-intensity map of the
        model of the rings. Computed on the device with the ANALYTIC samples."""
        from .profile import weights_of, QBITS
        torch = _torch()
        if blo is None or bhi is None:
            blo, bhi = (0, 0), (self.setup.dimx, self.setup.dimy)
        st = self.setup
        mode = self.ringmode if ringmode is None else ringmode_name(ringmode)
        mr = model_rings(rings, mode, st.radsep).copy()
        mr.dens = np.ones(mr.nr)
        sr = st.subrings(mr)[0]
        cols = []
        for r in range(rings.nr):
            unit = rings.copy()
            unit.dens = np.eye(rings.nr)[r]
            cols.append(st.subrings(model_rings(unit, mode, st.radsep))[0][5])
        widx, wval = weights_of(np.stack(cols, axis=1))
        by, bx = bhi[1] - blo[1], bhi[0] - blo[0]
        dev = torch.device("cuda", torch.cuda.current_device())
        counts = torch.empty((rings.nr, by, bx), dtype=torch.int64, device=dev)
        maps = torch.empty((rings.nr, by, bx), dtype=torch.float64, device=dev)
        nzq = nzq_for(sr)
        h = self.handle(nzq)
        L = _lib.lib()
        stp = _stream_ptr(torch)
        nsub = sr.shape[1]
        flux = [2 * math.pi * R * st.radsep / st.pixarea * st.pixarea / (naz_of(R, st.pixarea) * 2) for R in sr[0]]
        qunit = max(max(flux, default=1.), 1e-300) * 2. ** -QBITS
        p = self.azplan(sr[0], nzq)
        wb = int(L.bb_ring_maps_work_size(p, 1))
        ws = torch.empty(wb, dtype=torch.uint8, device=dev)
        d_sr = torch.from_numpy(np.ascontiguousarray(sr)).to(dev)
        d_wi = torch.from_numpy(widx).to(dev)
        d_wv = torch.from_numpy(wval).to(dev)
        box = _lib.bb_box(blo[0], blo[1], bhi[0], bhi[1])
        _lib.check(L.bb_ring_maps_deposit(h, p, box, d_sr.data_ptr(), nsub, 1, d_wi.data_ptr(), d_wv.data_ptr(),
                                          rings.nr, 0, rings.nr, qunit, ws.data_ptr(), wb, counts.data_ptr(), stp))
        unit = qunit * float(np.mean(st.cd2i)) * float(np.sum(st.relint))
        _lib.check(L.bb_ring_maps_finish(h, counts.data_ptr(), rings.nr, by, bx, unit, int(self.cfield is not None),
                                         maps.data_ptr(), stp))
        return maps.cpu().numpy()

    def mc_clouds(self, sr):
        """The clouds of each table (B, 16, nsub) of the MC model (galmod.py:103-108): (B,) int64 upper bounds
        (the clouds inside any box are at most these)."""
        st = self.setup
        sr = np.asarray(sr, np.float64)
        R0, dens = sr[:, 0], sr[:, 5]
        area = 2 * math.pi * R0 * st.radsep / st.pixarea
        dc = dens if st.cmode == 1 else dens * dens if st.cmode == 2 else dens ** st.cmode
        nc = np.floor(st.cdens * dc * area + 0.5)
        nc = np.where(np.floor(nc / np.maximum(area, 1e-300) + 0.5) < 10., np.floor(10 * area + 0.5), nc)
        nc = np.where(dens != 0, nc, 0)
        return (nc.sum(axis=1) + 2 * sr.shape[2]).astype(np.int64)

    def array_batch_mc(self, sr, blo, bhi, convolve=True, out=None, work=None, stream=None, d_sr=None, seed=None,
                       counts=None, max_clouds=None, vbox=None):
        """Monte-Carlo models (counter-based stream, bbgpu.h) of B sub-ring tables sr (B, 16, nsub) in the box
        [blo, bhi): a (B, by, bx, nz) float32 device tensor (yxz), convolved with the beam if convolve. counts:
        optional (B, by, bx, nz) int64 device tensor of the integer sums. max_clouds: clouds per vector to hold
        (default: the bound of mc_clouds on sr). vbox (B, 4) int: the model box [blo0, blo1, bhi0, bhi1) of each
        vector within [blo, bhi) (bb_galmod_mc_boxes; None: [blo, bhi) for all)."""
        torch = _torch()
        sr = np.ascontiguousarray(sr, np.float64)
        B, nrow, nsub = sr.shape
        if nrow != 16:
            raise ValueError("sub-ring tables must have 16 rows")
        if not np.all(np.isfinite(sr)):
            raise ValueError("sub-ring tables hold NaN or infinite values")
        st = self.setup
        by, bx, nz = bhi[1] - blo[1], bhi[0] - blo[0], st.nsubs
        dev = torch.device("cuda", torch.cuda.current_device())
        h = self.handle(nzq_for(sr))
        sp = _stream_ptr(torch, stream)
        L = _lib.lib()
        box = _lib.bb_box(blo[0], blo[1], bhi[0], bhi[1])
        cap = int(self.mc_clouds(sr).max()) if max_clouds is None else int(max_clouds)
        wb = int(L.bb_mc_work_size(h, box, nsub, B, cap))
        if wb < 0:
            raise _lib.BBError("bb_mc_work_size: bad arguments")
        ws = torch.empty(wb, dtype=torch.uint8, device=dev)
        flag = torch.zeros(1, dtype=torch.int32, device=dev)
        dep = torch.empty((B, by, bx, nz), dtype=torch.float32, device=dev) if work is None else work
        if d_sr is None:
            d_sr = torch.from_numpy(sr).to(dev)
        nv = st.nvpar if st.nvpar >= 1 else st.nsubs
        s32 = seed32(st.iseed if seed is None else seed)
        if vbox is None:
            _lib.check(L.bb_galmod_mc(h, box, d_sr.data_ptr(), nsub, B, nv, s32, cap, ws.data_ptr(), wb, dep.data_ptr(),
                                      None if counts is None else counts.data_ptr(), flag.data_ptr(), 0, sp))
        else:
            d_vb = torch.from_numpy(np.ascontiguousarray(vbox, np.int32).reshape(B, 4)).to(dev)
            _lib.check(L.bb_galmod_mc_boxes(h, box, d_vb.data_ptr(), d_sr.data_ptr(), nsub, B, nv, s32, cap,
                                            ws.data_ptr(), wb, dep.data_ptr(),
                                            None if counts is None else counts.data_ptr(), flag.data_ptr(), 0, sp))
        if int(flag.item()) != 0:
            raise _lib.BBError(f"bb_galmod_mc: more than {cap} clouds of a vector inside the box")
        if not convolve or self.cfield is None:
            return dep
        res = torch.empty_like(dep) if out is None else out
        _lib.check(L.bb_conv2d_same(h, dep.data_ptr(), res.data_ptr(), B, by, bx, sp))
        return res

    def orient_tile(self, sr, nzq=None):
        """Orient the rectangular pixel tile of the deposits (bb_model_set_tile) along the vertical columns of the
        tables sr (B, 16, nsub): a sample's column runs along the projected minor axis, (sin i cos p / cdelt0,
        -sin i sin p / cdelt1) in pixels (p = PA + CROTA2), and a tile that is long in that direction is crossed by
        fewer columns, so fewer samples are gathered by two tiles. Returns the (TX, TY) set."""
        sr = np.asarray(sr, np.float64).reshape(-1, 16, np.shape(sr)[-1])
        st = self.setup
        L = _lib.lib()
        h = self.handle(nzq_for(sr) if nzq is None else nzq)
        tx, ty = C.c_int(), C.c_int()
        _lib.check(L.bb_model_get_tile(h, C.byref(tx), C.byref(ty)))
        lo, hi = sorted((tx.value, ty.value))
        if lo == hi or sr.shape[2] == 0:
            return tx.value, ty.value
        sinc = np.abs(np.sin(sr[:, 9]))
        p = sr[:, 10] + st.crota2
        w = np.abs(sr[:, 5]) + 1e-300
        dx = float(np.sum(w * sinc * np.abs(np.cos(p))) / abs(st.cdelt[0]))
        dy = float(np.sum(w * sinc * np.abs(np.sin(p))) / abs(st.cdelt[1]))
        want = (lo, hi) if dy > dx else (hi, lo)
        if want != (tx.value, ty.value):
            _lib.check(L.bb_model_set_tile(h, want[0], want[1]))
        return want

    def array_batch(self, sr, blo, bhi, convolve=True, out=None, work=None, stream=None, d_sr=None, vbox=None):
        """ANALYTIC models of B sub-ring tables sr (B, 16, nsub) sharing their radii, in the box [blo, bhi):
        a (B, by, bx, nz) float32 device tensor (yxz), convolved with the beam if convolve. d_sr: the tables
        already on the device (sr then gives their radii and heights: the density row may differ). vbox (B, 4) int:
        the model box [blo0, blo1, bhi0, bhi1) of each vector within [blo, bhi) (bb_galmod_analytic_boxes)."""
        torch = _torch()
        sr = np.ascontiguousarray(sr, np.float64)
        B, nrow, nsub = sr.shape
        if nrow != 16:
            raise ValueError("sub-ring tables must have 16 rows")
        if not np.all(np.isfinite(sr)):
            raise ValueError("sub-ring tables hold NaN or infinite values")
        by, bx, nz = bhi[1] - blo[1], bhi[0] - blo[0], self.setup.nsubs
        dev = torch.device("cuda", torch.cuda.current_device())
        nzq = nzq_for(sr)
        h = self.handle(nzq)
        st = _stream_ptr(torch, stream)
        dep = torch.empty((B, by, bx, nz), dtype=torch.float32, device=dev) if work is None else work
        if nsub:
            if d_sr is None:
                d_sr = torch.from_numpy(sr).to(dev, non_blocking=False)
            self.orient_tile(sr, nzq)
            p = self.azplan(sr[0, 0], nzq)
            L = _lib.lib()
            wb = int(L.bb_analytic_work_size(p, B))
            ws = torch.empty(wb, dtype=torch.uint8, device=dev)
            box = _lib.bb_box(blo[0], blo[1], bhi[0], bhi[1])
            if vbox is None:
                _lib.check(L.bb_galmod_analytic(h, p, box, d_sr.data_ptr(), nsub, B, ws.data_ptr(), wb,
                                                dep.data_ptr(), 0, st))
            else:
                d_vb = torch.from_numpy(np.ascontiguousarray(vbox, np.int32).reshape(B, 4)).to(dev)
                _lib.check(L.bb_galmod_analytic_boxes(h, p, box, d_vb.data_ptr(), d_sr.data_ptr(), nsub, B,
                                                      ws.data_ptr(), wb, dep.data_ptr(), 0, st))
        else:
            dep.zero_()
        if not convolve or self.cfield is None:
            return dep
        return self.convolve_batch(dep, out=out, stream=stream, nzq=nzq)

    def convolve_batch(self, dep, out=None, stream=None, nzq=None):
        """The beam convolution of the device models dep (B, by, bx, nz) (galmod.py:1264-1283, final=True):
        a new tensor, or out."""
        torch = _torch()
        if self.cfield is None:
            return dep
        B, by, bx, _ = dep.shape
        h = self.handle(ZSAMPLES_MAX if nzq is None else nzq)
        res = torch.empty_like(dep) if out is None else out
        _lib.check(_lib.lib().bb_conv2d_same(h, dep.data_ptr(), res.data_ptr(), B, by, bx, _stream_ptr(torch, stream)))
        return res
