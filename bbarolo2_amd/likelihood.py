"""The Gaussian log-likelihood of Bayesian3DFIT (GLOBAL / BAYESIAN), evaluated for batches of parameter vectors on the
GPU by the library's likelihood context (bbgpu.h bb_ctx_create / bb_loglike_batch).

Mirrors LikelihoodMixin._log_likelihood (bbarolo/bayesian/likelihood.py:567-619) for the default Gaussian residuals
(_calculate_residuals / _residuals_fracerr, likelihood.py:457-531):

    -logL = likefactor * [ 0.5 (S_box + ss_out) / sigma^2 + 0.5 nvox ln(2 pi sigma^2) ]      (+ fracerr terms)

with S_box the masked residual sum inside the vector's model box, ss_out the squared data of the likelihood voxels of
the region outside that box (a summed-area table, likelihood.py:437-440) and nvox the number of likelihood voxels of
the region. theta -> rings follows _update_rings (bayesian/fitter.py:1109-1136): the free values are stored as
float32 in the ring table (bayesian/rings.py:54-61), then used as float64; the free edges of the disk (RADMAX,
RADMAXA, RADMAXR) are used as float64 and, with the profile, move each vector's exterior annulus.

This module does the per-fit set-up that the reference does in Python (the sub-ring plan of the fixed radii, the
constants of the deconvolved profile, the model box) and hands it to the context once; every batch then runs in the
library from theta (host) to log L (host): the sub-ring tables are built on the device, each vector gets its own model
box when XPOS, YPOS or Z0 is free, and the profile, deposit, convolution and residual follow without a host round
trip. The set-up that the reference does once per fit from the data (likelihood mask, region, noise) is taken as
given: the caller passes the likelihood mask, the region and sigma (see synth.likelihood_inputs).
"""

import ctypes as C
import math

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr
from .galmod import SubringPlan, LIKE_KEYS, ring_table, outer_edge, seed32, ZSAMPLES_MAX
from .header import arcsconv
from .profile import ProfileSetup, SMOOTH_DEFAULT, ANNULUS_KEYS

# free parameters that change the deconvolved density profile (fitter.py:614-615)
PROFILE_KEYS = ('inc', 'phi', 'xpos', 'ypos', 'radmax', 'z0', 'fthick')

# ring parameters that make the likelihood zero when negative (likelihood.py:532-535)
POSITIVE = ('vrot', 'vdisp', 'inc', 'xpos', 'ypos', 'z0', 'dens', 'zcyl', 'z0thick', 'fthick', 'vdispthick')

# the most sub-rings of one layer of a vector's exterior annulus (bb_like.hip BB_ANN_CAP)
ANN_CAP = 2048

# free parameters that move the model box (DiskModel.box: the centre of the last ring and the layer height)
BOX_KEYS = ('xpos', 'ypos', 'z0', 'z0thick')

# free parameters that are not ring values: SIGMA, FRACERR and the edges of the disk (bayesian/fitter.py:1112-1120)
GLOBAL_KEYS = ('sigma', 'fracerr', 'radmax', 'radmaxa', 'radmaxr')
EDGE_KEYS = ('radmax', 'radmaxa', 'radmaxr')


class FreePars:
    """The free parameters of a fit: an ordered list of (key, 'ring' | 'single'); 'ring' takes nr values of theta
    (one per ring), 'single' one value for all the rings (the FREE keys VROT and INC_SINGLE of the reference,
    bayesian/fitter.py:434-473). SIGMA, FRACERR, RADMAX, RADMAXA and RADMAXR are single values that are not ring
    values."""

    def __init__(self, spec, nr):
        self.spec = [(k.lower(), m) for k, m in spec]
        self.nr = nr
        self.index = {}
        i = 0
        for k, m in self.spec:
            if k in GLOBAL_KEYS and m == "ring":
                raise ValueError(f"{k.upper()} is one value, not one per ring")
            n = nr if m == "ring" else 1
            self.index[k] = np.arange(i, i + n)
            i += n
        self.ndim = i

    def values(self, theta):
        """dict key -> (B, nr) float64 ring values of the free ring keys, rounded to float32 as the ring table of the
        reference stores them."""
        theta = np.atleast_2d(np.asarray(theta, float))
        out = {}
        for k, m in self.spec:
            if k in GLOBAL_KEYS:
                continue
            v = theta[:, self.index[k]]
            if m != "ring":
                v = np.repeat(v, self.nr, axis=1)
            out[k] = v.astype(np.float32).astype(np.float64)
        return out

    def negative(self, theta):
        """True where the likelihood is zero (likelihood.py:537-559): a free positive-only value below 0, a free FTHICK
        above 1, SIGMA not above 0, RADMAX below 0."""
        theta = np.atleast_2d(np.asarray(theta, float))
        bad = np.zeros(theta.shape[0], bool)
        for k, _ in self.spec:
            v = theta[:, self.index[k]]
            if k in POSITIVE:
                bad |= np.any(v.astype(np.float32) < 0, axis=1)
            if k == "fthick":
                bad |= np.any(v.astype(np.float32) > 1, axis=1)
            if k == "sigma":
                bad |= ~(v[:, 0] > 0)
            if k == "radmax":
                bad |= v[:, 0] < 0
        return bad

    def offsets(self):
        """(free_off, free_ring) int32 (NKEY,): the offset of each key of the context's ring table in theta (-1: not
        free) and whether it takes one value per ring."""
        off = np.full(_lib.NKEY, -1, np.int32)
        ring = np.zeros(_lib.NKEY, np.int32)
        for k, m in self.spec:
            j = LIKE_KEYS.index(k)
            off[j] = self.index[k][0]
            ring[j] = 1 if m == "ring" else 0
        return off, ring


def float32_rings(rings):
    """The rings as the ring table of Bayesian3DFIT holds them: every value rounded to float32, radsep
    float32(R[-1] - R[-2]) (bayesian/rings.py:54-76)."""
    r = rings.copy()
    for k, v in r.as_dict().items():
        setattr(r, k, np.asarray(v, np.float32).astype(np.float64))
    if r.nr > 1:
        r.radsep = float(np.float32(r.radii[-1] - r.radii[-2]))
    return r


def _plan_struct(plan, ridx, from_fit, keep):
    """bb_ring_plan of a SubringPlan (arrays kept alive in keep)."""
    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data
    densw, side = plan.weights()
    p = _lib.bb_ring_plan()
    p.nmr, p.nsub, p.thick, p.from_fit = len(plan.rad), plan.nsub, int(plan.thick), from_fit
    p.ridx = arr(ridx, np.int32)
    p.rad = arr(plan.rad, np.float64)
    p.base = arr(ring_table(plan.proto), np.float64)
    p.idx = arr(plan.idx, np.int32)
    p.dr = arr(plan.dr, np.float64)
    p.cst = arr(plan.const, np.uint8)
    p.rads = arr(plan.rads, np.float64)
    p.densw = None if densw is None else arr(densw, np.float64)
    p.side = arr(side, np.float64)
    keep.append(p)
    return p


class GpuLikelihood:
    """Batched log-likelihood of a cube for the free parameters `free` around the rings `rings0`.

    dm: engine.DiskModel of the cube (ANALYTIC or MC). data (nz, ny, nx) float32. likemask (nz, ny, nx) bool (maskmode
    2: only those voxels; None: all). region (ny, nx) bool (None: the likelihood box). likebox ((blo), (bhi)): the
    likelihood region's box. sigma: noise per voxel. likefactor: 1 / correlation volume. fracerr: fixed fractional
    model error (None: 0). chunk: the most vectors of one launch. halves: the two halves of a launch on two streams.
    radmax, radmax_width: what a free RADMAX is (like_kwargs radmax, bayesian/fitter.py:556-569): 'edge' the outer edge
    of the disk, sharp; 'taper' a Gaussian edge of width radmax_width arcsec (default a quarter of the beam FWHM)."""

    def __init__(self, dm, data, rings0, free, likemask=None, region=None, likebox=None, sigma=1.0,
                 likefactor=1.0, fracerr=None, maskmode=2, chunk=256, densprof_smooth=SMOOTH_DEFAULT,
                 likefactor2d=None, beam_arcsec=None, graphs=False, halves=True, radmax="edge", radmax_width=None):
        self.torch = _torch()
        self.dm = dm
        self.free = free
        st = dm.setup
        for k in free.index:
            if k not in LIKE_KEYS:
                raise NotImplementedError(f"free {k.upper()} is not supported by this engine")
        r0 = rings0.copy()
        self.edgewidth = 0.
        self.ann_cap = None
        if "radmax" in free.index:
            mode = str(radmax or "edge")
            if mode == "ring":
                raise NotImplementedError("RADMAX as the radius of the last ring (like_kwargs radmax 'ring') moves the "
                                          "radii of the sub-ring plan: use 'edge' or 'taper'")
            if mode not in ("edge", "taper"):
                raise ValueError(f"radmax {mode!r}: 'edge', 'taper' or 'ring'")
            if mode == "taper":
                bm = dm.header.Bmaj()
                w = radmax_width if radmax_width is not None else 0.25 * abs(bm) * 3600.
                self.edgewidth = float(w)
            r0.edge, r0.edgewidth = None, 0.       # (the vector's edge replaces a fixed one)
        if "fthick" in free.index:
            r0.twolayer = True           # (a free FTHICK needs the thick layer, also where it starts at 0)
        self.rings0 = float32_rings(r0)
        self.sigma = float(sigma)
        self.likefactor = float(likefactor)
        self.eps = float(fracerr) if fracerr else 0.
        self.chunk = int(chunk)
        self.use_graphs = bool(graphs)
        nz, ny, nx = data.shape
        if likebox is None:
            likebox = ((0, 0), (nx, ny))
        Lb, Hb = (tuple(int(v) for v in likebox[0]), tuple(int(v) for v in likebox[1]))
        self.likebox = (Lb, Hb)
        if region is None:
            region = np.zeros((ny, nx), bool)
            region[Lb[1]:Hb[1], Lb[0]:Hb[0]] = True
        region = np.asarray(region, bool)
        if maskmode == 2:
            lm = None if likemask is None else np.ascontiguousarray(np.asarray(likemask, bool)).view(np.uint8)
        elif maskmode == 0:
            lm = None
        else:
            raise NotImplementedError("maskmode 1 (the data times the mask) is not supported by the likelihood context")
        data = np.ascontiguousarray(data, np.float32)
        nvmap, dm2, ymap = self._maps(data, lm)
        self.nvox = int(nvmap[region].sum())
        dm2 = np.where(region, dm2, 0.)
        sat = np.zeros((ny + 1, nx + 1))
        sat[1:, 1:] = np.cumsum(np.cumsum(dm2, axis=0), axis=1)
        self._sat = sat
        # the model box: that of the initial rings (fixed), or one per vector when a free parameter moves it
        self.box_free = any(k in free.index for k in BOX_KEYS)
        blo, bhi = dm.box(self.rings0, final=True)
        blo = [max(blo[0], Lb[0]), max(blo[1], Lb[1])]
        bhi = [max(min(bhi[0], Hb[0]), blo[0]), max(min(bhi[1], Hb[1]), blo[1])]
        self.fixbox = (blo, bhi)
        self.blo, self.bhi = (list(Lb), list(Hb)) if self.box_free else (blo, bhi)
        self.ss_out = max(float(sat[-1, -1]) - self._boxsum(blo, bhi), 0.)
        self.plan = SubringPlan(st, self.rings0, dm.ringmode, final=True)
        # DENS not fitted: the deconvolved profile of the initial rings, and of every parameter vector when the
        # geometry is free (fitter.py:606-615)
        self.useNorm = "dens" not in free.index
        self.update_prof = self.useNorm and any(k in free.index for k in PROFILE_KEYS)
        if self.update_prof and "fthick" in free.index:
            raise NotImplementedError("a free FTHICK changes the ring weights of the deconvolved profile: fit DENS")
        self.profile = None
        self.dens0 = None
        self.last_dens = None
        if self.useNorm:
            smap = self.sigma * np.sqrt(nvmap.astype(np.float64))
            sl = (slice(Lb[1], Hb[1]), slice(Lb[0], Hb[0]))
            ok = region[sl] & (nvmap[sl] > 0)
            if beam_arcsec is None:
                bm = dm.header.Bmaj()
                beam_arcsec = abs(bm) * 3600. if bm else None
            l2d = self.likefactor if likefactor2d is None else float(likefactor2d)
            self.profile = ProfileSetup(dm, self.rings0, self.plan, (Lb, Hb), ymap[sl], smap[sl], ok,
                                        smooth=densprof_smooth, corr=1. / l2d, beam_arcsec=beam_arcsec)
        self._data = (data, lm, region)
        if self.useNorm and not self.update_prof:
            # the profile of the initial rings, once (a context without free parameters that computes it)
            ctx0 = self._make_ctx(FreePars([], self.rings0.nr), 1, True, halves=False)
            ll = np.zeros(1)
            dens = np.zeros((1, self.rings0.nr), np.float32)
            th = np.zeros(1)
            try:
                _lib.check(_lib.lib().bb_loglike_batch(ctx0, th.ctypes.data, 1, ll.ctypes.data, dens.ctypes.data,
                                                       _stream_ptr(self.torch)))
            finally:
                _lib.lib().bb_ctx_destroy(ctx0)
            if not np.all(np.isfinite(dens)):
                raise RuntimeError("the density profile of the initial rings did not converge")
            self.dens0 = dens[0].astype(np.float64)
            self.rings0.dens = self.dens0.copy()
        self._ctx = self._make_ctx(free, self.chunk, self.update_prof, halves=halves)
        del self._data
        if self.use_graphs:
            self.set_graphs(True)

    # ------------------------------------------------------------------ set-up
    def _maps(self, data, lm):
        """(nvox_map int, dm2_map, mom0) of the likelihood voxels (likemask lm, uint8 or None: all; not blank) per
        pixel, from one pass over the cube on the device (bb_like_maps: the reference's sums, likelihood.py:176-185)."""
        torch = self.torch
        nz, ny, nx = data.shape
        dev = torch.device("cuda", torch.cuda.current_device())
        d = torch.from_numpy(data).to(dev)
        m = None if lm is None else torch.from_numpy(lm).to(dev)
        dm2 = torch.empty((ny, nx), dtype=torch.float64, device=dev)
        nv = torch.empty((ny, nx), dtype=torch.int32, device=dev)
        m0 = torch.empty((ny, nx), dtype=torch.float64, device=dev)
        _lib.check(_lib.lib().bb_like_maps(d.data_ptr(), None if m is None else m.data_ptr(), None, nz, ny, nx,
                                           _lib.bb_box(0, 0, nx, ny), None, dm2.data_ptr(), nv.data_ptr(),
                                           m0.data_ptr(), _stream_ptr(torch)))
        return nv.cpu().numpy(), dm2.cpu().numpy(), m0.cpu().numpy()

    def _boxsum(self, blo, bhi):
        S = self._sat
        return float(S[bhi[1], bhi[0]] - S[blo[1], bhi[0]] - S[bhi[1], blo[0]] + S[blo[1], blo[0]])

    def _nzq(self, free):
        r = self.rings0
        z = np.any(np.asarray(r.z0) > 0) or (r.thick() and np.any(r.thick_z0() > 0))
        return ZSAMPLES_MAX if (z or "z0" in free.index or "z0thick" in free.index) else 1

    def _make_ctx(self, free, chunk, with_profile, halves=True):
        """A bb_ctx of the fit for the free parameters free (the device buffers of chunk vectors)."""
        dm, st = self.dm, self.dm.setup
        data, lm, region = self._data
        nz, ny, nx = data.shape
        keep = []
        nzq = self._nzq(free)
        h = dm.handle(nzq)
        self.handle = h
        sr0 = self.plan.tables({})[0]
        dm.orient_tile(sr0, nzq)
        nr = self.rings0.nr
        main = _plan_struct(self.plan, self.plan.ring_index, 0xFFFF, keep)
        prof = None
        ann = None
        if with_profile:
            P = self.profile
            bits = 0
            for k in ANNULUS_KEYS:
                bits |= 1 << LIKE_KEYS.index(k)
            ann = _plan_struct(P.plan_ext, np.full(len(P.plan_ext.rad), nr - 1), bits, keep)
            prof = _lib.bb_like_profile()
            prof.d = P.desc()
            for name in ("wpix", "ypix", "lcoef", "lcol", "near", "widx", "wval", "widx_e", "wval_e"):
                a = np.ascontiguousarray(getattr(P, name))
                if name in ("lcol", "near", "widx", "widx_e"):
                    a = a.astype(np.int32)
                keep.append(a)
                if name in ("lcoef", "lcol", "near") and not P.prior():
                    setattr(prof, name, None)
                else:
                    setattr(prof, name, a.ctypes.data)
            prof.qunit, prof.unit = P.qunit, P.unit
            R = np.ascontiguousarray(P.radii[:-1], np.float64)
            keep.append(R)
            prof.radii = R.ctypes.data
            prof.dmin = 0.25 * float(np.diff(R).min()) if R.size > 1 else 0.
        d = _lib.bb_like_desc()
        d.nr, d.ndim, d.chunk = nr, max(free.ndim, 1), int(chunk)
        d.mc = 1 if st.galmodtype == "MC" else 0
        d.nvpar = st.nvpar if st.nvpar >= 1 else st.nsubs
        d.seed = seed32(st.iseed)
        d.box_free = 1 if any(k in free.index for k in BOX_KEYS) else 0
        d.halves = 1 if halves else 0
        (Lb, Hb), (fb, fh) = self.likebox, self.fixbox
        d.likebox = _lib.bb_box(Lb[0], Lb[1], Hb[0], Hb[1])
        d.fixbox = _lib.bb_box(fb[0], fb[1], fh[0], fh[1])
        d.sigma, d.likefactor, d.eps = self.sigma, self.likefactor, self.eps
        r = self.rings0
        ext = r.radii[-1] + r.radsep / 2.
        d.ext = max(ext, outer_edge(r, dm.ringmode))
        # the edges of the disk (the free ones come from theta; the rest are fixed here and in the plans)
        d.edgewidth = self.edgewidth
        d.oedge = outer_edge(r, dm.ringmode)
        d.edge0 = float(r.edge) if r.edge is not None else math.nan
        ea, er = st.side_edges(r)
        d.edgea0 = -1. if ea is None else float(ea)
        d.edger0 = -1. if er is None else float(er)
        d.empty = 1 if st.empty else 0
        if with_profile and any(k in free.index for k in EDGE_KEYS):
            d.annw = float(self.profile.ext.radsep)
            span = d.annw if st.empty else d.oedge + d.annw
            d.ann_cap = int(span / 60. / st.radsep) + 3
            if d.ann_cap > ANN_CAP:
                raise NotImplementedError(f"an exterior annulus of {d.ann_cap} sub-rings (at most {ANN_CAP})")
            self.ann_cap = d.ann_cap
        hd = dm.header
        conv = arcsconv(hd.Cunit(0))
        d.pixas0, d.pixas1 = abs(hd.Cdelt(0)) * conv, abs(hd.Cdelt(1)) * conv
        d.nx, d.ny = nx, ny
        if d.mc:
            d.mc_cap = max(2 * int(dm.mc_clouds(sr0).max()), 4096)
        off, ring = free.offsets()
        rt = np.ascontiguousarray(ring_table(r))
        rgu = np.ascontiguousarray(region, np.uint8)
        keep += [rt, off, ring, lm, rgu, data]
        out = C.c_void_p()
        _lib.check(_lib.lib().bb_ctx_create(h, C.byref(d), data.ctypes.data, None if lm is None else lm.ctypes.data,
                                            rgu.ctypes.data,
                                            rt.ctypes.data, off.ctypes.data, ring.ctypes.data, C.byref(main),
                                            C.byref(ann) if ann is not None else None,
                                            C.byref(prof) if prof is not None else None, C.byref(out)))
        return out

    def __del__(self):
        try:
            if getattr(self, "_ctx", None):
                _lib.lib().bb_ctx_destroy(self._ctx)
                self._ctx = None
        except Exception:
            pass

    # ------------------------------------------------------------------ evaluation
    def _theta(self, theta):
        theta = np.ascontiguousarray(np.atleast_2d(np.asarray(theta, float)))
        if self.free.ndim == 0:
            theta = np.zeros((theta.shape[0], 1))
        elif theta.shape[1] != self.free.ndim:
            raise ValueError(f"theta has {theta.shape[1]} columns, the fit {self.free.ndim}")
        return theta

    @property
    def ctx(self):
        """The bb_ctx of the fit (a C pointer): bb_loglike_batch(ctx, theta, B, loglike, dens, stream) evaluates a
        batch with no Python in the way."""
        return self._ctx

    def loglike(self, theta):
        """log-likelihood of each row of theta (B, ndim) -> (B,) float64 (-inf where the likelihood is zero)."""
        theta = self._theta(theta)
        B = theta.shape[0]
        out = np.empty(B)
        dens = np.empty((B, self.rings0.nr), np.float32) if self.update_prof else None
        _lib.check(_lib.lib().bb_loglike_batch(self._ctx, theta.ctypes.data, B, out.ctypes.data,
                                               None if dens is None else dens.ctypes.data, _stream_ptr(self.torch)))
        if dens is not None:
            self.last_dens = dens.astype(np.float64)
        return out

    def launch(self, theta, stream=None):
        """Launch the evaluation of the rows of theta (at most chunk rows) on the stream (default: the current one) and
        return at once; collect() gives the log-likelihoods."""
        theta = self._theta(theta)
        if theta.shape[0] > self.chunk:
            raise ValueError(f"{theta.shape[0]} vectors for a likelihood of chunk {self.chunk}")
        self._pending = theta.shape[0]
        _lib.check(_lib.lib().bb_loglike_launch(self._ctx, theta.ctypes.data, theta.shape[0],
                                                _stream_ptr(self.torch, stream)))

    def collect(self):
        """The log-likelihoods (B,) of the last launch (waits for it)."""
        B = self._pending
        out = np.empty(B)
        dens = np.empty((B, self.rings0.nr), np.float32) if self.update_prof else None
        _lib.check(_lib.lib().bb_loglike_collect(self._ctx, out.ctypes.data,
                                                 None if dens is None else dens.ctypes.data))
        if dens is not None:
            self.last_dens = dens.astype(np.float64)
        return out

    def device_tables(self, theta):
        """(sr (B, 16, nsub), sr_ext or None): the sub-ring tables of the model (and of the exterior annulus when the
        profile is computed per vector) that the context builds on the device for the rows of theta."""
        theta = self._theta(theta)
        B = theta.shape[0]
        nsub_t = self.plan.nsub * (2 if self.plan.thick else 1)
        sr = np.empty((B, 16, nsub_t))
        ext = None
        if self.update_prof:
            pe = self.profile.plan_ext
            ns = self.ann_cap or pe.nsub
            ext = np.empty((B, 16, ns * (2 if pe.thick else 1)))
        _lib.check(_lib.lib().bb_ctx_tables(self._ctx, theta.ctypes.data, B, sr.ctypes.data,
                                            None if ext is None else ext.ctypes.data))
        return sr, ext

    def tables(self, theta):
        """(sr (B, 16, nsub), values) of the rows of theta on the host (SubringPlan.tables): the sub-ring tables of the
        model (density row 1 when the profile is computed per vector) and the ring values of the free keys."""
        vals = self.free.values(theta)
        if self.useNorm and not self.update_prof:
            vals["dens"] = self.dens0
        sr, _ = self.plan.tables(vals)
        if self.update_prof:
            sr[:, 5, :] = 1.
        return sr, vals

    def info(self):
        """dict: nvox and the squared data of the region (as the context counts them), its deposit box, the vectors
        whose profile did not converge and the MC launches repeated with larger buffers."""
        nv, ss = C.c_double(), C.c_double()
        box = (C.c_int32 * 4)()
        nf, mr = C.c_int64(), C.c_int64()
        _lib.check(_lib.lib().bb_ctx_info(self._ctx, C.byref(nv), C.byref(ss), box, C.byref(nf), C.byref(mr)))
        return dict(nvox=nv.value, ss_total=ss.value, deposit_box=list(box), nnls_fail=nf.value, mc_retries=mr.value)

    def set_halves(self, on):
        _lib.check(_lib.lib().bb_ctx_set_halves(self._ctx, 1 if on else 0))

    def set_graphs(self, on):
        """Batches as replays of a captured HIP graph per batch size (bb_ctx_set_graphs)."""
        self.use_graphs = bool(on)
        _lib.check(_lib.lib().bb_ctx_set_graphs(self._ctx, 1 if on else 0))

    def graph_replays(self):
        return int(_lib.lib().bb_ctx_graph_replays(self._ctx))

    def trace(self, events):
        """Record 5 torch CUDA events at the stage boundaries of eager launches (None: stop)."""
        if events is None:
            _lib.check(_lib.lib().bb_ctx_trace(self._ctx, None))
            return
        arr = (C.c_void_p * 5)(*[e.cuda_event for e in events])
        _lib.check(_lib.lib().bb_ctx_trace(self._ctx, arr))
