"""The Gaussian log-likelihood of Bayesian3DFIT (GLOBAL / BAYESIAN), evaluated for batches of parameter
vectors on the GPU.

Mirrors LikelihoodMixin._log_likelihood (bbarolo/bayesian/likelihood.py:567-619) for the default Gaussian
residuals (_calculate_residuals / _residuals_fracerr, likelihood.py:457-531):

    -logL = likefactor * [ 0.5 (S_box + ss_out) / sigma^2 + 0.5 nvox ln(2 pi sigma^2) ]      (+ fracerr terms)

with S_box the masked residual sum inside the model box, ss_out the squared data of the likelihood voxels of
the region outside the box (a summed-area table, likelihood.py:437-440) and nvox the number of likelihood
voxels of the region. theta -> rings follows _update_rings (bayesian/fitter.py:1109-1136): the free values
are stored as float32 in the ring table (bayesian/rings.py:54-61), then used as float64.

The setup that the reference does once per fit (likelihood mask, region, noise) is taken as given: the
caller passes the likelihood mask, the region and sigma (see synth.likelihood_inputs for the conventions
of the synthetic configurations).
"""

import ctypes as C
import math

import numpy as np

from . import _lib
from .engine import _torch, _stream_ptr
from .galmod import SubringPlan, nzq_for
from .profile import DensityProfile, SMOOTH_DEFAULT
from .rings import Rings

# free parameters that change the deconvolved density profile (fitter.py:614-615)
PROFILE_KEYS = ('inc', 'phi', 'xpos', 'ypos', 'radmax', 'z0', 'fthick')

# ring parameters that make the likelihood zero when negative (likelihood.py:532-535)
POSITIVE = ('vrot', 'vdisp', 'inc', 'xpos', 'ypos', 'z0', 'dens', 'zcyl', 'z0thick', 'fthick', 'vdispthick')


class FreePars:
    """The free parameters of a fit: an ordered list of (ring key, 'ring' | 'single'); 'ring' takes nr
    values of theta (one per ring), 'single' one value for all the rings (the FREE keys VROT and
    INC_SINGLE of the reference, bayesian/fitter.py:434-473)."""

    def __init__(self, spec, nr):
        self.spec = [(k.lower(), m) for k, m in spec]
        self.nr = nr
        self.index = {}
        i = 0
        for k, m in self.spec:
            n = nr if m == "ring" else 1
            self.index[k] = np.arange(i, i + n)
            i += n
        self.ndim = i

    def values(self, theta):
        """dict key -> (B, nr) float64 ring values of the free keys, rounded to float32 as the ring table
        of the reference stores them."""
        theta = np.atleast_2d(np.asarray(theta, float))
        B = theta.shape[0]
        out = {}
        for k, m in self.spec:
            v = theta[:, self.index[k]]
            if m != "ring":
                v = np.repeat(v, self.nr, axis=1)
            out[k] = v.astype(np.float32).astype(np.float64)
        return out

    def negative(self, theta):
        """True where a free positive parameter is negative (likelihood.py:537-559)."""
        theta = np.atleast_2d(np.asarray(theta, float))
        bad = np.zeros(theta.shape[0], bool)
        for k, _ in self.spec:
            if k in POSITIVE:
                bad |= np.any(theta[:, self.index[k]].astype(np.float32) < 0, axis=1)
        return bad


def float32_rings(rings):
    """The rings as the ring table of Bayesian3DFIT holds them: every value rounded to float32, radsep
    float32(R[-1] - R[-2]) (bayesian/rings.py:54-76)."""
    r = rings.copy()
    for k, v in r.as_dict().items():
        setattr(r, k, np.asarray(v, np.float32).astype(np.float64))
    if r.nr > 1:
        r.radsep = float(np.float32(r.radii[-1] - r.radii[-2]))
    return r


class GpuLikelihood:
    """Batched log-likelihood of a cube for the free parameters `free` around the rings `rings0`.

    dm: engine.DiskModel of the cube (ANALYTIC or MC). data (nz, ny, nx) float32. likemask (nz, ny, nx) bool
    (maskmode 2: only those voxels; None: all). region (ny, nx) bool (None: all). likebox ((blo), (bhi)):
    the likelihood region's box. sigma: noise per voxel. likefactor: 1 / correlation volume.
    fracerr: fixed fractional model error (None: 0)."""

    def __init__(self, dm, data, rings0, free, likemask=None, region=None, likebox=None, sigma=1.0,
                 likefactor=1.0, fracerr=None, maskmode=2, chunk=256, densprof_smooth=SMOOTH_DEFAULT,
                 likefactor2d=None, beam_arcsec=None, graphs=False):
        torch = _torch()
        self.torch = torch
        self.dm = dm
        self.free = free
        self.rings0 = float32_rings(rings0)
        self.sigma = float(sigma)
        self.likefactor = float(likefactor)
        self.eps = float(fracerr) if fracerr else 0.
        self.chunk = int(chunk)
        nz, ny, nx = data.shape
        if likebox is None:
            likebox = ((0, 0), (nx, ny))
        Lb, Hb = likebox
        if region is None:
            region = np.zeros((ny, nx), bool)
            region[Lb[1]:Hb[1], Lb[0]:Hb[0]] = True
        blank = ~np.isfinite(data) | (data == 0)
        lm = np.ones(data.shape, bool) if likemask is None else np.asarray(likemask, bool)
        if maskmode == 2:
            use = lm & ~blank & region[None]
            dused = np.where(use, data, np.nan)
        elif maskmode == 1:
            use = ~blank & region[None]
            dused = np.where(use, np.where(lm, data, 0.), np.nan)
        else:
            use = ~blank & region[None]
            dused = np.where(use, data, np.nan)
        self.nvox = int(use.sum())
        d2 = np.where(use, np.nan_to_num(dused), 0.).astype(np.float64) ** 2
        dm2 = d2.sum(axis=0)
        self._sat = np.zeros((ny + 1, nx + 1))
        self._sat[1:, 1:] = np.cumsum(np.cumsum(dm2, axis=0), axis=1)
        # the model box (fixed: the centre, heights and radii of the batch are those of rings0)
        for k in ("xpos", "ypos", "z0", "z0thick", "radii"):
            if k in free.index:
                raise NotImplementedError(f"free {k.upper()} moves the model box; not supported by this engine yet")
        blo, bhi = dm.box(self.rings0, final=True)
        blo = [max(blo[0], Lb[0]), max(blo[1], Lb[1])]
        bhi = [max(min(bhi[0], Hb[0]), blo[0]), max(min(bhi[1], Hb[1]), blo[1])]
        self.blo, self.bhi = blo, bhi
        S = self._sat
        inbox = S[bhi[1], bhi[0]] - S[blo[1], bhi[0]] - S[bhi[1], blo[0]] + S[blo[1], blo[0]]
        self.ss_out = max(float(S[-1, -1] - inbox), 0.)
        # the data of the box in the layout of the models (yxz), NaN where not summed
        dbox = dused[:, blo[1]:bhi[1], blo[0]:bhi[0]].transpose(1, 2, 0)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.dvalid = torch.from_numpy(np.ascontiguousarray(dbox, np.float32)).to(self.dev)
        self.nvox_box = int(np.prod(dbox.shape))
        self.plan = SubringPlan(dm.setup, self.rings0, dm.ringmode, final=True)
        self._bufs = {}
        self._streams = None
        # graphs: a batch of a given size runs as one captured HIP graph (ANALYTIC models), replayed with the
        # tables of every new batch copied into its input buffers
        self.use_graphs = bool(graphs) and dm.setup.galmodtype != "MC"
        self._graphs = {}
        # DENS not fitted: the deconvolved profile of the initial rings, and of every parameter vector when the
        # geometry is free (fitter.py:606-615)
        self.useNorm = "dens" not in free.index
        self.update_prof = self.useNorm and any(k in free.index for k in PROFILE_KEYS)
        self.profile = None
        self.dens0 = None
        if self.useNorm:
            v = lm & ~blank
            ymap = np.where(v, data, 0.).astype(np.float64).sum(axis=0)
            nvmap = v.sum(axis=0)
            smap = self.sigma * np.sqrt(nvmap.astype(np.float64))
            sl = (slice(Lb[1], Hb[1]), slice(Lb[0], Hb[0]))
            ok = region[sl] & (nvmap[sl] > 0)
            if beam_arcsec is None:
                bm = dm.header.Bmaj()
                beam_arcsec = abs(bm) * 3600. if bm else None
            l2d = self.likefactor if likefactor2d is None else float(likefactor2d)
            self.profile = DensityProfile(dm, self.rings0, self.plan, (Lb, Hb), ymap[sl], smap[sl], ok,
                                          smooth=densprof_smooth, corr=1. / l2d, beam_arcsec=beam_arcsec)
            sr = self.plan.tables({})[0]
            sr[:, 5, :] = 1.
            d_sr = torch.from_numpy(np.ascontiguousarray(sr)).to(self.dev)
            dens = self.profile.run(sr, d_sr, {}).cpu().numpy()[0].astype(np.float64)
            if not np.all(np.isfinite(dens)):
                raise RuntimeError("the density profile of the initial rings did not converge")
            self.dens0 = dens
            self.rings0.dens = dens.copy()

    def _buffers(self, B, slot=0):
        """Device buffers of a chunk of B vectors; slot: an independent set (chunks in flight on two streams)."""
        b = self._bufs.get((B, slot))
        if b is None:
            torch = self.torch
            by, bx, nz = self.bhi[1] - self.blo[1], self.bhi[0] - self.blo[0], self.dm.setup.nsubs
            L = _lib.lib()
            wl = int(L.bb_residual_work_size(B, self.nvox_box))
            cwl = int(L.bb_conv_residual_work_size(self.dm.handle(), B, by, bx)) if self.dm.cfield is not None else 1
            b = dict(dep=torch.empty((B, by, bx, nz), dtype=torch.float32, device=self.dev),
                     work=torch.empty(wl, dtype=torch.float64, device=self.dev),
                     cwork=torch.empty(max(cwl, 1), dtype=torch.float64, device=self.dev), cwl=cwl,
                     ss=torch.empty(B, dtype=torch.float64, device=self.dev),
                     lg=torch.empty(B, dtype=torch.float64, device=self.dev), wl=wl)
            if self.dm.setup.galmodtype == "MC":
                # (the MC model is convolved into its own buffer, then compared)
                b["mod"] = torch.empty((B, by, bx, nz), dtype=torch.float32, device=self.dev)
            self._bufs = {k: v for k, v in self._bufs.items() if k[0] == B}
            self._bufs[(B, slot)] = b
        return b

    def buffers_in_use(self):
        """References to the device buffers of the current batch size (this likelihood's and its profile's)."""
        return (dict(self._bufs), dict(self.profile._bufs) if self.profile is not None else None)

    def _stream_pair(self):
        if self._streams is None:
            self._streams = (self.torch.cuda.Stream(), self.torch.cuda.Stream())     # (default priority)
        return self._streams

    def tables(self, theta):
        """(sr (B, 16, nsub), values) of the rows of theta: the sub-ring tables of the model (density row 1 when
        the profile is computed per vector) and the ring values of the free keys."""
        vals = self.free.values(theta)
        if self.useNorm and not self.update_prof:
            vals["dens"] = self.dens0
        sr, _ = self.plan.tables(vals)
        if self.update_prof:
            sr[:, 5, :] = 1.
        return sr, vals

    def _device_part(self, sr, d_sr, vals, slot, sr_ext=None, d_sr_ext=None, after_maps=None):
        """Launch one chunk on the current stream with the buffers of slot: (ss, lg, dens) device tensors.
        after_maps: an event recorded once the chunk's ring maps are launched (or, without a profile, at once)."""
        B = sr.shape[0]
        bufs = self._buffers(B, slot)
        dens = None
        if self.update_prof:
            prof = self.profile
            if sr_ext is None:
                sr_ext = prof.ext_tables(vals if vals is not None else {}, B)
            maps = prof.maps(sr, d_sr, sr_ext, None, d_sr_ext, slot)
            if after_maps is not None:
                after_maps.record()
            dens = prof.set_dens(prof.solve(maps, None, slot), d_sr, None, slot)
        elif after_maps is not None:
            after_maps.record()
        if self.dm.setup.galmodtype == "MC":
            tab = sr
            if dens is not None:
                # the clouds of the MC model follow the densities: bound them with those of the profile
                v2 = dict(vals or {})
                v2["dens"] = dens.cpu().numpy().astype(np.float64)
                tab = self.plan.tables(v2)[0]
            mod = self.dm.array_batch_mc(sr, self.blo, self.bhi, convolve=True, out=bufs["mod"], work=bufs["dep"],
                                         d_sr=d_sr, max_clouds=int(self.dm.mc_clouds(tab).max()))
        else:
            dep = self.dm.array_batch(sr, self.blo, self.bhi, convolve=False, work=bufs["dep"], d_sr=d_sr)
            if self.dm.cfield is not None:
                # the beam convolution and the residual in one pass: the convolved models are never stored
                by, bx = self.bhi[1] - self.blo[1], self.bhi[0] - self.blo[0]
                h = self.dm.handle(nzq_for(sr.reshape(16, -1)))
                _lib.check(_lib.lib().bb_conv_residual_gauss(h, dep.data_ptr(), self.dvalid.data_ptr(), B, by, bx,
                                                             self.sigma ** 2, self.eps ** 2, bufs["cwork"].data_ptr(),
                                                             bufs["cwl"], bufs["ss"].data_ptr(),
                                                             bufs["lg"].data_ptr(), _stream_ptr(self.torch)))
                return bufs["ss"], bufs["lg"], dens
            mod = dep
        s2 = self.sigma ** 2
        _lib.check(_lib.lib().bb_residual_gauss(mod.data_ptr(), self.dvalid.data_ptr(), B, self.nvox_box, s2,
                                                self.eps ** 2, bufs["work"].data_ptr(), bufs["wl"],
                                                bufs["ss"].data_ptr(), bufs["lg"].data_ptr(),
                                                _stream_ptr(self.torch)))
        return bufs["ss"], bufs["lg"], dens

    def device_batch(self, parts):
        """Launch the chunks parts (dicts: sr, d_sr, vals, optionally sr_ext, d_sr_ext: host / device sub-ring
        tables) on two streams in turn, so that the latency-bound kernels of one chunk (the profile's least
        squares) run beside the deposit of the other; the current stream waits for all of them. Returns the
        (ss, lg, dens) device tensors of every chunk (valid once the current stream reaches this point)."""
        torch = self.torch
        cur = torch.cuda.current_stream()
        streams = self._stream_pair() if len(parts) > 1 else (cur,)
        out = []
        prev = None
        for i, p in enumerate(parts):
            s = streams[i % len(streams)]
            if s is not cur:
                s.wait_stream(cur)
            if prev is not None:
                # (the ring maps of this chunk start when those of the previous one are launched: its least
                # squares, latency-bound, then run beside them)
                s.wait_event(prev)
            ev = torch.cuda.Event() if len(parts) > 1 else None
            with torch.cuda.stream(s):
                out.append(self._device_part(p["sr"], p["d_sr"], p.get("vals"), i % 2, p.get("sr_ext"),
                                             p.get("d_sr_ext"), ev))
            if s is not cur:
                # (the tables are read on s: their memory must not return to the caller's stream before that)
                for k in ("d_sr", "d_sr_ext"):
                    if k in p:
                        p[k].record_stream(s)
            prev = ev
        for s in streams:
            if s is not cur:
                cur.wait_stream(s)
        return out

    def split_parts(self, sr, vals, halves=True, pinned=False):
        """The chunk (sr, vals) as two halves with their device tables (one part below 64 vectors or without
        halves). pinned: the tables go through page-locked host buffers and are copied asynchronously on the
        current stream (the host does not wait; use the parts on that stream)."""
        torch = self.torch
        B = sr.shape[0]
        cuts = [0, B] if B < 64 or not halves else [0, B // 2, B]

        def up(a):
            t = torch.from_numpy(np.ascontiguousarray(a))
            return t.pin_memory().to(self.dev, non_blocking=True) if pinned else t.to(self.dev)

        parts = []
        for a, b in zip(cuts[:-1], cuts[1:]):
            p = dict(sr=sr[a:b], d_sr=up(sr[a:b]), vals=vals)
            if self.update_prof:
                # (the values of the free keys are per vector, (B, nr))
                v = {k: np.asarray(x)[a:b] for k, x in (vals or {}).items()}
                p["vals"] = v
                p["sr_ext"] = self.profile.ext_tables(v, b - a)
                p["d_sr_ext"] = up(p["sr_ext"])
            parts.append(p)
        return parts

    def _graph_batch(self, sr, vals):
        """The launches of device_batch for the tables sr (and vals), as a HIP graph per batch size: captured on the
        first batch of that size (after one eager run that allocates the buffers), replayed afterwards with the new
        tables copied into the graph's input buffers. The ring orientation of the deposit tiles is that of the first
        batch (it changes the speed, not the result). Returns device_batch's output tensors."""
        torch = self.torch
        B = sr.shape[0]
        parts = self.split_parts(sr, vals)
        g = self._graphs.get(B)
        if g is None:
            self.device_batch(parts)               # (eager: allocates the buffers, sets the kernel attributes)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                out = self.device_batch(parts)
            torch.cuda.synchronize()
            # (the graph writes into the buffers of this batch size: hold them, since a batch of another size
            # replaces them in _bufs)
            g = (graph, parts, out, self.buffers_in_use())
            if len(self._graphs) >= 4:                # (a few batch sizes: a sampler's batches vary in size)
                self._graphs.pop(next(iter(self._graphs)))
            self._graphs[B] = g
        graph, static, out, _ = g
        for p, q in zip(static, parts):
            p["d_sr"].copy_(q["d_sr"])
            if "d_sr_ext" in p:
                p["d_sr_ext"].copy_(q["d_sr_ext"])
        graph.replay()
        return out

    def residuals(self, sr, vals=None):
        """(S_box, lg) per table: device launches for one chunk of sub-ring tables (B, 16, nsub), in parts on
        two streams; with a per-vector profile, vals are the ring values of the free keys (for the annulus)."""
        if self.use_graphs:
            res = self._graph_batch(sr, vals)
        else:
            res = self.device_batch(self.split_parts(sr, vals))
        self.torch.cuda.current_stream().synchronize()
        return self.collect(res)

    def launch(self, theta, halves=False):
        """Launch the evaluation of the rows of theta (at most chunk rows, no negative positive-only parameter) on
        the current stream and return at once: the device tensors for collect(), valid once that stream reaches
        this point. halves: as residuals, two halves on two streams."""
        sr, vals = self.tables(np.atleast_2d(np.asarray(theta, float)))
        return self.device_batch(self.split_parts(sr, vals, halves))

    def collect(self, res):
        """(S_box, lg) of the device tensors of launch (or device_batch), as numpy arrays; the density profiles of
        the chunk are kept in last_dens."""
        ss = np.concatenate([r[0].cpu().numpy() for r in res])
        lg = np.concatenate([r[1].cpu().numpy() for r in res])
        if res[0][2] is not None:
            self.last_dens = np.concatenate([r[2].cpu().numpy() for r in res])
            if not np.all(np.isfinite(self.last_dens)):
                raise RuntimeError("the density profile did not converge (non-negative least squares)")
        return ss, lg

    def loglike(self, theta):
        """log-likelihood of each row of theta (B, ndim) -> (B,) float64 (-inf where a positive parameter
        is negative)."""
        theta = np.atleast_2d(np.asarray(theta, float))
        out = np.full(theta.shape[0], -np.inf)
        ok = ~self.free.negative(theta)
        idx = np.flatnonzero(ok)
        for c in range(0, len(idx), self.chunk):
            ii = idx[c:c + self.chunk]
            out[ii] = self.finish(*self.residuals(*self.tables(theta[ii])))
        return out

    def finish(self, ss, lg):
        """The log-likelihoods of the residual sums (S_box, lg) of residuals (likelihood.py:567-619)."""
        s2 = self.sigma ** 2
        if self.eps > 0:
            nll = 0.5 * (ss + self.ss_out / s2) + 0.5 * self.nvox * math.log(2 * math.pi * s2) + 0.5 * lg
        else:
            nll = 0.5 * (ss + self.ss_out) / s2 + 0.5 * self.nvox * math.log(2 * math.pi * s2)
        return -self.likefactor * nll
