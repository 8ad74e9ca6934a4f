"""Many galaxies of uneven cube sizes evaluated together on one GPU.

Each galaxy keeps its own likelihood (GpuLikelihood: its cube, model box, density profile and device buffers).
A step evaluates one batch of parameter vectors for every galaxy: largest cube first, each batch's host tables are
built, copied to the device asynchronously from page-locked buffers and the batch launched, on the next of a few
streams, so that the host work of one galaxy overlaps the device work of the ones before it and the deposits of
small galaxies fill the CUs that the large ones leave idle; one synchronisation ends the step. The reference fits galaxies one after the other, each with its own process pool (bbarolo/galfit.py,
bbarolo/bayesian/fitter.py); it has no batch over galaxies.
"""

import time

import numpy as np

from .bayes import fit_of
from .synth import multi_galaxy_configs
from .tuner import tune_tiles


def build_fits(cfgs, chunk=64):
    """synthetic fits (bayes.fit_of dicts) of configuration tuples, on the current GPU; returns (fits, seconds of
    set-up per galaxy)."""
    fits, secs = [], []
    for cfg in cfgs:
        t = time.perf_counter()
        fits.append(fit_of(cfg, chunk=chunk))
        secs.append(time.perf_counter() - t)
    return fits, np.array(secs)


def draw_thetas(fits, B, seed=1):
    """B parameter vectors per galaxy from its uniform priors (default_rng(seed))."""
    rng = np.random.default_rng(seed)
    out = []
    for f in fits:
        u = rng.random((B, len(f["lo"])))
        out.append(f["lo"] + u * (f["hi"] - f["lo"]))
    return out


class GalaxyBatch:
    """The likelihoods likes (GpuLikelihood, one per galaxy) evaluated together: loglike(thetas) takes one array
    of parameter vectors per galaxy (each at most that likelihood's chunk rows) and returns their log-likelihoods
    (-inf where a positive-only parameter is negative), as the galaxies' own loglike would."""

    def __init__(self, likes, n_streams=4, graphs=False):
        self.likes = list(likes)
        torch = self.likes[0].torch
        self.torch = torch
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]       # (default priority)
        size = [l.dvalid.numel() for l in self.likes]
        self.order = np.argsort(size, kind="stable")[::-1]                  # largest model box first
        self.graphs = bool(graphs)
        self._graph = None          # (key: rows per galaxy, CUDAGraph, static parts, outputs)
        self.n_replays = 0

    def tune(self, thetas, reps=2):
        """Pick the deposit tile of every galaxy for batches like thetas (tuner.tune_tiles); returns the shapes."""
        out = []
        for like, th in zip(self.likes, thetas):
            th = np.atleast_2d(th)[~like.free.negative(th)]
            out.append(tune_tiles(like, th, reps=reps)[0] if len(th) else None)
        return out

    def launch(self, thetas, events=False):
        """Launch the step without waiting: (pending, end events or None). Galaxy by galaxy, largest first, the
        host tables are built, copied and launched on the next of the streams: the host work of a galaxy overlaps
        the device work of the ones before it. The pending record keeps the tables until collect()."""
        torch = self.torch
        for like, th in zip(self.likes, thetas):
            if np.atleast_2d(th).shape[0] > like.chunk:
                raise ValueError(f"{np.atleast_2d(th).shape[0]} vectors for a likelihood of chunk {like.chunk}")
        cur = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(cur)
        pending = [None] * len(self.likes)
        ends = [None] * len(self.likes) if events else None
        for j, g in enumerate(self.order):
            like = self.likes[g]
            th = np.atleast_2d(np.asarray(thetas[g], float))
            ok = np.flatnonzero(~like.free.negative(th))
            s = self.streams[j % len(self.streams)]
            parts = res = None
            with torch.cuda.stream(s):
                if len(ok):
                    # (the tables are copied on s from page-locked buffers, without a host wait, and read there)
                    sr, vals = like.tables(th[ok])
                    parts = like.split_parts(sr, vals, halves=False, pinned=True)
                    res = like.device_batch(parts)
                if events:
                    ends[g] = torch.cuda.Event(enable_timing=True)
                    ends[g].record()
            pending[g] = (th.shape[0], ok, res, parts)
        for s in self.streams:
            cur.wait_stream(s)
        return pending, ends

    def _launch_parts(self, parts):
        """The device work of prepared parts (one list per galaxy, or None) on the streams, largest galaxy first;
        returns each galaxy's output tensors."""
        torch = self.torch
        cur = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(cur)
        res = [None] * len(self.likes)
        for j, g in enumerate(self.order):
            if parts[g] is None:
                continue
            with torch.cuda.stream(self.streams[j % len(self.streams)]):
                res[g] = self.likes[g].device_batch(parts[g])
        for s in self.streams:
            cur.wait_stream(s)
        return res

    def graph_step(self, thetas):
        """The step as one captured HIP graph: the first step with these row counts runs eagerly (allocating the
        buffers), then its launches on all the streams are captured once; every later step copies its tables into
        the graph's input buffers and replays it. Steps with a row whose positive-only parameter is negative (their
        batches shrink) run eagerly. Returns the log-likelihoods, as loglike."""
        torch = self.torch
        ths = [np.atleast_2d(np.asarray(t, float)) for t in thetas]
        if any(np.any(l.free.negative(t)) for l, t in zip(self.likes, ths)):
            return self.loglike(thetas)
        key = tuple(t.shape[0] for t in ths)
        parts = []
        for like, th in zip(self.likes, ths):
            if th.shape[0] > like.chunk:
                raise ValueError(f"{th.shape[0]} vectors for a likelihood of chunk {like.chunk}")
            sr, vals = like.tables(th)
            parts.append(like.split_parts(sr, vals, halves=False) if th.shape[0] else None)
        if self._graph is None or self._graph[0] != key:
            self._graph = None
            self._launch_parts(parts)           # (eager: allocates every buffer, sets the kernel attributes)
            torch.cuda.synchronize()
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                out = self._launch_parts(parts)
            torch.cuda.synchronize()
            # (the graph writes into the likelihoods' buffers of these batch sizes: hold them, since a batch of
            # another size replaces them)
            self._graph = (key, graph, parts, out, [l.buffers_in_use() for l in self.likes])
        else:
            _, graph, static, out, _ = self._graph
            for ps, pn in zip(static, parts):
                for a, b in zip(ps or (), pn or ()):
                    a["d_sr"].copy_(b["d_sr"])
                    if "d_sr_ext" in a:
                        a["d_sr_ext"].copy_(b["d_sr_ext"])
        self._graph[1].replay()
        self.n_replays += 1
        torch.cuda.current_stream().synchronize()
        res = self._graph[3]
        return [like.finish(*like.collect(r)) if r is not None else np.zeros(0)
                for like, r in zip(self.likes, res)]

    def collect(self, pending):
        """The log-likelihoods of a launched step (waits for it)."""
        self.torch.cuda.current_stream().synchronize()
        out = []
        for like, (n, ok, res, _) in zip(self.likes, pending):
            v = np.full(n, -np.inf)
            if res is not None:
                v[ok] = like.finish(*like.collect(res))
            out.append(v)
        return out

    def loglike(self, thetas):
        return self.collect(self.launch(thetas)[0])

    def timed_graph_step(self, thetas):
        """graph_step between HIP events: (log-likelihoods, ms of the step from the table copies to the end)."""
        torch = self.torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = self.graph_step(thetas)
        e1.record()
        e1.synchronize()
        return out, e0.elapsed_time(e1)

    def timed_step(self, thetas):
        """One step with HIP events: (log-likelihoods, ms of the whole step, ms from the step's start to the end of
        each galaxy's batch, ms of host time to launch it)."""
        torch = self.torch
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        t0 = time.perf_counter()
        pending, ends = self.launch(thetas, events=True)
        host = (time.perf_counter() - t0) * 1e3
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        out = self.collect(pending)
        return out, e0.elapsed_time(e1), np.array([e0.elapsed_time(e) for e in ends]), host

    def alone_ms(self, thetas, reps=1):
        """ms of each galaxy's batch run by itself on the idle GPU (after one untimed run)."""
        torch = self.torch
        out = []
        for like, th in zip(self.likes, thetas):
            th = np.atleast_2d(th)[~like.free.negative(th)]
            if not len(th):
                out.append(np.nan)
                continue
            sr, vals = like.tables(th)
            parts = like.split_parts(sr, vals, halves=False)
            like.device_batch(parts)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                like.device_batch(parts)
            e1.record()
            e1.synchronize()
            out.append(e0.elapsed_time(e1) / reps)
        return np.array(out)


def multi_galaxy(n=64, seed=11, chunk=64):
    """The uneven multi-galaxy batch (synth.multi_galaxy_configs(n, seed)) on the current GPU: (GalaxyBatch, fits,
    set-up seconds per galaxy)."""
    fits, secs = build_fits(multi_galaxy_configs(n, seed), chunk)
    return GalaxyBatch([f["like"] for f in fits]), fits, secs
