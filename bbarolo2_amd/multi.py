"""Many galaxies of uneven cube sizes evaluated together on one GPU.

Each galaxy keeps its own likelihood (GpuLikelihood: its context with its cube, model box, profile and device
buffers). A step evaluates one batch of parameter vectors for every galaxy: largest cube first, each batch is staged
in its context's page-locked buffer and launched (sub-ring tables built on the device), on the next of a few
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
    (-inf where the likelihood is zero), as the galaxies' own loglike would. graphs: every galaxy's launches replay
    a captured HIP graph of its batch size (bb_ctx_set_graphs)."""

    def __init__(self, likes, n_streams=4, graphs=False):
        self.likes = list(likes)
        torch = self.likes[0].torch
        self.torch = torch
        self.streams = [torch.cuda.Stream() for _ in range(n_streams)]       # (default priority)
        size = [(l.bhi[0] - l.blo[0]) * (l.bhi[1] - l.blo[1]) * l.dm.setup.nsubs for l in self.likes]
        self.order = np.argsort(size, kind="stable")[::-1]                  # largest model box first
        self.graphs = bool(graphs)
        for l in self.likes:
            l.set_graphs(self.graphs)

    def tune(self, thetas, reps=2):
        """Pick the deposit tile of every galaxy for batches like thetas (tuner.tune_tiles); returns the shapes."""
        out = []
        for like, th in zip(self.likes, thetas):
            th = np.atleast_2d(th)[~like.free.negative(th)]
            out.append(tune_tiles(like, th, reps=reps)[0] if len(th) else None)
        return out

    def launch(self, thetas, events=False):
        """Launch the step without waiting: (pending, end events or None). Galaxy by galaxy, largest first, the
        likelihood context stages the vectors (host) and launches the batch on the next of the streams: the host
        work of a galaxy overlaps the device work of the ones before it."""
        torch = self.torch
        for like, th in zip(self.likes, thetas):
            if np.atleast_2d(th).shape[0] > like.chunk:
                raise ValueError(f"{np.atleast_2d(th).shape[0]} vectors for a likelihood of chunk {like.chunk}")
        cur = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(cur)
        ends = [None] * len(self.likes) if events else None
        for j, g in enumerate(self.order):
            s = self.streams[j % len(self.streams)]
            self.likes[g].launch(thetas[g], stream=s)
            if events:
                ends[g] = torch.cuda.Event(enable_timing=True)
                ends[g].record(s)
        for s in self.streams:
            cur.wait_stream(s)
        return True, ends

    def collect(self, pending=None):
        """The log-likelihoods of a launched step (waits for it)."""
        return [like.collect() for like in self.likes]

    def loglike(self, thetas):
        return self.collect(self.launch(thetas)[0])

    def graph_step(self, thetas):
        """A step whose galaxies replay their captured graphs (GalaxyBatch(graphs=True)); as loglike."""
        if not self.graphs:
            raise RuntimeError("GalaxyBatch(graphs=True) runs the graph step")
        return self.loglike(thetas)

    def timed_graph_step(self, thetas):
        """graph_step between HIP events: (log-likelihoods, ms of the step)."""
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
        torch.cuda.synchronize()
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
            like.loglike(th)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                like.launch(th)
                like.collect()
            e1.record()
            e1.synchronize()
            out.append(e0.elapsed_time(e1) / reps)
        return np.array(out)


def multi_galaxy(n=64, seed=11, chunk=64):
    """The uneven multi-galaxy batch (synth.multi_galaxy_configs(n, seed)) on the current GPU: (GalaxyBatch, fits,
    set-up seconds per galaxy)."""
    fits, secs = build_fits(multi_galaxy_configs(n, seed), chunk)
    return GalaxyBatch([f["like"] for f in fits]), fits, secs
