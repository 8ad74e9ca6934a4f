"""Benchmark: likelihood evaluations per second of the 160x160x96 GLOBAL configuration on one GPU, the wall time
of a Bayesian fit of the 64x64x40 configuration with and without the surrogate screen, and the evaluations per
second of an uneven batch of 64 galaxies (tile shapes picked by the tuner; per-galaxy batch times p50 / p99).

One step = one batch of B parameter vectors (free VROT, VDISP per ring, INC, PA: 82-D) through the likelihood
context, from theta on the host to the log-likelihoods on the host (GpuLikelihood.loglike -> bb_loglike_batch): the
sub-ring tables built on the device, the deconvolved density profile of every vector (ring maps, non-negative least
squares), ANALYTIC GALMOD -> beam convolution fused with the masked Gaussian residual; two halves of the batch on two
streams. Every step takes new vectors (a pool of prior draws). The synthetic cube is built on the GPU (ANALYTIC GALMOD
of the truth rings, beam, noise 3% of the peak). Prints one JSON line (see --help).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu] [--no-bayes] [--no-multi] [--no-mc]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # MI355X HBM3E, bytes/s
# HBM bytes per launch of k_an_tile at batch 256 on the 160x160x96 box, from two rocprofv3 counter passes
# (FETCH_SIZE, WRITE_SIZE) scaled by the same counters of a plain 1 GiB copy (bench.py --copy-calibration)
TRAFFIC_SHAPE = (256, 160, 160, 96)
TRAFFIC_TILE_BYTES = (2850600.0 * 2.0 + 7313122.0 * 1.0) * 1024   # KiB x scale -> bytes
# LDS instructions (wave level) of one such launch (rocprofv3 --pmc SQ_INSTS_LDS) and the rate of one CU for
# 64-bit LDS atomics measured by a probe of ds_add_u64 (6.4-6.8 cycles per wave instruction): one of the deposit's
# two limiters
LDS_INSTS_TILE = 2.286e9
LDS_ATOMIC_PEAK = 256 * 2.4e9 / 6.6          # wave instructions per second over the 256 CUs at 2.4 GHz
# VALU instructions (wave level) of that launch (SQ_INSTS_VALU) and the issue rate: one wave64 instruction per 4
# cycles per SIMD (16 lanes a cycle), 1024 SIMDs at 2.4 GHz: the other limiter
VALU_INSTS_TILE = 1.709e10
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4


def workload(n_theta, B, seed=1):
    """The 160x160x96 configuration on the GPU (truth cube: ANALYTIC GALMOD of the truth rings on the device, beam;
    noise 3% of the peak, seed 7), its likelihood (GpuLikelihood, chunk B) and n_theta parameter vectors drawn from
    the default priors (default_rng(seed))."""
    from bbarolo2_amd.bayes import synthetic_fit
    fit = synthetic_fit(2, chunk=B)
    u = np.random.default_rng(seed).random((n_theta, fit["free"].ndim))
    fit["theta"] = fit["lo"] + u * (fit["hi"] - fit["lo"])
    return fit


_W = {}


def _cpu_worker_init(wl):
    # one core per worker: no BLAS / OpenMP threads of its own (the pool is the parallelism)
    from threadpoolctl import threadpool_limits
    _W["limits"] = threadpool_limits(1)
    _W.update(wl)


def _cpu_eval(i):
    """One likelihood evaluation with the FP64 oracle pipeline (oracle.pipeline: deconvolved profile with FFT
    convolutions, model in the vector's box, FFT beam convolution, residual); returns its wall time."""
    from oracle import pipeline
    w = _W
    if "orc" not in w:
        w["orc"] = pipeline.Likelihood(w["h"], w["par"], w["truth"], w["spec"], w["data"], w["likemask"], w["region"],
                                       w["box"], w["sigma"], w["likefactor"], fft=True)
    th = w["theta"][i % len(w["theta"])]
    t0 = time.perf_counter()
    w["orc"].loglike(th)
    return time.perf_counter() - t0


def _host_cpus():
    """(cores this process may run on, os.cpu_count(), CPU model name)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return avail, os.cpu_count(), model


def cpu_baseline(wl, per_worker=3):
    """The oracle timed on the host cores: one worker process per core this process may run on (one evaluation per
    core, as the reference's sampler pools; started by spawn, since this process holds the GPU), each doing 1
    untimed + per_worker timed evaluations; evals/s = P / median time."""
    import multiprocessing as mp
    P, ncpu, model = _host_cpus()
    ctx = mp.get_context("spawn")
    small = {k: wl[k] for k in ("h", "truth", "par", "data", "likemask", "region", "theta", "box", "sigma",
                                "likefactor")}
    small["spec"] = wl["free"].spec
    with ctx.Pool(P, initializer=_cpu_worker_init, initargs=(small,)) as pool:
        pool.map(_cpu_eval, range(P), chunksize=1)
        t0 = time.perf_counter()
        times = pool.map(_cpu_eval, range(P * per_worker), chunksize=per_worker)
        wall = time.perf_counter() - t0
    med = float(np.median(times))
    return dict(value=P / med, unit="evals/s", cores=P, kind="port",
                sample=f"{P * per_worker} FP64 oracle likelihood evaluations (deconvolved density profile, ANALYTIC "
                       f"model, beam, residual) of the 160x160x96 workload on {P} worker processes (all the cores this "
                       f"process may run on; os.cpu_count() {ncpu}; {model}), 1 untimed each; median "
                       f"{med * 1e3:.0f} ms/eval/core, wall {wall:.1f} s; single core {1 / med:.2f} evals/s",
                single_core=1 / med, cpu_model=model, cpu_count=ncpu)


# The reference's published time of one likelihood call on this geometry, ANALYTIC, INC and PA free, one BLAS
# thread (CHANGES.md:1276-1278, docs/tasks/3dfit/likelihood.md:586-587)
REF_NGC2403_MS = 22.0


def cpu_calibration(n_eval=6, seed=1):
    """The oracle on the geometry of the reference's own likelihood benchmark (performance/bench.py:46-49), one
    worker process, one core: ms per call beside the reference's published 22 ms, so that the CPU baseline can be
    read against the reference's own speed."""
    import multiprocessing as mp
    from bbarolo2_amd.bayes import synthetic_fit
    fit = synthetic_fit("ngc2403", chunk=16)
    u = np.random.default_rng(seed).random((n_eval, fit["free"].ndim))
    fit["theta"] = fit["lo"] + u * (fit["hi"] - fit["lo"])
    small = {k: fit[k] for k in ("h", "truth", "par", "data", "likemask", "region", "theta", "box", "sigma",
                                 "likefactor")}
    small["spec"] = fit["free"].spec
    with mp.get_context("spawn").Pool(1, initializer=_cpu_worker_init, initargs=(small,)) as pool:
        pool.map(_cpu_eval, [0], chunksize=1)
        times = pool.map(_cpu_eval, range(n_eval), chunksize=n_eval)
    ms = 1e3 * float(np.median(times))
    return {"geometry": "150x150x62, 16\" pixels, 5.15 km/s, beam 59\", 6 rings (60:700:120), ANALYTIC, free VROT "
                        "VDISP_SINGLE INC_SINGLE PHI_SINGLE, synthetic truth and noise",
            "oracle_ms_per_call": round(ms, 2), "reference_published_ms_per_call": REF_NGC2403_MS,
            "oracle_over_reference": round(ms / REF_NGC2403_MS, 3), "evaluations": n_eval}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size (an elementwise kernel that reads nbytes and writes nbytes): FETCH_SIZE and
    WRITE_SIZE of its dispatches give the counters' scale."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}), flush=True)


def stage_times(like, theta, reps):
    """ms per batch of each stage (ring maps, profile, deposit, convolution + residual) with the batch on one stream,
    and of the deposit kernel k_an_tile alone (events the library records around its launch)."""
    import torch
    from bbarolo2_amd import _lib
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    tev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for e in ev + tev:
        e.record()
    torch.cuda.synchronize()
    like.set_halves(False)
    like.loglike(theta)
    like.trace(ev)
    _lib.check(_lib.lib().bb_trace_deposit(C.c_void_p(tev[0].cuda_event), C.c_void_p(tev[1].cuda_event)))
    st = np.zeros(4)
    tile = 0.
    try:
        for _ in range(reps):
            like.loglike(theta)
            torch.cuda.synchronize()
            st += [ev[k].elapsed_time(ev[k + 1]) for k in range(4)]
            tile += tev[0].elapsed_time(tev[1])
    finally:
        _lib.check(_lib.lib().bb_trace_deposit(None, None))
        like.trace(None)
        like.set_halves(True)
    return st / reps, tile / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-bayes", action="store_true", help="skip the Bayesian fits (plain and surrogate-screened)")
    ap.add_argument("--no-multi", action="store_true", help="skip the uneven multi-galaxy batch")
    ap.add_argument("--no-mc", action="store_true", help="skip the MC likelihood leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device buffers 5 times (the reference kernel of the HBM counter passes)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (the likelihood batches "
              "shard across GPUs by parameter vectors, one process per GPU).", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    B = a.batch
    import torch
    torch.cuda.set_device(0)
    npool = B * 8
    wl = workload(npool, B)
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(wl)
        print(f"[bench] cpu baseline {cpu['value']:.2f} evals/s on {cpu['cores']} cores", file=sys.stderr, flush=True)
        cpu["calibration"] = cpu_calibration()
        print(f"[bench] cpu calibration {cpu['calibration']}", file=sys.stderr, flush=True)
    like = wl["like"]
    if not like.update_prof:
        raise RuntimeError("the workload fits the geometry without DENS: the profile must be per vector")
    pool = wl["theta"]
    steps = [pool[(k * B) % npool:(k * B) % npool + B] for k in range(max(a.steps, a.warmup, 1))]
    blo, bhi = like.blo, like.bhi
    by, bx, nz = bhi[1] - blo[1], bhi[0] - blo[0], wl["dm"].setup.nsubs
    nvox = by * bx * nz
    for k in range(a.warmup):
        like.loglike(steps[k % len(steps)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ll = like.loglike(steps[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    if not np.all(np.isfinite(ll)):
        raise RuntimeError("a log-likelihood of the timed batch is not finite")
    ms = dt * 1e3 / a.steps
    value = B * a.steps / dt
    print(f"[bench] {value:.1f} evals/s ({ms:.2f} ms per batch of {B}, theta -> log L)", file=sys.stderr, flush=True)
    # diagnostics (not the value): the stages of one batch in sequence on one stream, the deposit kernel alone
    stage, tile_ms = stage_times(like, steps[0], max(2, min(a.steps, 5)))
    names = ["ring_maps", "profile_solve", "galmod_analytic", "conv_residual"]
    # the MC likelihood of the same cube and vectors (GALMODTYPE MC, counter-based stream)
    mc = None
    if not a.no_mc:
        mc = mc_likelihood(wl, steps, B)
        print(f"[bench] MC likelihood {mc['evals_per_s']:.1f} evals/s", file=sys.stderr, flush=True)
    nmap = (like.rings0.nr + 1) * by * bx
    ach = 4 * nvox * B / (tile_ms * 1e-3)
    traffic = TRAFFIC_TILE_BYTES if (B, by, bx, nz) == TRAFFIC_SHAPE else None
    lds = valu = None
    if (B, by, bx, nz) == TRAFFIC_SHAPE:
        rate = LDS_INSTS_TILE / (tile_ms * 1e-3)
        lds = {"insts_per_launch": LDS_INSTS_TILE, "achieved": rate, "peak": LDS_ATOMIC_PEAK,
               "unit": "LDS wave instructions/s", "frac": rate / LDS_ATOMIC_PEAK}
        vrate = VALU_INSTS_TILE / (tile_ms * 1e-3)
        valu = {"insts_per_launch": VALU_INSTS_TILE, "achieved": vrate, "peak": VALU_ISSUE_PEAK,
                "unit": "VALU wave instructions/s", "frac": vrate / VALU_ISSUE_PEAK}
    out = {
        "metric": "likelihood evals/s (160\u00b2\u00d796 cube); Bayes fit wall-time, surrogate vs plain", "value": value,
        "unit": "evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 geometry and sums)",
        "data": "synthetic",
        "config": {"workload": "GLOBAL likelihood, 160x160x96, 8\" pixels, 5.2 km/s, beam 30\", 40 rings, "
                               "ANALYTIC, free VROT VDISP INC PA (82-D), DENS from the deconvolved profile of "
                               "every vector; timed from theta on the host to log L on the host, new vectors every "
                               "step", "batch": B,
                   "model_box": [int(blo[0]), int(blo[1]), int(bhi[0]), int(bhi[1])]},
        "stages_ms_per_step": dict(zip(names, [round(float(x), 4) for x in stage])),
        "stages_note": "one batch, stages in sequence on one stream (the timed steps run two halves on two streams)",
        "roofline": {"bound": "hbm", "kernel": "k_an_tile", "kernel_ms": round(tile_ms, 4), "achieved": ach / 1e9,
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": ach / PEAK_HBM, "traffic": traffic,
                     "traffic_unit": "bytes per launch", "algorithmic_bytes": 4 * nvox * B,
                     "limiter": "LDS atomics and VALU issue, not HBM (lds and valu give their fractions)",
                     "lds": lds, "valu": valu},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample",
                                                                      "calibration", "cpu_model", "cpu_count")},
        "loglike_sample": [float(x) for x in ll[:4]],
        "map_pixels_per_eval": nmap,
    }
    if mc is not None:
        out["mc_likelihood"] = mc
    if cpu is not None:
        out["gpu_over_cpu"] = value / cpu["value"]
    if not a.no_bayes:
        out["bayes_fit"] = bayes_fits()
    if not a.no_multi:
        out["multi_galaxy"] = multi_galaxy_bench()
    print(json.dumps(out), flush=True)


def mc_likelihood(wl, steps, B, n=3):
    """The same cube, vectors and free parameters with the Monte-Carlo model (GALMODTYPE MC, counter-based stream):
    evaluations per second of the likelihood from theta to log L (ring maps and profile as ANALYTIC, MC deposit)."""
    import torch
    from bbarolo2_amd.engine import DiskModel
    from bbarolo2_amd.galmod import GalmodPar
    from bbarolo2_amd.likelihood import GpuLikelihood
    par = GalmodPar(**{k: getattr(wl["par"], k) for k in wl["par"].__dataclass_fields__})
    par.GALMODTYPE = "MC"
    dm = DiskModel(wl["h"], par)
    like = GpuLikelihood(dm, wl["data"], wl["truth"], wl["free"], likemask=wl["likemask"], region=wl["region"],
                         likebox=wl["box"], sigma=wl["sigma"], likefactor=wl["likefactor"], chunk=B)
    like.loglike(steps[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        ll = like.loglike(steps[k % len(steps)])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"evals_per_s": B * n / dt, "ms_per_step": dt * 1e3 / n, "steps": n, "finite": bool(np.all(np.isfinite(ll))),
            "mc_retries": like.info()["mc_retries"]}


def multi_galaxy_bench(n=64, B=32, steps=3, seed=11):
    """The uneven multi-galaxy batch: n synthetic galaxies (cube sides 48-256 pixels, 32-128 channels, 5-40 rings,
    default_rng(seed)), each with its GLOBAL likelihood (free VROT VDISP per ring, INC, PA; DENS from the profile of
    every vector); the deposit tile of each picked by the tuner; then steps of B vectors per galaxy, all galaxies
    launched together on four streams. Reports evals/s over all galaxies and the per-galaxy batch times."""
    from bbarolo2_amd.multi import multi_galaxy, draw_thetas
    t0 = time.perf_counter()
    gb, fits, secs = multi_galaxy(n, seed, chunk=B)
    t_setup = time.perf_counter() - t0
    print(f"[bench] multi-galaxy set-up: {t_setup:.1f} s", file=sys.stderr, flush=True)
    thetas = draw_thetas(fits, B, seed=1)
    t0 = time.perf_counter()
    tiles = gb.tune(thetas, reps=1)
    t_tune = time.perf_counter() - t0
    print(f"[bench] multi-galaxy tile tuning: {t_tune:.1f} s", file=sys.stderr, flush=True)
    gb.loglike(thetas)
    step_ms, done, host = [], [], []
    for _ in range(steps):
        ll, ms, d, hm = gb.timed_step(thetas)
        if not all(np.all(np.isfinite(v)) for v in ll):
            raise RuntimeError("multi-galaxy batch: a log-likelihood is not finite")
        step_ms.append(ms)
        done.append(d)
        host.append(hm)
    # the same step with every galaxy's chain replayed as a captured HIP graph (eager, captured, then replayed)
    from bbarolo2_amd.multi import GalaxyBatch
    gg = GalaxyBatch(gb.likes, n_streams=len(gb.streams), graphs=True)
    gg.graph_step(thetas)
    ref = gg.graph_step(thetas)
    if not all(np.array_equal(a, b) for a, b in zip(ref, ll)):
        raise RuntimeError("multi-galaxy batch: the graph replay differs from the eager step")
    graph_ms = float(np.median([gg.timed_graph_step(thetas)[1] for _ in range(steps)]))
    for l in gb.likes:
        l.set_graphs(False)
    alone = gb.alone_ms(thetas)
    ms = float(np.median(step_ms))
    d = np.median(np.array(done), axis=0)
    vox = [f["h"].DimAx(0) * f["h"].DimAx(1) * f["h"].DimAx(2) for f in fits]
    shapes = {}
    for t in tiles:
        shapes[f"{t[0]}x{t[1]}"] = shapes.get(f"{t[0]}x{t[1]}", 0) + 1
    print(f"[bench] multi-galaxy step {ms:.1f} ms ({n * B / ms * 1e3:.0f} evals/s), as a graph {graph_ms:.1f} ms, "
          f"sum of alone {alone.sum():.1f} ms", file=sys.stderr, flush=True)
    return {"workload": f"{n} galaxies (seed {seed}), cube sides {min(f['h'].DimAx(0) for f in fits)}-"
                        f"{max(f['h'].DimAx(0) for f in fits)} px, {sum(vox) / 1e6:.1f} M voxels in all; GLOBAL "
                        f"likelihood, {B} vectors per galaxy per step, four streams; also with every galaxy's chain as a "
                        f"captured HIP graph",
            "evals_per_s": n * B / ms * 1e3, "ms_per_step": round(ms, 3), "steps": steps,
            "host_launch_ms": round(float(np.median(host)), 3),
            "graph_ms_per_step": round(graph_ms, 3), "graph_evals_per_s": n * B / graph_ms * 1e3,
            "per_galaxy_ms_in_step": {"p50": round(float(np.percentile(d, 50)), 3),
                                      "p99": round(float(np.percentile(d, 99)), 3)},
            "per_galaxy_ms_alone": {"p50": round(float(np.percentile(alone, 50)), 3),
                                    "p99": round(float(np.percentile(alone, 99)), 3),
                                    "sum": round(float(alone.sum()), 3)},
            "setup_s": round(t_setup, 2), "setup_s_per_galaxy_p50": round(float(np.median(secs)), 3),
            "tune_s": round(t_tune, 2), "tiles": shapes}


SCREEN = dict(low_pass="retrain", walk_min_pass=0.1, surrogate_walks=100)


def _bayes_pair(cfg, n_live, walks, seed, train_max, graphs):
    """The plain and the screened nested sampling run of configuration cfg (same seed): wall times and samplers."""
    from bbarolo2_amd import bayes
    fit = bayes.synthetic_fit(cfg, chunk=n_live, graphs=graphs)
    res = {}
    for scr in (False, True):
        t0 = time.perf_counter()
        kw = dict(SCREEN, train_max=train_max) if scr else {}
        s = bayes.nested_fit(fit, n_live=n_live, n_batch=n_live, seed=seed, surrogate=scr, bound="rwalk",
                             walks=walks, **kw)
        res[scr] = (time.perf_counter() - t0, s)
        print(f"[bench] config {cfg} bayes fit surrogate={scr}: {res[scr][0]:.1f} s, {s.n_like} likelihood "
              "evaluations", file=sys.stderr, flush=True)
    (tp, sp), (ts, ss) = res[False], res[True]
    shift, width = bayes.compare(sp, ss)
    return {"plain_s": round(tp, 3), "surrogate_s": round(ts, 3), "speedup": round(tp / ts, 3),
            "plain_n_like": sp.n_like, "surrogate_n_like": ss.n_like, "surrogate_walk_steps": ss.n_screened,
            "surrogate_train_walk_s": round(ss.t_surrogate, 3), "surrogate_trainings": ss.n_train,
            "walk_end_points_judged": ss.n_walk_judged, "walk_end_points_passed": ss.n_walk_passed,
            "plain_like_s": round(sp.t_like, 3), "surrogate_like_s": round(ss.t_like, 3),
            "log_z": [sp.log_z, ss.log_z], "median_shift_sigma_max": shift, "ci68_width_change_max": width}


def bayes_fits(walks=25, seed=3):
    """The second half of the metric: wall time of the batched Bayesian fit of the 128x128x64 cube (8 rings, free
    VROT VDISP: 16-D; 2000 live points per batch, batches as captured HIP graphs), the nested sampler with
    constrained random walks, plain and with its walks screened by the BF16 surrogate trained on the device; and how
    far the screened posterior is from the plain one. The same for the 64x64x40 surrogate configuration (13-D,
    500 live points), where an exact evaluation costs ~3 us and the network's overhead is of the same order."""
    out = {"workload": "128x128x64 cube, 10 km/s channels, beam 40\", 8 rings, ANALYTIC, free VROT VDISP (16-D), DENS "
                       "from the deconvolved profile of every vector; nested sampling, n_live = n_batch = 2000, "
                       f"random walks of {walks} steps, f_live 0.01, batches as captured HIP graphs; screened: walks "
                       "on the surrogate's region on the device, end points judged exactly, retrained at once when "
                       "under 10% pass (most recent 16384 points)"}
    out.update(_bayes_pair(3, 2000, walks, seed, 16384, True))
    small = {"workload": "64x64x40 cube, 20 km/s channels, beam 4x3 px at PA 30, 6 rings, free VROT VDISP INC "
                         "(13-D), n_live = n_batch = 500, otherwise as above (most recent 8192 points)"}
    small.update(_bayes_pair(5, 500, walks, seed, 8192, False))
    out["config_64x64x40"] = small
    return out


if __name__ == "__main__":
    main()
