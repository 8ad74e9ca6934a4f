"""Tile shapes of the ANALYTIC deposit picked by measurement.

The deposit (bb_galmod_analytic) gives each workgroup a rectangle of TX x TY pixels whose spectra it accumulates
in LDS. The shape trades the LDS per workgroup (TX TY (nz | 1) 8 bytes: at most one workgroup per CU when it is
large) against the samples that two tiles both gather (smaller tiles: more of them), and the best one depends on
the cube (channels, box, ring count, thickness) and on the batch size. bb_model_create takes the tile of the most
pixels (at most 256, sides at most 3:2) that fits, the squarest of those; tune_tiles times the likelihood batch of a
fit at every shape the deposit accepts and keeps the fastest. The result of the deposit does not depend on the shape (its sums are exact integers in any
order), only its time does. There is no counterpart in the reference (its GALMOD runs per cloud on the CPU).
"""

import ctypes as C

import numpy as np

from . import _lib

# (TX, TY) with TX >= TY; the deposit orients each shape along the projected minor axis (DiskModel.orient_tile)
TILE_SHAPES = ((4, 4), (8, 4), (8, 8), (16, 4), (16, 8), (32, 4), (16, 16), (32, 8), (64, 4), (32, 16))

_CACHE = {}


def _key(like, B):
    st = like.dm.setup
    return (like.bhi[0] - like.blo[0], like.bhi[1] - like.blo[1], st.nsubs, len(st.relint), int(like.plan.nsub),
            bool(like.plan.thick), int(B))


def get_tile(h):
    tx, ty = C.c_int(), C.c_int()
    _lib.check(_lib.lib().bb_model_get_tile(h, C.byref(tx), C.byref(ty)))
    return tx.value, ty.value


def tile_candidates(h, shapes=TILE_SHAPES):
    """The shapes of shapes that the deposit of the model h accepts (its LDS holds the tile), the default first."""
    L = _lib.lib()
    d = get_tile(h)
    out = [tuple(sorted(d, reverse=True))]
    for s in shapes:
        if s in out:
            continue
        if L.bb_model_set_tile(h, s[0], s[1]) == 0:
            out.append(s)
    _lib.check(L.bb_model_set_tile(h, d[0], d[1]))
    return out


def tune_tiles(like, theta, reps=2, shapes=TILE_SHAPES, cache=_CACHE):
    """Set the deposit tile of the likelihood like (a GpuLikelihood) to the fastest of shapes for batches of the
    rows of theta (no negative positive-only parameter): each shape runs the batch once, then reps times between
    HIP events on the current stream. Results are kept in cache by (box, channels, lines, sub-rings, batch size).
    Returns (shape, {shape: ms per batch})."""
    torch = like.torch
    theta = np.atleast_2d(np.asarray(theta, float))
    h = like.handle
    key = _key(like, theta.shape[0])
    if key in cache:
        best, times = cache[key]
        _lib.check(_lib.lib().bb_model_set_tile(h, best[0], best[1]))
        return best, times
    times = {}
    L = _lib.lib()
    for s in tile_candidates(h, shapes):
        _lib.check(L.bb_model_set_tile(h, s[0], s[1]))
        like.loglike(theta)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            like.launch(theta)
            like.collect()
        e1.record()
        e1.synchronize()
        times[s] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    _lib.check(L.bb_model_set_tile(h, best[0], best[1]))
    cache[key] = (best, times)
    return best, times
