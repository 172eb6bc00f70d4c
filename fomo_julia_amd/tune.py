"""Tile-shape autotuner for batches of shots.

The fused step's tile is 64 x cols_per_lane columns by rows_per_tile rows per
wavefront (rows_per_tile 0: sized so the tiles fill the GPU's resident
wavefronts in whole rounds).  Which shape is fastest depends on the FD order
(register windows grow with it), the storage type (bytes per load) and the
grid sizes of the batch.  `autotune_batch` times each candidate on the batch's
own members for a few steps, keeps the fastest, and leaves the members reset
(zero wavefields, zero gathers).  Results are cached per (fd_order, storage,
members, cells) class for the life of the process.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import elastic2d as E

CANDIDATES = ((1, 0), (2, 0), (4, 0), (2, 16), (1, 48))
_cache: dict = {}


def _key(dms, n: int, storage: str) -> tuple:
    cells = sum(int(dm.M.nx) * int(dm.M.nz) for dm in dms)
    return (dms[0].p.fd_order, storage, n, round(cells, -5))


def trial_steps(nt: int, steps: int, warm: int) -> int:
    """Timed steps of a trial on members of `nt` steps: the warm-up and timed
    steps (1 .. warm + steps) stay inside every member's nt, since a kernel
    launched past a shot's last step returns at once; even, at least 2."""
    steps = min(steps, nt - warm) // 2 * 2
    if steps < 2:
        raise ValueError(f"members of {nt} steps are too short to time {warm} + 2 steps")
    return steps


def time_batch(b: E.MI355XBackend, dms, wfs, dss, steps: int, warm: int = 4) -> float:
    """ms per step of one batch over `steps` steps (HIP events on the library's stream)."""
    steps = trial_steps(min(int(d.nt) for d in dss), steps, warm)
    lib = b.lib
    n = len(wfs)
    arr = lambda hs: (ctypes.c_void_p * n)(*[h.value for h in hs])
    bh = ctypes.c_void_p()
    b.check(lib.fomo_batch_create(b.ctx, n, arr([d.h for d in dms]), arr([w._h for w in wfs]),
                                  arr([d.h for d in dss]), ctypes.byref(bh)), "fomo_batch_create")
    try:
        stream = torch.cuda.ExternalStream(b.stream_handle())
        b.check(lib.fomo_batch_run(bh, 1, warm), "warm")
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        b.check(lib.fomo_batch_run(bh, 1 + warm, steps), "timed")
        e1.record(stream)
        b.synchronize()
        return e0.elapsed_time(e1) / steps
    finally:
        lib.fomo_batch_destroy(bh)


def autotune_batch(b: E.MI355XBackend, dms, wfs, dss, steps: int = 32,
                   candidates=CANDIDATES, graph_steps: int = 16) -> dict:
    """Chooses cols_per_lane and rows_per_tile for this batch, sets them on
    the backend and returns {"tiling": (cpt, tz), "ms_per_step": ..., "trials": {...}}."""
    storage = getattr(dms[0].M, "storage", "f32")
    key = _key(dms, len(wfs), storage)
    if key in _cache:
        best = _cache[key]
        b.set_tiling(best["tiling"][0], best["tiling"][1], graph_steps)
        return best
    trials = {}
    steps = max(2, steps // 2 * 2)
    for cpt, tz in candidates:
        b.set_tiling(cpt, tz, graph_steps)
        try:
            trials[(cpt, tz)] = time_batch(b, dms, wfs, dss, steps)
        except E._lib.FomoError:
            continue
        for w in wfs:
            E.reset(b, w)
    for d in dss:
        b.check(b.lib.fomo_shot_reset(d.h), "shot_reset")
    b.synchronize()
    cpt, tz = min(trials, key=trials.get)
    b.set_tiling(cpt, tz, graph_steps)
    best = {"tiling": (cpt, tz), "ms_per_step": trials[(cpt, tz)],
            "trials": {f"{c}x{t}": round(v, 4) for (c, t), v in trials.items()}}
    _cache[key] = best
    return best
