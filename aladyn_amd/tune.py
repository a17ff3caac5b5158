"""Picks the tile shape and re-layout cadence of the particle store for a
workload by timing them on the device.

The tile shapes with a compiled kernel are 8x8x4, 8x8x8 and 16x8x4 in 3-D
and 16x16, 32x16 in 2-D (a shape whose workgroup does not fit the LDS in
the chosen precision is refused when the context is made, and skipped);
the cadence is the number of steps between re-layouts of the store
(sort_every).  Each candidate runs `warmup` steps untimed, then `steps`
timed ones, from the same input; the fastest mean step wins.
"""
from __future__ import annotations

import time

from ._lib import PREC_F32, AladynError
from .lp_step import LPStep

SHAPES_3D = ((8, 8, 4), (8, 8, 8), (16, 8, 4))
SHAPES_2D = ((16, 16), (32, 16))
CADENCES = (32, 64, 128)


def candidates(cfg):
    shapes = SHAPES_2D if cfg.ndim == 2 else SHAPES_3D
    return [(t, k) for t in shapes for k in CADENCES]


def time_steps(cfg, parts, ebf, precision, tile, sort_every, *, warmup=10, steps=20):
    """Mean ms per lpf2_evolve step of one configuration (None if the shape
    does not fit the grid)."""
    try:
        gpu = LPStep(cfg, precision, tile=tile, sort_every=sort_every)
    except AladynError:
        return None
    with gpu:
        for ic, p in enumerate(parts):
            gpu.put_species(ic, p)
        gpu.put_fields(ebf)
        for it in range(warmup):
            gpu.lpf2_evolve(it, initial_time=(it == 0))
        gpu.synchronize()
        t0 = time.perf_counter()
        for it in range(warmup, warmup + steps):
            gpu.lpf2_evolve(it)
        gpu.synchronize()
        return 1e3 * (time.perf_counter() - t0) / steps


def autotune(cfg, parts, ebf, precision=PREC_F32, *, warmup=10, steps=20, log=None):
    """Times every candidate; returns (best (tile, sort_every), {candidate: ms})."""
    results = {}
    for tile, k in candidates(cfg):
        ms = time_steps(cfg, parts, ebf, precision, tile, k, warmup=warmup, steps=steps)
        results[(tile, k)] = ms
        if log:
            log(f"tile {tile} sort_every {k}: " + ("does not fit" if ms is None else f"{ms:.3f} ms"))
    timed = {c: v for c, v in results.items() if v is not None}
    return min(timed, key=timed.get), results
