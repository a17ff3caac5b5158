"""Picks the particle store's tile shape, the particle kernel's threads per
tile and the re-layout cadence for a workload by timing them on the device.

The tile shapes with a compiled kernel are 8x8x4, 8x8x8 and 16x8x4 in 3-D
and 16x16, 32x16 in 2-D (a shape whose workgroup does not fit the LDS in
the chosen precision is refused when the context is made, and skipped).
Threads per tile (aladyn_config.block) is one per node-cell, or two per
node-cell for 8x8x4 (FP32) and 16x16: with twice the threads a tile's sorted
particles are cut into slices half as long, so a workgroup processes the
same particles with twice the wavefronts in flight.  The cadence is the
number of steps between re-layouts of the store (sort_every).  Each
candidate runs `warmup` steps untimed, then `steps` timed ones, from the
same input; the fastest mean step wins.

`python -m aladyn_amd.tune --workload lwfa3d` tunes the 3-D LWFA workload
that bench.py measures and records the pick in aladyn_amd/tuned.json,
which bench.py then uses.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

from ._lib import PREC_F32, AladynError
from .lp_step import LPStep

SHAPES_3D = ((8, 8, 4), (8, 8, 8), (16, 8, 4))
SHAPES_2D = ((16, 16), (32, 16))
CADENCES = (32, 64, 128)
# shapes compiled with two threads per node-cell as well (3-D: FP32 only)
TWO_THREADS = {(8, 8, 4), (16, 16)}
TUNED = Path(__file__).resolve().parent / "tuned.json"


def candidates(cfg, precision=PREC_F32):
    """(tile, sort_every, block) triples; block 0 = one thread per node-cell."""
    shapes = SHAPES_2D if cfg.ndim == 2 else SHAPES_3D
    out = []
    for t in shapes:
        blocks = [0]
        if tuple(t) in TWO_THREADS and (cfg.ndim == 2 or precision == PREC_F32):
            blocks.append(2 * math.prod(t))
        out += [(t, k, b) for b in blocks for k in CADENCES]
    return out


def time_steps(cfg, parts, ebf, precision, tile, sort_every, block=0, *, warmup=10, steps=20):
    """Mean ms per lpf2_evolve step of one configuration (None if the shape
    does not fit the grid or the LDS)."""
    try:
        gpu = LPStep(cfg, precision, tile=tile, sort_every=sort_every, block=block)
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
    """Times every candidate; returns (best (tile, sort_every, block),
    {candidate: ms or None})."""
    results = {}
    for tile, k, b in candidates(cfg, precision):
        ms = time_steps(cfg, parts, ebf, precision, tile, k, b, warmup=warmup, steps=steps)
        results[(tile, k, b)] = ms
        if log:
            log(f"tile {tile} sort_every {k} block {b}: "
                + ("does not fit" if ms is None else f"{ms:.3f} ms"))
    timed = {c: v for c, v in results.items() if v is not None}
    return min(timed, key=timed.get), results


def save_pick(workload: str, best, results, path: Path = TUNED, **extra):
    """Records the pick for `workload` in the tuning file (other workloads kept)."""
    data = json.loads(path.read_text()) if path.exists() else {}
    tile, k, b = best
    data[workload] = {"tile": list(tile), "sort_every": k, "block": b, "ms": results[best],
                      "candidates": {f"{list(t)} sort_every={kk} block={bb}": ms
                                     for (t, kk, bb), ms in results.items()}, **extra}
    path.write_text(json.dumps(data, indent=1) + "\n")


def load_pick(workload: str, path: Path = TUNED):
    """The recorded pick for `workload` ({tile, sort_every, block, ...}) or None."""
    if not path.exists():
        return None
    return json.loads(path.read_text()).get(workload)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", choices=("lwfa3d", "lwfa2d"), default="lwfa3d")
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--steps", type=int, default=20)
    args = ap.parse_args(argv)
    from .inputs import lwfa_2d_case, lwfa_3d_case

    cfg, part, ebf = lwfa_2d_case() if args.workload == "lwfa2d" else lwfa_3d_case(seed=0)
    best, res = autotune(cfg, [part], ebf, PREC_F32, warmup=args.warmup, steps=args.steps,
                         log=lambda m: print(m, file=sys.stderr, flush=True))
    save_pick(args.workload, best, res, steps=f"{args.warmup}..{args.warmup + args.steps - 1}")
    print(json.dumps({"workload": args.workload, "tile": list(best[0]), "sort_every": best[1],
                      "block": best[2], "ms": res[best]}), flush=True)


if __name__ == "__main__":
    main()
