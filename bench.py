#!/usr/bin/env python3
"""Throughput of the LP PIC step (gather + push + Esirkepov + Yee) on 1 GPU.

Workload: 3-D LWFA, grid 256x128x128, dx = 0.05, dy = dz = 0.2 (micron),
n/nc = 0.005, 8 particles per cell on a 2x2x2 sub-grid with seeded jitter
(33,554,432 electrons), linearly polarised Gaussian pulse a0 = 3, waist 10.
A step is one full lpf2_evolve: field ghosts, gather, push, removal,
Esirkepov deposit, current fold/normalisation and the Yee update.  The
default run times 200 steps (after 5 untimed ones) and, as the
configuration has it, moves the window from step 100 on (every 20 steps, at
c), with the entering plasma injected on the device from a target table;
--no-window leaves the box fixed.

The FP32 tile path is the measured value; the FP64 tile path (the same
kernels on 64-byte double records) runs the same workload for the FP32/FP64
ratio.  Inputs are resident in HBM before
timing starts.  `roofline` refers to the dominant kernel (the FP32 particle
kernel), its duration measured with HIP events on the context stream over the
timed steps; `cpu_baseline` times the reference's own Fortran gather + push +
deposit (oracle/_ref, built from the reference sources) over a bounded
particle sample, one single-rank process per host core (at most 16) at once,
with the one-core rate beside it.

`roofline.traffic` is the particle kernel's measured HBM bytes per launch
(rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain
device copy of known size), read from the committed profile summary
profiles/k_tile_3d_lwfa_256x128x128.md when it covers this workload; null
otherwise.  `--calibrate-copy` runs only that copy (1 GiB read + 1 GiB
written) for the calibration pass.

--workload lwfa2d measures the 2-D LWFA configuration instead (512x256,
458,752 particles, no window, no CPU leg).

--workload env3d measures the envelope model's step (env_lpf2_evolve) on the
same 3-D plasma with a Gaussian envelope pulse (no CPU leg).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-fp64] [--no-cpu] [--no-window]
                       [--workload lwfa3d|lwfa2d|env3d] [--graphs] [--no-tune]
       python bench.py --calibrate-copy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s (gather+Boris+Esirkepov), FP32 vs FP64 on 1 GPU; % HBM peak"
HBM_PEAK_GBS = 8000.0
PROFILE = ROOT / "profiles" / "k_tile_3d_lwfa_256x128x128.md"


def measured_traffic(nx, ny, nz):
    """HBM bytes per particle-kernel launch from the committed profile summary
    (a line `traffic_bytes_per_launch: <value>` for grid `<nx>x<ny>x<nz>`)."""
    if not PROFILE.exists() or (nx, ny, nz) != (256, 128, 128):
        return None
    for line in PROFILE.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def calibrate_copy():
    """The reference kernels for the HBM byte counters: one device-to-device
    copy of 1 GiB (torch: 1 GiB read, 1 GiB written), then two of the
    library's own kernels on the 256x128x128 grid in FP32: k_field_energy
    (reads one field component's interior, 256*128*128*4 bytes, once per
    component) and k_fill (writes jc whole, ncell*3*4 bytes)."""
    import numpy as np
    import torch

    from aladyn_amd import PREC_F32, LPStep
    from aladyn_amd.config import LPConfig

    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    cfg = LPConfig(ndim=3, nx=256, ny=128, nz=128, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.8,
                   n_over_nc=0.005)
    with LPStep(cfg, PREC_F32) as gpu:
        gpu.put_fields(np.ones(cfg.field_shape, order="F"))
        gpu.field_energy()
        gpu.curr_clean()
        gpu.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4,
                      "k_field_energy_read_bytes": cfg.nx * cfg.ny * cfg.nz * 4,
                      "k_fill_write_bytes": cfg.n1p * cfg.n2p * cfg.n3p * 3 * 4}), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Moving window of the 3-D LWFA configuration ("enabled after step 100"):
# lp_run's call (pic_evolve_in_time.f90:124-131) every WINDOW_EVERY steps from
# WINDOW_START on, at the speed of light; the cadence w_sh is not given by the
# configuration and is chosen here.
WINDOW_START, WINDOW_EVERY, WINDOW_SPEED = 100, 20, 1.0


def run_path(cfg, part, ebf, precision, steps, warmup, table=None, graphs=False, sort_every=0,
             tile=None, block=0):
    """Times `steps` lpf2_evolve calls after `warmup` untimed ones.  With a
    target table the box moves as lp_run moves it and the entering plasma is
    injected on the device from the table.  graphs: steps replayed as HIP
    graphs (the per-kernel HIP-event timing is then not recorded)."""
    from aladyn_amd import LPStep

    gpu = LPStep(cfg, precision, sort_every=sort_every, tile=tile, block=block)
    gpu.put_species(0, part)
    gpu.put_fields(ebf)
    gpu.set_graphs(graphs)
    moves = 0
    if table is not None:
        gpu.set_injection(0, table, nptx=0, targ_in=cfg.xmn,
                          targ_end=float(table["xpt"][-1]) + cfg.dx)

    def step(it):
        nonlocal moves
        gpu.lpf2_evolve(it, initial_time=(it == 0))
        if table is not None and it >= WINDOW_START and it % WINDOW_EVERY == 0:
            gpu.lp_window_xshift(WINDOW_EVERY, it, WINDOW_SPEED)
            moves += 1

    gpu.synchronize()
    for it in range(warmup):
        step(it)
    gpu.synchronize()
    n0 = gpu.species_count(0)
    gpu.kernel_timing(reset=True)
    moves = 0
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        step(it)
    gpu.synchronize()
    t1 = time.perf_counter()
    part_ms, field_ms = gpu.kernel_timing()
    n1 = gpu.species_count(0)
    ek, _ = gpu.energy_momenta(0)
    fe = gpu.field_energy()
    gpu.close()
    wall = t1 - t0
    n_avg = 0.5 * (n0 + n1)
    return dict(wall=wall, ms_per_step=1e3 * wall / steps, n0=n0, n1=n1, moves=moves,
                pps=n_avg * steps / wall, part_ms=part_ms, field_ms=field_ms,
                kinetic=float(ek[6]), field_energy=float(fe.sum()))


_CPU = {}  # the sample, shared with forked workers (copy-on-write)


def _cpu_timed(q, k):
    """Seconds for k reference steps (gather+push+deposit) of particles q."""
    import numpy as np

    t0 = time.perf_counter()
    _CPU["ref"].gather_push_deposit(np.asfortranarray(q), _CPU["ebf"], k)
    return time.perf_counter() - t0


def _cpu_worker(w):
    """Worker w of the many-core leg: its own share of the sample, its own
    copy of the reference harness (one single-rank process per core, as the
    reference would run without MPI)."""
    from oracle.pyoracle import RefHarness

    _CPU["ref"] = RefHarness(_CPU["cfg"])
    q = _CPU["shares"][w]
    t1 = _cpu_timed(q, 1)
    tk = _cpu_timed(q, _CPU["steps"])
    return q.shape[0], t1, tk


def cpu_baseline(cfg, part, ebf, budget_s=15.0):
    """Reference Fortran gather+push+deposit (FP64) on the host's cores.

    The reference routines have no box test, so the sample is drawn from
    particles at least 24 cells from every face and runs 20 steps (under one
    cell per step on each axis, since dt < dx): every index stays inside the
    grid.  Each call copies ebf and clears jc first, so a rate is taken from
    the difference of a 20-step and a 1-step call.  First one core over a
    sample of about budget_s seconds; then one process per core (the box's
    share: at most 16), each over its own share of the inner particles, run
    at the same time, their rates summed.  Must run before the process
    touches the GPU: the workers are forked."""
    import multiprocessing as mp

    import numpy as np

    from oracle.pyoracle import REF_LIB, RefHarness

    if not REF_LIB.exists():
        return None
    steps, margin = 20, 24
    lo = np.array([cfg.xmn, cfg.ymn, cfg.zmn])[: cfg.curr_ndim]
    hi = np.array([cfg.xmx, cfg.ymx, cfg.zmx])[: cfg.curr_ndim]
    dd = np.array([cfg.dx, cfg.dy, cfg.dz])[: cfg.curr_ndim]
    x = part[:, : cfg.curr_ndim]
    inner = np.all((x > lo + margin * dd) & (x < hi - margin * dd), axis=1)
    pool = part[inner]
    _CPU.update(cfg=cfg, ebf=ebf, steps=steps, ref=RefHarness(cfg))

    n0 = min(200_000, pool.shape[0])
    per = max(_cpu_timed(pool[:n0], 3) - _cpu_timed(pool[:n0], 1), 1e-6) / (2 * n0)
    n = int(min(pool.shape[0], max(n0, budget_s / (per * steps))))
    t1 = _cpu_timed(pool[:n], 1)
    tk = _cpu_timed(pool[:n], steps)
    one = n * (steps - 1) / (tk - t1)

    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    share = min(pool.shape[0] // cores, n)
    _CPU["shares"] = [pool[w * share:(w + 1) * share] for w in range(cores)]
    workers = mp.get_context("fork").Pool(cores)
    try:
        res = workers.map(_cpu_worker, range(cores))
        workers.close()
    finally:
        workers.terminate()
        workers.join()  # no worker outlives the leg
    many = sum(m * (steps - 1) / (b - a) for m, a, b in res)
    wall = max(b for _, _, b in res)
    return dict(value=many, unit="particle-steps/s", cores=cores, kind="reference",
                value_one_core=one,
                sample=f"{cores} processes at once, each {share} particles of this workload "
                       f"(at least {margin} cells from every face) x {steps} steps minus 1 "
                       f"step, rates summed ({wall:.1f} s); one core alone: {n} particles, "
                       f"{one:.3g} particle-steps/s. Gather+push+Esirkepov, FP64, reference "
                       f"Fortran routines compiled -O2 -fdefault-real-8")


def envelope_pulse(cfg, a0=3.0, w0=10.0, length=2.0, xc=None):
    """A Gaussian envelope A_R = a0 exp(-((x-xc)/length)^2 - (r/w0)^2) on every
    node, A_I = 0 and A^{n-1} = A^n: env(n1p, n2p, n3p, 4)."""
    import numpy as np

    xc = 0.25 * (cfg.xmn + cfg.xmx) if xc is None else xc
    x = cfg.xmn + (np.arange(cfg.n1p) - 2.0) * cfg.dx
    y = cfg.ymn + (np.arange(cfg.n2p) - 2.0) * cfg.dy
    z = cfg.zmn + (np.arange(cfg.n3p) - 2.0) * cfg.dz if cfg.ndim == 3 else np.zeros(1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a = a0 * np.exp(-((X - xc) / length) ** 2 - (Y * Y + Z * Z) / (w0 * w0))
    env = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 4), order="F")
    env[..., 0] = a
    env[..., 2] = a
    return env


def run_env(cfg, part, env, precision, steps, warmup):
    """Times `steps` env_lpf2_evolve calls after `warmup` untimed ones (the
    envelope model's step, species 0), wake fields starting at zero."""
    import numpy as np

    from aladyn_amd import LPStep

    oml = 2.0 * np.pi / cfg.lam0
    gpu = LPStep(cfg, precision)
    gpu.put_species(0, part)
    gpu.put_fields(np.zeros(cfg.field_shape, order="F"))
    gpu.put_envelope(env, oml)
    for it in range(warmup):
        gpu.env_lpf2_evolve(it)
    gpu.synchronize()
    n0 = gpu.species_count(0)
    gpu.env_kernel_timing(reset=True)
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        gpu.env_lpf2_evolve(it)
    gpu.synchronize()
    wall = time.perf_counter() - t0
    push_ms, move_ms, _ = gpu.env_kernel_timing()
    n1 = gpu.species_count(0)
    gpu.close()
    return {"ms_per_step": 1e3 * wall / steps, "pps": 0.5 * (n0 + n1) * steps / wall,
            "push_ms": push_ms, "move_ms": move_ms, "n0": n0, "n1": n1}


def bench_env(args):
    """The envelope model's step on the 3-D LWFA plasma (--workload env3d)."""
    from aladyn_amd import PREC_F32, PREC_F64
    from aladyn_amd.inputs import lwfa_3d_case

    cfg, part, _ = lwfa_3d_case(args.nx, args.ny, args.nz, seed=0)
    env = envelope_pulse(cfg)
    r32 = run_env(cfg, part, env, PREC_F32, args.steps, args.warmup)
    log(f"env fp32: {r32['ms_per_step']:.3f} ms/step, push {r32['push_ms']:.3f} ms, "
        f"move {r32['move_ms']:.3f} ms")
    r64 = None if args.no_fp64 else run_env(cfg, part, env, PREC_F64, args.steps, args.warmup)
    n_p, cells = 0.5 * (r32["n0"] + r32["n1"]), cfg.nx * cfg.ny * cfg.nz
    # algorithmic bytes, FP32: push reads the record (32 B) and writes p
    # (12 B), gathers E, B and |A|^2/2 with its gradient once per cell (40 B)
    # and deposits the FP64 density (8 B/cell); move reads the record and
    # writes it to its tile (32 + 32 B), gathers |A|^2/2 and its gradient (16
    # B/cell) and writes J (12 B/cell)
    push_b, move_b = n_p * 44 + cells * 48, n_p * 64 + cells * 28
    k_ms = max(r32["push_ms"], r32["move_ms"])
    k_b = push_b if r32["push_ms"] >= r32["move_ms"] else move_b
    achieved = k_b / (k_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC + " (envelope model step)",
        "value": r32["pps"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r32["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded plasma + Gaussian envelope pulse, wake fields from zero)",
        "config": {"workload": f"3-D envelope (ponderomotive) LWFA {cfg.nx}x{cfg.ny}x{cfg.nz}, "
                               f"dx=0.05 dy=dz=0.2, n/nc=0.005, 8 ppc, envelope a0=3 w0=10, "
                               f"env_lpf2_evolve",
                   "particles": int(r32["n0"]), "cells": cells},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": ("k_env_push_tile<3, float> with the density fold"
                                if k_b == push_b else "k_tile<3, 8, 8, 4, 256, float, EnvMove>"),
                     "kernel_ms": k_ms, "algorithmic_bytes_per_launch": k_b},
        "kernels_ms": {"push phase": r32["push_ms"], "move phase": r32["move_ms"]},
        "cpu_baseline": None,
    }
    if r64:
        out["fp64"] = {"value": r64["pps"], "ms_per_step": r64["ms_per_step"],
                       "push_ms": r64["push_ms"], "move_ms": r64["move_ms"]}
        out["fp32_over_fp64"] = r32["pps"] / r64["pps"]
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--no-window", action="store_true")
    ap.add_argument("--workload", choices=("lwfa3d", "lwfa2d", "env3d"), default="lwfa3d",
                    help="lwfa3d: the 3-D LWFA configuration (default); lwfa2d: the 2-D "
                         "LWFA configuration, 512x256 (no window, CPU leg skipped); env3d: "
                         "the envelope model's step on the 3-D LWFA plasma")
    ap.add_argument("--sort-every", type=int, default=0,
                    help="steps between re-layouts of the particle store (0: the library's "
                         "default)")
    ap.add_argument("--graphs", action="store_true",
                    help="also time the FP32 path with steps replayed as HIP graphs")
    ap.add_argument("--no-tune", action="store_true",
                    help="use the library's default tile shape, threads per tile and cadence "
                         "instead of the autotuner's pick recorded in aladyn_amd/tuned.json")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library runs one sub-domain per GPU "
              "and has no multi-GPU exchange (the step would shard by spatial domain, as "
              "the reference's MPI decomposition does)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "env3d":
        bench_env(args)
        return

    from aladyn_amd import PREC_F32, PREC_F64
    from aladyn_amd.inputs import lwfa_2d_case, lwfa_3d_case, plasma_table

    t = time.perf_counter()
    two_d = args.workload == "lwfa2d"
    if two_d:
        cfg, part, ebf = lwfa_2d_case()
        args.no_window, args.no_cpu = True, True
    else:
        cfg, part, ebf = lwfa_3d_case(args.nx, args.ny, args.nz, seed=0)
    log(f"inputs: {part.shape[0]} particles, grid {cfg.nx}x{cfg.ny}x{cfg.nz} "
        f"({time.perf_counter() - t:.1f} s)")

    # the CPU leg forks its workers, so it runs before anything touches the GPU
    cpu = None if args.no_cpu else cpu_baseline(cfg, part, ebf)
    if cpu:
        log(f"cpu baseline: {cpu['value']:.4g} particle-steps/s on {cpu['cores']} cores "
            f"({cpu['sample']})")

    # the autotuner's pick for this workload (python -m aladyn_amd.tune), FP32
    from aladyn_amd.tune import load_pick

    pick = None if args.no_tune else load_pick(args.workload)
    tile32, block32 = None, 0
    se32 = args.sort_every
    if pick:
        tile32, block32 = tuple(pick["tile"]), int(pick["block"])
        se32 = args.sort_every or int(pick["sort_every"])
        log(f"tuned configuration: tile {tile32}, block {block32}, sort_every {se32}")

    # the plasma continues beyond the box: the columns the window brings in
    table = None if args.no_window else plasma_table(cfg, (2, 2, 2), x_start_cell=cfg.nx,
                                                     x_end_cell=cfg.nx + 400)
    r32 = run_path(cfg, part, ebf, PREC_F32, args.steps, args.warmup, table,
                   sort_every=se32, tile=tile32, block=block32)
    log(f"fp32: {r32['ms_per_step']:.3f} ms/step, {r32['pps']:.4g} particle-steps/s, "
        f"particle kernel {r32['part_ms']:.3f} ms, fields {r32['field_ms']:.3f} ms, "
        f"{r32['moves']} window moves")
    rg = None
    if args.graphs:
        rg = run_path(cfg, part, ebf, PREC_F32, args.steps, args.warmup, table, graphs=True,
                      sort_every=se32, tile=tile32, block=block32)
        log(f"fp32 graphs: {rg['ms_per_step']:.3f} ms/step, {rg['pps']:.4g} particle-steps/s")
    r64 = None
    if not args.no_fp64:
        r64 = run_path(cfg, part, ebf, PREC_F64, args.steps, args.warmup, table,  # same steps
                       sort_every=args.sort_every)
        log(f"fp64: {r64['ms_per_step']:.3f} ms/step, {r64['pps']:.4g} particle-steps/s, "
            f"particle kernel {r64['part_ms']:.3f} ms, fields {r64['field_ms']:.3f} ms, "
            f"{r64['moves']} window moves")

    n_p = 0.5 * (r32["n0"] + r32["n1"])
    cells = cfg.nx * cfg.ny * cfg.nz
    # dominant kernel = FP32 particle kernel: particle state read and written
    # (3-D: 32 B read, 28 B written; 2-D: 24 + 20) + field gather + current
    # (3-D: 24 + 12 B/cell; 2-D: 12 + 8)
    bp32, bc32, bp64, bc64 = (44.0, 20.0, 72.0, 40.0) if two_d else (60.0, 36.0, 104.0, 72.0)
    kern_bytes = n_p * bp32 + cells * bc32
    achieved = kern_bytes / (r32["part_ms"] * 1e-3) / 1e9 if r32["part_ms"] > 0 else 0.0
    step_bytes = cfg.algorithmic_bytes(n_p, 32)
    step_gbs = step_bytes / (r32["ms_per_step"] * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": r32["pps"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded plasma + Gaussian laser, no checkpoint)",
        "config": {
            "workload": (f"2-D LWFA {cfg.nx}x{cfg.ny}, dx=dy=0.05, n/nc=0.01, 2x2 ppc for x > "
                         f"{cfg.nx // 8} cells (jittered, seed 0), a0=2 w0=8, quadratic shapes, "
                         f"Yee 2nd order") if two_d else
                        (f"3-D LWFA {cfg.nx}x{cfg.ny}x{cfg.nz}, dx=0.05 dy=dz=0.2, n/nc=0.005, "
                         f"8 ppc (2x2x2, jittered, seed 0), a0=3 w0=10, quadratic shapes, Yee 2nd order"),
            "particles": int(r32["n0"]),
            "cells": cells,
            "path": "fp32 tile-sorted (int32 cell + fp32 offset)",
            "steps_timed": f"{args.warmup}..{args.warmup + args.steps - 1}",
            "tiles": {"tile": list(tile32) if tile32 else ([16, 16] if two_d else [8, 8, 4]),
                      "threads_per_tile": block32 or None,
                      "sort_every": se32 or 64,
                      "source": ("autotuner pick, aladyn_amd/tuned.json" if pick
                                 else "library defaults")},
            "moving_window": None if table is None else
            f"from step {WINDOW_START}, every {WINDOW_EVERY} steps at c, plasma injected from "
            f"the target table ({r32['moves']} moves timed)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None if two_d else measured_traffic(cfg.nx, cfg.ny, cfg.nz),
            "kernel": (f"k_tile<{2 if two_d else 3}, "
                       f"{'x'.join(map(str, tile32 or ((16, 16) if two_d else (8, 8, 4))))} tiles, "
                       f"{block32 or 'one thread per node-cell'}, float>"),
            "kernel_ms": r32["part_ms"],
            "algorithmic_bytes_per_launch": kern_bytes,
        },
        "step_hbm_frac": step_gbs / HBM_PEAK_GBS,
        "cpu_baseline": cpu,
    }
    if r64:
        # FP64 particle kernel: 3-D 104 B/particle (56 R + 48 W) + gather 48 +
        # J 24 B/cell; 2-D 72 B/particle + 40 B/cell
        n64 = 0.5 * (r64["n0"] + r64["n1"])
        b64 = n64 * bp64 + cells * bc64
        a64 = b64 / (r64["part_ms"] * 1e-3) / 1e9 if r64["part_ms"] > 0 else 0.0
        out["fp64"] = {"value": r64["pps"], "ms_per_step": r64["ms_per_step"],
                       "particle_kernel_ms": r64["part_ms"], "field_ms": r64["field_ms"],
                       "particle_kernel_gbs": a64, "particle_kernel_frac": a64 / HBM_PEAK_GBS}
        out["fp32_over_fp64"] = r32["pps"] / r64["pps"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
