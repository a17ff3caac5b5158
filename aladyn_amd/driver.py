"""The LP time loop around the step (ALaDyn.F90:66-107, lp_run in
pic_evolve_in_time.f90:113-140, timing in run_data_info.f90:38-72), for
callers that drive the library from Python instead of the Fortran code.
"""
from __future__ import annotations

import dataclasses
import math


def dt_loc(dt: float, tmax: float) -> float:
    """dt_loc = tmax/int(tmax/dt) when tmax > 0, else dt
    (start_all.F90:219-226)."""
    if tmax > 0.0:
        return tmax / float(int(tmax / dt))
    return dt


def with_dt_loc(cfg, tmax: float):
    """The configuration with its time step replaced by dt_loc."""
    return dataclasses.replace(cfg, dt=dt_loc(cfg.dt, tmax))


def lp_cycle(gpu, *, tmax: float, tstart: float = 0.0, w_speed: float = 0.0,
             wi_time: float = 0.0, wf_time: float = math.inf, w_sh: int = 1, on_step=None):
    """Lp_cycle (ALaDyn.F90:89-107) on an LPStep whose configuration carries
    dt_loc: while tnow < tmax, lp_run(tnow, iter) -- lpf2_evolve(iter), then
    lp_window_xshift(w_sh, iter) when w_speed > 0, wi_time <= tnow < wf_time
    and mod(iter, w_sh) == 0 -- and timing's tnow += dt_loc, iter += 1
    (run_data_info.f90:68-71); then, when tnow + dt_loc >= tmax, dt_loc =
    tmax - tnow for the steps left (set_dt on the device).  initial_time
    holds for the first step when tstart < dt_loc (ALaDyn.F90:66-67).
    on_step(iter, tnow) runs after each step (the place of Data_out).
    Returns (iter, tnow)."""
    dt = gpu.cfg.dt
    tnow, it = tstart, 0
    initial = tnow < dt
    while tnow < tmax:
        gpu.lpf2_evolve(it, initial_time=initial)
        if w_speed > 0.0 and wi_time <= tnow < wf_time and it % w_sh == 0:
            gpu.lp_window_xshift(w_sh, it, w_speed)
        tnow += dt
        it += 1
        if on_step is not None:
            on_step(it, tnow)
        if tnow + dt >= tmax and tmax - tnow != dt:
            dt = tmax - tnow
            if dt > 0.0:
                gpu.set_dt(dt)
        initial = False
    return it, tnow


def run_interleaved(steppers, steps: int, *, start_iter: int = 0, initial_time: bool = True,
                    graphs: bool = False):
    """Advances independent runs (one LPStep each) by `steps` lpf2_evolve
    steps, one step of every run in turn.  Each context has its own stream
    and a step waits on the host only for that context's previous step, so
    the runs' kernels overlap on the GPU: small grids, which leave most of
    the chip idle one at a time, fill it together.  initial_time applies to
    the first step of every run.  graphs: replay the steps as captured HIP
    graphs (aladyn_set_graphs), which cuts the host's launch work per step.
    The runs share nothing; the result of each is that of running it alone."""
    if graphs:
        for gpu in steppers:
            gpu.set_graphs(True)
    for k in range(steps):
        it = start_iter + k
        first = initial_time and k == 0
        for gpu in steppers:
            gpu.lpf2_evolve(it, initial_time=first)


def run_batched(steppers, steps: int, *, start_iter: int = 0, initial_time: bool = True,
                graphs: bool = True, groups: int = 1):
    """Advances independent runs (one LPStep each, one species, same precision
    and tile configuration) by `steps` lpf2_evolve steps with every kernel of
    a step launched once for all runs of a batch (aladyn_batch_lpf2_evolve:
    the particle kernel over the work list of every run's tiles), the step's
    launches replayed as one HIP graph.  groups > 1 splits the runs into that
    many batches, each on its own stream, stepped in turn, so that one
    batch's kernels fill the GPU while another's drain.  The result of each
    run is that of running it alone."""
    from contextlib import ExitStack

    from .lp_step import LPBatch

    groups = max(1, min(groups, len(steppers)))
    parts = [steppers[g::groups] for g in range(groups)]
    with ExitStack() as stack:
        batches = [stack.enter_context(LPBatch(p, graphs=graphs)) for p in parts]
        for k in range(steps):
            for b in batches:
                b.lpf2_evolve(start_iter + k, initial_time=initial_time and k == 0)
        for b in batches:
            b.synchronize()
