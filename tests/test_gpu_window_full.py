"""The 3-D LWFA workload at its full size (256x128x128 cells, 2x2x2
jittered particles per cell, n/nc = 0.005, a0 = 3) with the moving window
as the benchmark runs it: lp_window_xshift every 20 steps from step 100 at
the speed of light (pic_evolve_in_time.f90:124-131), the entering plasma
injected on the device from its target table (particles_inject,
window.f90:106-139), 170 steps, so four moves.  FP32 against FP64 on the
same input: the same box position and particle counts within 1e-6, kinetic
energy, charge, momentum sums and field energy within 1e-3; in each path
the grid charge (set_grid_charge) sums to the particles' charge, since the
quadratic weights are a partition of unity: the injected particles are
live records in the store."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS, START, EVERY = 170, 100, 20


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_lwfa3d_window_fp32_against_fp64_full_size():
    from aladyn_amd import PREC_F32, PREC_F64, LPStep
    from aladyn_amd.inputs import lwfa_3d_case, plasma_table

    cfg, part, ebf = lwfa_3d_case(seed=0)
    table = plasma_table(cfg, (2, 2, 2), x_start_cell=cfg.nx, x_end_cell=cfg.nx + 200)
    out = {}
    for prec in (PREC_F64, PREC_F32):
        print(f"lwfa3d window precision={prec}: {part.shape[0]} particles", flush=True)
        with LPStep(cfg, prec) as gpu:
            gpu.put_species(0, part)
            gpu.put_fields(ebf)
            gpu.set_injection(0, table, nptx=0, targ_in=cfg.xmn,
                              targ_end=float(table["xpt"][-1]) + cfg.dx)
            t = time.perf_counter()
            counts, moved = [], 0
            for it in range(NSTEPS):
                gpu.lpf2_evolve(it, initial_time=(it == 0))
                if it >= START and it % EVERY == 0:
                    moved += gpu.lp_window_xshift(EVERY, it, 1.0)
                    counts.append(gpu.species_count(0))
            ek, _ = gpu.energy_momenta(0)
            rho = gpu.grid_charge(0)
            box = gpu.box()
            print(f"precision={prec}: {NSTEPS} steps in {time.perf_counter() - t:.1f} s, "
                  f"moved {moved} cells, counts {counts}, grid charge {rho.sum():.9e}, "
                  f"particle charge {ek[5]:.9e}", flush=True)
            assert moved > 0 and len(counts) == 4
            assert abs(rho.sum() - ek[5]) <= 1e-10 * abs(ek[5])
            out[prec] = (counts, ek, gpu.field_energy(), box)
    c64, ek64, fe64, b64 = out[PREC_F64]
    c32, ek32, fe32, b32 = out[PREC_F32]
    assert np.allclose(b32, b64, rtol=0, atol=1e-12)
    for n32, n64 in zip(c32, c64):
        assert abs(n32 - n64) <= max(2, n64 // 1000000)
    print(f"kinetic {_rel(ek32[6], ek64[6]):.2e}, charge {_rel(ek32[5], ek64[5]):.2e}, "
          f"field {np.max(np.abs(fe32 - fe64)) / np.sum(fe64):.2e}", flush=True)
    assert ek64[6] > 0
    assert _rel(ek32[6], ek64[6]) <= 1e-3
    assert _rel(ek32[5], ek64[5]) <= 1e-3
    pscale = max(abs(ek64[i]) for i in range(3))
    assert max(abs(ek32[i] - ek64[i]) for i in range(3)) / pscale <= 1e-3
    assert np.max(np.abs(fe32 - fe64)) / np.sum(fe64) <= 1e-3
