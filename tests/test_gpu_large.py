"""The large 3-D FP32 study at its full size: 1024x256x256 cells, 536,870,912
particles (2x2x2 per cell, Maxwellian momenta 1e-3, linear weight up-ramp
over the first 128 cells, n/nc = 0.002), a0 = 4 pulse, 500 steps.  The FP32
path's moments after 500 steps are compared with the FP64 path's on the same
input, as that study specifies: kinetic energy, charge, momentum sums and
field energy within 1e-3.  After the last step each path's grid charge
(set_grid_charge, grid_part_util.f90:240-335) sums to its particles' charge
(the quadratic weights are a partition of unity).

The FP64 run holds about 250 GB of device memory (two 64-byte record
buffers of 3.5 slots per particle, 120 GB each), the FP32 run half of that;
the input is about 30 GB of host memory.  The whole test takes a few
minutes, about 45 s of it on the GPU.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NSTEPS = 500


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_large_3d_fp32_against_fp64_full_size():
    from aladyn_amd import PREC_F32, PREC_F64, LPStep
    from aladyn_amd.inputs import lwfa_3d_large_case

    t = time.perf_counter()
    cfg, part, ebf = lwfa_3d_large_case(workers=min(16, os.cpu_count() or 1))
    npart = part.shape[0]
    print(f"large 3-D study: {cfg.nx}x{cfg.ny}x{cfg.nz}, {npart} particles, "
          f"input made in {time.perf_counter() - t:.1f} s", flush=True)
    out = {}
    for prec in (PREC_F64, PREC_F32):
        with LPStep(cfg, prec) as gpu:
            t = time.perf_counter()
            gpu.put_species(0, part)
            gpu.put_fields(ebf)
            assert gpu.species_count(0) == npart
            print(f"precision={prec}: put in {time.perf_counter() - t:.1f} s", flush=True)
            gpu.kernel_timing(reset=True)
            t = time.perf_counter()
            for it in range(NSTEPS):
                gpu.lpf2_evolve(it, initial_time=(it == 0))
                if (it + 1) % 50 == 0:
                    n = gpu.species_count(0)
                    print(f"precision={prec}: step {it + 1}, {n} particles, "
                          f"{time.perf_counter() - t:.1f} s", flush=True)
            wall = time.perf_counter() - t
            kms, fms = gpu.kernel_timing()
            ek, _ = gpu.energy_momenta(0)
            rho = gpu.grid_charge(0)
            # FP64 sums over 5.4e8 particles in either order: 1e-10 (one
            # particle of the 5.4e8 is 2e-9 of the charge)
            qtol = 1e-10
            print(f"precision={prec}: grid charge {rho.sum():.9e}, particle charge "
                  f"{ek[5]:.9e}", flush=True)
            assert abs(rho.sum() - ek[5]) <= qtol * abs(ek[5])
            del rho
            out[prec] = (gpu.species_count(0), ek, gpu.field_energy())
            print(f"precision={prec}: {NSTEPS} steps in {wall:.2f} s, "
                  f"{npart * NSTEPS / wall:.3e} particle-steps/s, particle kernel {kms:.2f} ms "
                  f"(last steps), field update {fms:.3f} ms", flush=True)
    n64, ek64, fe64 = out[PREC_F64]
    n32, ek32, fe32 = out[PREC_F32]
    print(f"count {n32} vs {n64}; kinetic {_rel(ek32[6], ek64[6]):.2e}; "
          f"field {np.max(np.abs(fe32 - fe64)) / np.sum(fe64):.2e}", flush=True)
    assert abs(n32 - n64) <= max(2, n64 // 1000)
    assert ek64[6] > 0
    assert _rel(ek32[6], ek64[6]) <= 1e-3
    assert _rel(ek32[5], ek64[5]) <= 1e-3
    pscale = max(abs(ek64[i]) for i in range(3))
    assert max(abs(ek32[i] - ek64[i]) for i in range(3)) / pscale <= 1e-3
    assert np.max(np.abs(fe32 - fe64)) / np.sum(fe64) <= 1e-3
