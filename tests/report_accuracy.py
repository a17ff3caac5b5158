"""Accuracy of the GPU paths against the FP64 oracle, as numbers (the tests
check bars; this prints what is observed).  Needs a GPU:

    python tests/report_accuracy.py

For the 2-D and 3-D laser-plasma cases of the parity tests (48x32 with 2x2
ppc; 24x16x16 with 2x2x2 ppc; a0 = 2, thermal momenta 0.05), FP64 and FP32:

* one step: max-norm relative error of positions (over the box size),
  momenta, fields and raw current;
* ten steps: the same for positions, momenta and fields;
* 200 steps: relative error of the kinetic energy, total charge and field
  energy (per component, over the total);
* the continuity residual of one stage on the device, over max|drho|.

Prints a markdown table.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

from aladyn_amd import PREC_F32, PREC_F64, LPStep  # noqa: E402
from cases import random_fields, random_particles, small_2d, small_3d  # noqa: E402
from conftest import rel_maxerr  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402


def steps_errors(cfg, part, ebf, precision, nsteps):
    orc = Oracle(cfg)
    p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
    with LPStep(cfg, precision, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        for it in range(nsteps):
            p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
            gpu.lpf2_evolve(it, initial_time=(it == 0))
        pg, eg, jg = gpu.get_species(0), gpu.get_fields(), gpu.get_current()
    D = cfg.curr_ndim
    span = max((cfg.xmx - cfg.xmn), (cfg.ymx - cfg.ymn))
    pos = np.max(np.abs(pg[:, :D] - p[:, :D])) / span if pg.shape == p.shape else float("nan")
    mom = rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) if pg.shape == p.shape else float("nan")
    return pos, mom, rel_maxerr(eg, e), rel_maxerr(jg, jc)


def moments_errors(cfg, part, ebf, precision, nsteps=200):
    orc = Oracle(cfg)
    p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
    for it in range(nsteps):
        p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
    ek0, _ = orc.energy_momenta(p)
    fe0 = orc.field_energy(e)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        for it in range(nsteps):
            gpu.lpf2_evolve(it, initial_time=(it == 0))
        ek1, _ = gpu.energy_momenta(0)
        fe1 = gpu.field_energy()
    rel = lambda a, b: abs(a - b) / abs(b)
    return rel(ek1[6], ek0[6]), rel(ek1[5], ek0[5]), float(np.max(np.abs(fe1 - fe0)) / fe0.sum())


def continuity(cfg, precision):
    part = random_particles(cfg, 20000, pmax=3.0, margin=4.0)
    ebf = random_fields(cfg, scale=20.0)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.set_diagnostics(True)
        gpu.curr_clean()
        gpu.particle_stage(0, initial_time=True)
        r, dmax = gpu.continuity_residual()
    return r / dmax


def main():
    rows = []
    for name, mk in (("2-D 48x32", small_2d), ("3-D 24x16x16", small_3d)):
        cfg, part, ebf = mk()
        for prec in (PREC_F64, PREC_F32):
            print(f"{name} precision={prec}", flush=True)
            one = steps_errors(cfg, part, ebf, prec, 1)
            ten = steps_errors(cfg, part, ebf, prec, 10)
            mo = moments_errors(cfg, part, ebf, prec)
            cr = continuity(cfg, prec)
            rows.append((name, prec, one, ten, mo, cr))
    print("| case | path | 1 step: x, p, E/B, J | 10 steps: x, p, E/B | 200 steps: kinetic, "
          "charge, field energy | continuity |")
    print("|---|---|---|---|---|---|")
    f = lambda v: f"{v:.1e}"
    for name, prec, one, ten, mo, cr in rows:
        print(f"| {name} | FP{prec} | {', '.join(map(f, one))} | {', '.join(map(f, ten[:3]))} | "
              f"{', '.join(map(f, mo))} | {f(cr)} |")


if __name__ == "__main__":
    main()
