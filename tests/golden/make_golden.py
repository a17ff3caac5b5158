"""Writes tests/golden/ref_vectors_*.npz from the reference's own routines.

Also writes tests/golden/ref_se_coeff.json: set_field_param's se_coeff and
its envelope-solver constants hord_der2 and opt_der1
(grid_field_param.f90:36-99) for der_ord 2, 3, 4 over a set of (ndim, cfl,
dy/dx).

Runs the reference Fortran gather (set_part2d/3d_hcell_acc), push
(init_lpf_momenta + lpf_momenta_and_positions), Esirkepov deposit
(esirkepov_2d/3d_curr) and set_grid_charge -- compiled from the reference
sources into oracle/_ref/libaladyn_ref.so by `make -C oracle ref` -- on small
seeded inputs and stores inputs and outputs.  Usage:

    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from aladyn_amd.config import LPConfig  # noqa: E402
from aladyn_amd.inputs import uniform_plasma  # noqa: E402
from cases import random_fields, random_particles  # noqa: E402
from oracle.pyoracle import RefHarness  # noqa: E402


def configs():
    return {
        "2d": LPConfig(ndim=2, nx=20, ny=12, k0=20.0, yx_rat=1.0, cfl=0.9, n_over_nc=0.01),
        "3d": LPConfig(ndim=3, nx=10, ny=8, nz=8, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.9,
                       n_over_nc=0.005),
    }


def make(name, cfg):
    ref = RefHarness(cfg)
    ppc = (2, 2) if cfg.curr_ndim == 2 else (1, 1, 1)
    plasma = uniform_plasma(cfg, ppc, seed=11, thermal=0.3)
    rand = random_particles(cfg, 400, seed=12, pmax=4.0, margin=0.02)
    part = np.asfortranarray(np.concatenate([plasma, rand], axis=0))
    ebf = random_fields(cfg, seed=13, scale=40.0)
    D = cfg.curr_ndim
    out = dict(part=part, ebf=ebf, dt=cfg.dt)
    gathered = ref.gather(part, ebf)
    out["gather"] = gathered[:, : cfg.nfield].copy()
    # push input: charge-multiplied fields (charge = -1 for every particle)
    ebfp = np.zeros_like(gathered, order="F")
    ebfp[:, : cfg.nfield] = -gathered[:, : cfg.nfield]
    for initial in (0, 1):
        p = part.copy(order="F")
        e = ebfp.copy(order="F")
        ref.push(p, e, initial=bool(initial))
        out[f"push{initial}_part"] = p
        out[f"push{initial}_ebfp"] = e
    p = out["push0_part"].copy(order="F")
    e = out["push0_ebfp"].copy(order="F")
    jc = np.zeros(cfg.current_shape, order="F")
    ref.esirkepov(p, e, jc)
    out["esirkepov_jc"] = jc
    out["grid_charge"] = ref.grid_charge(part)
    np.savez_compressed(HERE / f"ref_vectors_{name}.npz", **out)
    print(name, {k: np.asarray(v).shape for k, v in out.items()})


SE_CASES = [(nd, der, cfl, rat) for nd in (2, 3) for der in (2, 3, 4)
            for cfl in (0.5, 0.8, 0.9) for rat in (1.0, 2.0, 4.0)]


def make_se_coeff():
    import json

    rows = []
    for nd, der, cfl, rat in SE_CASES:
        c = RefHarness.se_coeff(nd, der, cfl, rat)
        e = RefHarness.field_param(nd, der, cfl, rat)
        rows.append({"ndim": nd, "der_ord": der, "cfl": cfl, "yx_rat": rat,
                     "se_coeff": [float(c[0]).hex(), float(c[1]).hex()],
                     "hord_der2_opt_der1": [float(e[0]).hex(), float(e[1]).hex()]})
    (HERE / "ref_se_coeff.json").write_text(json.dumps(rows, indent=1) + "\n")
    print("se_coeff", len(rows), "cases")


def make_envelope():
    """ref_env_{2d,3d}.npz: the envelope particle routines of the reference
    (set_env_acc, lpf_env_momenta, set_env_density, set_env_grad_interp,
    lpf_env_positions) on the inputs of tests/test_oracle_envelope.py."""
    from cases import small_2d, small_3d
    from test_oracle_envelope import _chain, _inputs

    for name, mk in (("2d", small_2d), ("3d", small_3d)):
        cfg, _, _ = mk()
        part, ebf, av0 = _inputs(cfg, n=800)
        _, _, av1 = _inputs(cfg, seed=9, n=800)
        out = _chain(RefHarness(cfg), cfg, part, ebf, av0, av1)
        np.savez_compressed(HERE / f"ref_env_{name}.npz", part=part, ebf=ebf, av0=av0, av1=av1,
                            dt=cfg.dt, **{f"out_{k}": v for k, v in out.items()})
        print("envelope", name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    for name, cfg in configs().items():
        make(name, cfg)
    make_se_coeff()
    make_envelope()
