"""The envelope model's step on the device (aladyn_env_lpf2_evolve, one
species, one rank, one laser colour) against the oracle's env_lpf2_evolve
(env_evolve_in_time.f90:196-361; its particle routines are pinned bit-equal
to the reference's own and its grid routines by their plane-wave symbols and
face rules, tests/test_oracle_envelope*.py).

Each case: a thermal electron plasma, small random E and B, and a Gaussian
envelope pulse (A^{n-1} = A^n) of amplitude a0 = 1.5.  After one and after
five steps the particles (positions relative to the box, momenta), the
fields, the envelope and the current (after curr_mpi_collect) are compared
in max-norm relative error: FP64 <= 1e-12 after one step and <= 1e-10 after
five; FP32 <= 1e-5 and <= 1e-4.  der_ord 3 and 4 (the high-order Laplacian,
the der_ord 3 first-derivative stencil and the fourth-order gradient) are
checked in 2-D.
"""
import math

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from aladyn_amd.config import LPConfig
from aladyn_amd.inputs import uniform_plasma
from cases import random_fields
from conftest import rel_maxerr
from oracle.pyoracle import Oracle, OrcWindow

pytestmark = pytest.mark.gpu

OML = 2.0 * math.pi  # laser frequency 2 pi/lam0 in c/micron (set_init_param.f90:291)


def _case(ndim, der_ord=2):
    if ndim == 2:
        cfg = LPConfig(ndim=2, nx=48, ny=32, k0=20.0, yx_rat=2.0, cfl=0.8, n_over_nc=0.01,
                       der_ord=der_ord)
        part = uniform_plasma(cfg, (2, 2), x_start_cell=4, seed=11, thermal=0.05)
    else:
        cfg = LPConfig(ndim=3, nx=24, ny=16, nz=16, k0=20.0, yx_rat=3.0, zx_rat=3.0, cfl=0.8,
                       n_over_nc=0.005, der_ord=der_ord)
        part = uniform_plasma(cfg, (2, 2, 2), x_start_cell=2, seed=12, thermal=0.05)
    ebf = random_fields(cfg, seed=13, scale=0.02)
    # Gaussian envelope on every node (ghosts included)
    x = cfg.xmn + (np.arange(cfg.n1p) - 2.0) * cfg.dx
    y = cfg.ymn + (np.arange(cfg.n2p) - 2.0) * cfg.dy
    z = cfg.zmn + (np.arange(cfg.n3p) - 2.0) * cfg.dz if ndim == 3 else np.zeros(1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    xc, lx, w = 0.5 * (cfg.xmn + cfg.xmx), 0.25 * (cfg.xmx - cfg.xmn), 0.3 * (cfg.ymx - cfg.ymn)
    a = 1.5 * np.exp(-((X - xc) / lx) ** 2 - (Y / w) ** 2 - ((Z / w) ** 2 if ndim == 3 else 0.0))
    env = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 4), order="F")
    env[..., 0] = a
    env[..., 1] = 0.3 * a * np.sin(2.0 * math.pi * (X - xc) / lx)
    env[..., 2] = env[..., 0]
    env[..., 3] = env[..., 1]
    return cfg, part, ebf, env


def _run(cfg, part, ebf, env, precision, nsteps, tile=None, block=0):
    orc = Oracle(cfg)
    orc.set_om0(OML)
    p = part.copy(order="F")
    e = ebf.copy(order="F")
    a = env.copy(order="F")
    jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, cfg.curr_ndim + 1), order="F")
    with LPStep(cfg, precision, track_ids=True, tile=tile, block=block) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.put_envelope(env, OML)
        for it in range(nsteps):
            p = orc.env_lpf2_evolve(p, e, a, jc)
            gpu.env_lpf2_evolve(it)
        out = dict(p=p, e=e, a=a, j=jc[..., :cfg.curr_ndim], pg=gpu.get_species(0),
                   eg=gpu.get_fields(), ag=gpu.get_envelope(), jg=gpu.get_current())
    return out


def _check(cfg, r, tol):
    D = cfg.curr_ndim
    assert r["pg"].shape == r["p"].shape, (r["pg"].shape, r["p"].shape)
    for d in range(D):
        span = (cfg.xmx, cfg.ymx, cfg.zmx)[d] - (cfg.xmn, cfg.ymn, cfg.zmn)[d]
        err = np.max(np.abs(r["pg"][:, d] - r["p"][:, d])) / span
        assert err <= tol, f"position {d}: {err}"
    errs = {k: rel_maxerr(r[k + "g"] if k != "p" else r["pg"][:, D:2 * D],
                          r[k] if k != "p" else r["p"][:, D:2 * D])
            for k in ("p", "e", "a", "j")}
    print("errors", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, v in errs.items():
        assert v <= tol, (k, v)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("precision,nsteps,tol", [(PREC_F64, 1, 1e-12), (PREC_F64, 5, 1e-10),
                                                  (PREC_F32, 1, 1e-5), (PREC_F32, 5, 1e-4)])
def test_env_step_matches_oracle(ndim, precision, nsteps, tol):
    print(f"envelope {ndim}-D precision={precision} steps={nsteps}", flush=True)
    cfg, part, ebf, env = _case(ndim)
    _check(cfg, _run(cfg, part, ebf, env, precision, nsteps), tol)


@pytest.mark.parametrize("der_ord", [3, 4])
def test_env_step_high_order_matches_oracle(der_ord):
    print(f"envelope 2-D der_ord={der_ord}", flush=True)
    cfg, part, ebf, env = _case(2, der_ord)
    _check(cfg, _run(cfg, part, ebf, env, PREC_F64, 3), 1e-10)


@pytest.mark.parametrize("ndim,tile,block", [(3, (8, 8, 8), 0), (3, (16, 8, 4), 0),
                                             (3, (8, 8, 4), 512), (2, (32, 16), 0),
                                             (2, (16, 16), 512)])
def test_env_step_other_tiles_match_oracle(ndim, tile, block):
    """The other tile shapes the move phase's tile kernel is compiled for,
    and two threads per node-cell, for which it is not (the move phase then
    runs on the envelope's own tile workgroups): FP32, three steps."""
    print(f"envelope {ndim}-D tile={tile} block={block}", flush=True)
    cfg, part, ebf, env = _case(ndim)
    _check(cfg, _run(cfg, part, ebf, env, PREC_F32, 3, tile=tile, block=block), 1e-4)


def test_env_steps_with_moving_window_match_oracle():
    """lp_window_xshift after each envelope step: the envelope's four
    components shift with the fields (window.f90:327-331)."""
    cfg, part, ebf, env = _case(2)
    orc = Oracle(cfg)
    orc.set_om0(OML)
    win = OrcWindow()
    p, e, a = part.copy(order="F"), ebf.copy(order="F"), env.copy(order="F")
    jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, cfg.curr_ndim + 1), order="F")
    shifts = []
    with LPStep(cfg, PREC_F64, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.put_envelope(env, OML)
        orc.window_xshift(win, 1, 0, 2.0, e, p)
        gpu.lp_window_xshift(1, 0, 2.0)
        for it in range(1, 6):
            p = orc.env_lpf2_evolve(p, e, a, jc)
            gpu.env_lpf2_evolve(it)
            nsh, p = orc.window_xshift(win, 1, it, 2.0, e, p)
            orc.fields_left_xshift(a, nsh)
            assert gpu.lp_window_xshift(1, it, 2.0) == nsh
            shifts.append(nsh)
        r = dict(p=p, e=e, a=a, j=jc[..., :cfg.curr_ndim], pg=gpu.get_species(0),
                 eg=gpu.get_fields(), ag=gpu.get_envelope(), jg=gpu.get_current())
    print("shifts", shifts, flush=True)
    assert sum(shifts) >= 3
    _check(cfg, r, 1e-10)


def test_env_step_needs_the_envelope():
    from aladyn_amd._lib import AladynError

    cfg, part, ebf, _ = _case(2)
    with LPStep(cfg, PREC_F64) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        with pytest.raises(AladynError, match="ESTATE"):
            gpu.env_lpf2_evolve(0)
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.put_envelope(np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 4)), -1.0)
