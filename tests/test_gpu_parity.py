"""GPU parity: the HIP library (through the C-ABI) against the C oracle.

Bars (max-norm relative errors):
  FP64 path: one step <= 1e-12 (particles, fields, current), ten steps
      <= 1e-10, continuity residual <= 1e-12 of max|d rho|.
  FP32 path: one step <= 1e-5 (momenta, fields), continuity <= 1e-5.
The tile store keeps particle identities when track_ids=True so get_species
returns them in input order (survivors only), like the reference's stable
compaction in reset_all_part_dist.
"""
import dataclasses

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from cases import random_fields, random_particles, small_2d, small_3d
from conftest import rel_maxerr
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {"2d": small_2d, "3d": small_3d}


def _run(cfg, part, ebf, precision, nsteps, **kw):
    orc = Oracle(cfg)
    p = part.copy(order="F")
    e = ebf.copy(order="F")
    jc = orc.zeros_current()
    gpu = LPStep(cfg, precision, track_ids=True, **kw)
    gpu.put_species(0, part)
    gpu.put_fields(ebf)
    for it in range(nsteps):
        p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
        gpu.lpf2_evolve(it, initial_time=(it == 0))
    out = dict(p_orc=p, e_orc=e, j_orc=jc, p_gpu=gpu.get_species(0), e_gpu=gpu.get_fields(),
               j_gpu=gpu.get_current())
    gpu.close()
    return out


def _check(cfg, r, tol, check_current=True):
    D = cfg.curr_ndim
    assert r["p_gpu"].shape == r["p_orc"].shape, (r["p_gpu"].shape, r["p_orc"].shape)
    for d in range(D):
        span = (cfg.xmx, cfg.ymx, cfg.zmx)[d] - (cfg.xmn, cfg.ymn, cfg.zmn)[d]
        err = np.max(np.abs(r["p_gpu"][:, d] - r["p_orc"][:, d])) / span
        assert err <= tol, f"position {d}: {err}"
    assert rel_maxerr(r["p_gpu"][:, D:2 * D], r["p_orc"][:, D:2 * D]) <= tol
    assert rel_maxerr(r["e_gpu"], r["e_orc"]) <= tol
    if check_current:
        assert rel_maxerr(r["j_gpu"], r["j_orc"]) <= tol


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f64_one_step(dim):
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F64, 1), 1e-12)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f64_ten_steps(dim):
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F64, 10), 1e-10)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f32_one_step(dim):
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F32, 1), 1e-5)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f32_ten_steps(dim):
    """Ten FP32 steps, the current included: positions agree with the FP64
    oracle to ~1e-8 of the box, so the particles' windows agree, and the
    current's error is the FP32 rounding of its fluxes and fields, like the
    fields' (bar 1e-4 of max|J|, as the fields)."""
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F32, 10), 1e-4)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f32_relayout_every_step(dim):
    """The tile store re-laid out after every step (one-pass count + scatter
    into the other buffer) gives the same ten steps."""
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F32, 10, sort_every=1), 1e-4)


@pytest.mark.parametrize("dim,tile", [("3d", (8, 8, 8)), ("3d", (16, 8, 4)), ("2d", (32, 16))])
def test_f32_other_tile_shapes(dim, tile):
    cfg, part, ebf = CASES[dim]()
    _check(cfg, _run(cfg, part, ebf, PREC_F32, 3, tile=tile), 1e-5)


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-5)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_continuity_residual(dim, precision, tol):
    """(rho^{n+1}-rho^n) + (1/D) div J_raw on the interior, FP64 on device."""
    cfg, _, _ = CASES[dim]()
    part = random_particles(cfg, 20000, pmax=3.0, margin=4.0)
    ebf = random_fields(cfg, scale=20.0)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.set_diagnostics(True)
        gpu.curr_clean()
        gpu.particle_stage(0, initial_time=True)
        r, dmax = gpu.continuity_residual()
    assert dmax > 0
    assert r / dmax <= tol, (r, dmax)


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-5)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_particle_stage_current(dim, precision, tol):
    """The species-loop body alone: raw Esirkepov current vs the oracle."""
    cfg, part, ebf = CASES[dim]()
    orc = Oracle(cfg)
    jc = orc.zeros_current()
    p, _ = orc.particle_stage(part.copy(order="F"), ebf, jc, initial_time=True)
    with LPStep(cfg, precision, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.curr_clean()
        gpu.particle_stage(0, initial_time=True)
        jg = gpu.get_current()
        pg = gpu.get_species(0)
    assert pg.shape == p.shape
    assert rel_maxerr(jg, jc) <= tol
    D = cfg.curr_ndim
    assert rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) <= tol


@pytest.mark.parametrize("precision", [PREC_F64, PREC_F32])
def test_particles_leaving_box_are_removed(precision):
    """cell_part_dist(.false.): particles crossing an open face are dropped
    before deposit and the survivors keep their order."""
    cfg, _, _ = small_2d()
    part = random_particles(cfg, 3000, pmax=0.0, margin=0.6)
    D = cfg.curr_ndim
    rng = np.random.default_rng(5)
    # a band within half a cell of x = xmin moving in -x, one near y = ymax
    # moving in +y: each moves ~0.64 cells in the step and most leave
    part[:200, 0] = cfg.xmn + rng.uniform(0.01, 0.5, 200) * cfg.dx
    part[:200, D] = -50.0
    part[200:400, 1] = cfg.ymx - rng.uniform(0.0, 0.5, 200) * cfg.dy
    part[200:400, D + 1] = 50.0
    ebf = np.zeros(cfg.field_shape, order="F")
    orc = Oracle(cfg)
    jc = orc.zeros_current()
    p = orc.lpf2_evolve(part.copy(order="F"), ebf.copy(order="F"), jc)
    with LPStep(cfg, precision, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.lpf2_evolve(0)
        pg = gpu.get_species(0)
        assert gpu.species_count(0) == p.shape[0]
    assert p.shape[0] < part.shape[0]
    assert pg.shape == p.shape
    tol = 1e-12 if precision == PREC_F64 else 1e-5
    assert rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) <= tol


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-5)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_moments(dim, precision, tol):
    """energy_momenta and the Envar field energy against the oracle."""
    cfg, part, ebf = CASES[dim]()
    orc = Oracle(cfg)
    ek0, em0 = orc.energy_momenta(part)
    fe0 = orc.field_energy(ebf)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        ek, em = gpu.energy_momenta(0)
        fe = gpu.field_energy()
    for i in (0, 1, 5, 6):
        assert abs(ek[i] - ek0[i]) <= tol * max(abs(ek0[i]), 1e-30) + 1e-12 * abs(ek0[5]), i
    assert abs(em - em0) <= tol * em0
    assert rel_maxerr(fe, fe0) <= (1e-12 if precision == PREC_F64 else 1e-6)


def test_error_paths():
    from aladyn_amd import AladynError

    cfg, part, ebf = small_2d()
    bad = part.copy(order="F")
    bad[5, 0] = cfg.xmx + 1.0
    with LPStep(cfg, PREC_F64) as gpu:
        with pytest.raises(AladynError, match="ERANGE"):
            gpu.put_species(0, bad)
        with pytest.raises(AladynError, match="ESTATE"):
            gpu.lpf2_evolve(0)
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.put_species(7, part)


@pytest.mark.parametrize("precision", [PREC_F64, PREC_F32])
def test_error_paths_window_injection_restart(precision):
    """The documented codes of the window, injection and restart calls."""
    from aladyn_amd import AladynError
    from aladyn_amd.inputs import pack_weight, plasma_table

    cfg, part, ebf = small_2d()
    table = plasma_table(cfg, (2, 2), x_start_cell=cfg.nx, x_end_cell=cfg.nx + 10)
    with LPStep(cfg, precision) as gpu:
        # a table before any put: the move reports the missing species
        gpu.put_fields(ebf)
        gpu.set_injection(0, table, nptx=0, targ_in=cfg.xmn, targ_end=100.0)
        assert gpu.lp_window_xshift(4, 0, 1.0) == 0
        with pytest.raises(AladynError, match="ESTATE"):  # moves ~6 cells, then injects
            gpu.lp_window_xshift(10, 10, 1.0)
    with LPStep(cfg, precision) as gpu:
        gpu.set_box_origin(10 * cfg.dx)                  # allowed before put
        assert gpu.box()[0] == pytest.approx(10 * cfg.dx)
        with pytest.raises(AladynError, match="ERANGE"):  # plasma from 4 cells: outside
            gpu.put_species(0, part)
        gpu.set_box_origin(0.0)
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        with pytest.raises(AladynError, match="ESTATE"):  # after put
            gpu.set_box_origin(cfg.dx)
        bad = dict(table)
        bad["ypt"] = np.array([cfg.ymx + 1.0])
        bad["wghyz"] = np.ones((1, 1))
        with pytest.raises(AladynError, match="ERANGE"):
            gpu.set_injection(0, bad)
        rev = dict(table)
        rev["xpt"] = table["xpt"][::-1].copy()
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.set_injection(0, rev)
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.set_injection(0, table, nptx=len(table["xpt"]) + 1)
        # a shift of the whole box is the reference's ier = 2
        assert gpu.lp_window_xshift(4, 0, 1.0) == 0
        with pytest.raises(AladynError, match="ERANGE"):
            gpu.lp_window_xshift(100000, 1, 1.0)
        # injected particles must carry the species' charge
        other = part[:4].copy(order="F")
        D = cfg.curr_ndim
        other[:, 2 * D] = pack_weight(1, 0, np.float32([0.5] * 4))
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.inject_species(0, other)


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_f64_charge_9_current(dim):
    """A species of charge 9 (an Al9+ like ion) with non-dyadic FP32 weights:
    esirkepov_*_curr forms real(sp) :: wght = charge*wgh
    (grid_part_connect.f90:983, 1205), so the product is rounded to FP32
    before it scales the current.  Random momenta up to 3 m c move
    particles across nodes on one, two and three axes, so the register
    sums, the edge layer and the union-window queue are all exercised.  The
    FP64 path's raw current must match the oracle at 1e-12."""
    cfg, _, _ = CASES[dim]()
    part = random_particles(cfg, 20000, pmax=3.0, margin=4.0, charge=9)
    D = cfg.curr_ndim
    w = (part[:, 2 * D].view(np.uint64) >> np.uint64(32)).astype(np.uint32).view(np.float32)
    # the test is sensitive: most 9*w products are inexact in FP32
    assert np.mean((np.float32(9) * w).astype(np.float64) != 9.0 * w.astype(np.float64)) > 0.5
    ebf = random_fields(cfg, scale=2.0)
    orc = Oracle(cfg, lorentz_fact=1.0 / 1000.0)
    jc = orc.zeros_current()
    p, _ = orc.particle_stage(part.copy(order="F"), ebf, jc, initial_time=True)
    cfg9 = dataclasses.replace(cfg, lorentz_fact=(1.0 / 1000.0, 1.0, 1.0, 1.0))
    with LPStep(cfg9, PREC_F64, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.curr_clean()
        gpu.particle_stage(0, initial_time=True)
        jg = gpu.get_current()
        pg = gpu.get_species(0)
    assert pg.shape == p.shape
    assert rel_maxerr(jg, jc) <= 1e-12
    assert rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) <= 1e-12
