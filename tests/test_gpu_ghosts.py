"""The field ghosts (aladyn_pfields_prepare: field_xyzbd,
mpi_field_interface.f90:53-198) and the current fold (aladyn_curr_mpi_collect:
jc_xyzbd and the djc normalisation, mpi_curr_interface.f90:551-643,
curr_and_fields_util.f90:104-152) on the device, each a single launch for
open boundaries, against the oracle.

Ghosts: random values on every node, ghosts included; in FP64 the device's
ghost values equal the oracle's to the bit (both form 2a - b from the same
nodes in the same order), in FP32 they agree to 1e-6 of max|ebf| (the FP32
grid rounded once).  Fold: particles placed within a cell of every face and
edge deposit into every ghost layer the fold reads; after the stage and the
collect the current matches the oracle's at FP64 1e-12 and FP32 1e-5 of
max|jc|, ghosts included (zeroed where the fold emptied them).
"""
import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from cases import random_fields, random_particles, small_2d, small_3d
from conftest import rel_maxerr
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {"2d": small_2d, "3d": small_3d}


@pytest.mark.parametrize("precision", [PREC_F64, PREC_F32])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_field_ghosts_match_oracle(dim, precision):
    cfg, _, _ = CASES[dim]()
    print(f"field ghosts {dim} precision={precision}", flush=True)
    ebf = random_fields(cfg, seed=31)
    want = ebf.copy(order="F")
    Oracle(cfg).field_xyzbd(want)
    with LPStep(cfg, precision) as gpu:
        gpu.put_fields(ebf)
        gpu.pfields_prepare()
        got = gpu.get_fields()
    assert not np.array_equal(want, ebf)
    if precision == PREC_F64:
        assert np.array_equal(got, want)
    else:
        assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want))


def _edge_particles(cfg, n, seed):
    """Particles within one cell of the faces (a third near each pair of
    faces, so edges and corners are reached), moving outward at |p| ~ 3."""
    rng = np.random.default_rng(seed)
    part = random_particles(cfg, n, seed=seed, pmax=3.0, margin=0.2)
    D = cfg.curr_ndim
    lo = (cfg.xmn, cfg.ymn, cfg.zmn)
    hi = (cfg.xmx, cfg.ymx, cfg.zmx)
    dd = (cfg.dx, cfg.dy, cfg.dz)
    for q in range(n):
        for d in rng.choice(D, size=rng.integers(1, D + 1), replace=False):
            side = rng.integers(2)
            off = rng.uniform(0.05, 0.95) * dd[d]
            part[q, d] = lo[d] + off if side == 0 else hi[d] - off
            part[q, D + d] = -abs(part[q, D + d]) if side == 0 else abs(part[q, D + d])
    return part


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-5)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_current_fold_matches_oracle(dim, precision, tol):
    cfg, _, ebf = CASES[dim]()
    print(f"current fold {dim} precision={precision}", flush=True)
    part = _edge_particles(cfg, 4000, seed=32)
    orc = Oracle(cfg)
    jc = orc.zeros_current()
    orc.particle_stage(part.copy(order="F"), ebf, jc, initial_time=True)
    raw = jc.copy(order="F")
    orc.curr_collect(jc)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.curr_clean()
        gpu.particle_stage(0, initial_time=True)
        graw = gpu.get_current()
        gpu.curr_mpi_collect()
        got = gpu.get_current()
    # the stage reached the ghost layers the fold empties
    ghost = np.ones(raw.shape, bool)
    ghost[cfg.ix1 - 1:cfg.ix2, cfg.jy1 - 1:cfg.jy2, (slice(cfg.kz1 - 1, cfg.kz2)
                                                    if cfg.ndim == 3 else slice(None))] = False
    assert np.max(np.abs(raw[ghost])) > 0.0
    assert rel_maxerr(graw, raw) <= tol
    assert rel_maxerr(got, jc) <= tol
