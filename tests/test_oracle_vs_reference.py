"""Pins the C oracle to the reference's own Fortran routines.

oracle/_ref/libaladyn_ref.so is built from the unmodified reference sources
(grid_part_connect.f90, grid_part_lib.f90, boris_push.f90, grid_part_util.f90)
plus oracle/ref_harness.f90.  The oracle and the reference are both compiled
without FMA contraction, so equal operation order gives equal bits; the bound
below (1e-14 max-norm relative) allows only last-bit differences.  When the
library is absent (e.g. away from the reference checkout) the committed
golden vectors in tests/golden/ (made by tests/golden/make_golden.py from the
same library) pin the oracle instead (test_oracle_golden.py).
"""
import numpy as np
import pytest

from cases import random_fields, random_particles, small_2d, small_3d
from conftest import rel_maxerr
from oracle.pyoracle import REF_LIB, Oracle, RefHarness

pytestmark = pytest.mark.skipif(not REF_LIB.exists(), reason="reference harness not built")

TOL = 1e-14


def _cases():
    c2, p2, _ = small_2d()
    c3, p3, _ = small_3d()
    return [("2d", c2, p2), ("3d", c3, p3)]


@pytest.mark.parametrize("name,cfg,part", _cases())
def test_gather_matches_reference(name, cfg, part):
    ebf = random_fields(cfg)
    o = Oracle(cfg).gather(part, ebf)
    r = RefHarness(cfg).gather(part, ebf)
    nf = cfg.nfield
    assert rel_maxerr(o[:, :nf], r[:, :nf]) <= TOL


@pytest.mark.parametrize("initial", [False, True])
@pytest.mark.parametrize("name,cfg,part", _cases())
def test_push_matches_reference(name, cfg, part, initial):
    ebf = random_fields(cfg, scale=30.0)
    orc = Oracle(cfg)
    ebfp = orc.gather(part, ebf)
    orc.charge_multiply(part, ebfp)
    po, pr = part.copy(order="F"), part.copy(order="F")
    eo, er = ebfp.copy(order="F"), ebfp.copy(order="F")
    if initial:
        orc.init_momenta(po, eo)
    orc.push(po, eo)
    RefHarness(cfg).push(pr, er, initial=initial)
    D = cfg.curr_ndim
    assert rel_maxerr(po[:, D:2 * D], pr[:, D:2 * D]) <= TOL
    for d in range(D):
        assert rel_maxerr(po[:, d], pr[:, d]) <= TOL
    # stage buffer: dt*v (1:D), old positions (D+1:2D), weight slot (2D+1)
    assert rel_maxerr(eo[:, :2 * D], er[:, :2 * D]) <= TOL
    assert np.array_equal(eo[:, 2 * D].view(np.int64), er[:, 2 * D].view(np.int64))


@pytest.mark.parametrize("name,cfg,part", _cases())
def test_esirkepov_matches_reference(name, cfg, part):
    ebf = random_fields(cfg, scale=30.0)
    orc = Oracle(cfg)
    p = part.copy(order="F")
    ebfp = orc.gather(p, ebf)
    orc.charge_multiply(p, ebfp)
    orc.push(p, ebfp)
    jo = orc.zeros_current()
    jr = orc.zeros_current()
    eo, er = ebfp.copy(order="F"), ebfp.copy(order="F")
    orc.esirkepov(p, eo, jo)
    RefHarness(cfg).esirkepov(p, er, jr)
    assert np.max(np.abs(jr)) > 0
    assert rel_maxerr(jo, jr) <= TOL


@pytest.mark.parametrize("name,cfg,part", _cases())
def test_grid_charge_matches_reference(name, cfg, part):
    ro = Oracle(cfg).grid_charge(part)
    rr = RefHarness(cfg).grid_charge(part)
    assert rel_maxerr(ro, rr) <= TOL


def test_random_particles_against_reference():
    """Random positions/momenta over the whole box (2-D and 3-D)."""
    for cfg in (small_2d()[0], small_3d()[0]):
        part = random_particles(cfg, 4000, margin=0.01, pmax=5.0)
        ebf = random_fields(cfg, seed=9, scale=50.0)
        orc, ref = Oracle(cfg), RefHarness(cfg)
        assert rel_maxerr(orc.gather(part, ebf)[:, :cfg.nfield],
                          ref.gather(part, ebf)[:, :cfg.nfield]) <= TOL


def _se_cases():
    return [(nd, der, cfl, rat) for nd in (2, 3) for der in (2, 3, 4)
            for cfl in (0.5, 0.8, 0.9) for rat in (1.0, 2.0, 4.0)]


@pytest.mark.parametrize("ndim,der_ord,cfl,rat", _se_cases())
def test_se_coeff_matches_reference(ndim, der_ord, cfl, rat):
    """The oracle's Yee x-derivative coefficients against set_field_param
    (grid_field_param.f90:36-99) itself, bit for bit.  The oracle derives
    dy/dx from the cell sizes (dy = yx_rat*dx, set_grid); for these ratios
    that quotient is exact."""
    from aladyn_amd.config import LPConfig

    kw = dict(nz=8, zx_rat=rat) if ndim == 3 else {}
    cfg = LPConfig(ndim=ndim, nx=16, ny=8, k0=20.0, yx_rat=rat, cfl=cfl, der_ord=der_ord, **kw)
    assert cfg.dy / cfg.dx == rat
    want = RefHarness.se_coeff(ndim, der_ord, cfl, rat)
    got = np.array(Oracle(cfg).g.se_coeff[:])
    assert np.array_equal(got, want), (got, want)
    if der_ord > 2:  # the 4-point stencil is really in use
        assert want[1] != 0.0
