"""lpf2_evolve steps replayed as captured HIP graphs (aladyn_set_graphs)
against the ordinary path, step by step, including window moves and steps
with diagnostics (which run outside the graph) in between.

Two contexts run the same input in lock step, one with graphs.  After every
step the host reads both (a synchronisation between replays) and compares
fields and particle counts.  Re-layouts every 5 steps, window moves and two
species make the graphs be captured again several times.  Bars: FP64
fields within 1e-12 relative and equal particle counts each step (the two
differ only in the order of floating-point atomic adds); FP32 fields within
1e-4 after 20 steps, as the 10-step FP32 parity runs.
"""
import dataclasses

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from aladyn_amd.inputs import plasma_table, uniform_plasma
from cases import small_2d, small_3d
from conftest import rel_maxerr

pytestmark = pytest.mark.gpu


def _pair(cfg, precision, parts, ebf):
    out = []
    for graphs in (False, True):
        g = LPStep(cfg, precision, sort_every=5)
        for ic, p in enumerate(parts):
            g.put_species(ic, p)
        g.put_fields(ebf)
        g.set_graphs(graphs)
        out.append(g)
    return out


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-4)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_graph_steps_match_eager_steps(dim, precision, tol):
    cfg, elec, ebf = {"2d": small_2d, "3d": small_3d}[dim]()
    cfg = dataclasses.replace(cfg, nsp=2, lorentz_fact=(1.0, 0.01, 1.0, 1.0))
    ppc = (1, 1) if cfg.ndim == 2 else (1, 1, 1)
    ions = uniform_plasma(cfg, ppc, x_start_cell=cfg.nx // 4, seed=9, charge=1, thermal=0.2)
    print(f"graphs {dim} precision={precision}", flush=True)
    eager, graph = _pair(cfg, precision, [elec, ions], ebf)
    try:
        for it in range(20):
            eager.lpf2_evolve(it, initial_time=(it == 0))
            graph.lpf2_evolve(it, initial_time=(it == 0))
            e0, e1 = eager.get_fields(), graph.get_fields()
            if precision == PREC_F64:
                assert rel_maxerr(e1, e0) <= tol, (it, rel_maxerr(e1, e0))
                for ic in range(2):
                    assert graph.species_count(ic) == eager.species_count(ic), (it, ic)
        assert rel_maxerr(e1, e0) <= tol
        for ic in range(2):
            n0, n1 = eager.species_count(ic), graph.species_count(ic)
            assert abs(n1 - n0) <= (0 if precision == PREC_F64 else 2)
            k0, k1 = eager.energy_momenta(ic)[0][6], graph.energy_momenta(ic)[0][6]
            assert abs(k1 - k0) <= tol * abs(k0)
    finally:
        eager.close()
        graph.close()


@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-12), (PREC_F32, 1e-4)])
def test_graph_steps_follow_window_moves(precision, tol):
    """Window moves (the box and the stores change) between graph steps."""
    cfg, part, ebf = small_2d()
    table = plasma_table(cfg, (2, 2), x_start_cell=cfg.nx, x_end_cell=cfg.nx + 40)
    print(f"graphs window precision={precision}", flush=True)
    eager, graph = _pair(cfg, precision, [part], ebf)
    try:
        for g in (eager, graph):
            g.set_injection(0, table)
        moves = 0
        for it in range(24):
            for g in (eager, graph):
                g.lpf2_evolve(it, initial_time=(it == 0))
            if it % 4 == 3:
                n = [g.lp_window_xshift(1, it, 1.0) for g in (eager, graph)]
                assert n[0] == n[1]
                moves += n[0] > 0
            e0, e1 = eager.get_fields(), graph.get_fields()
            if precision == PREC_F64:
                assert rel_maxerr(e1, e0) <= tol, (it, rel_maxerr(e1, e0))
                assert graph.species_count(0) == eager.species_count(0), it
            assert np.array_equal(eager.box(), graph.box())
        assert moves >= 3
        assert rel_maxerr(e1, e0) <= tol
    finally:
        eager.close()
        graph.close()


@pytest.mark.parametrize("precision,tol,ctol", [(PREC_F64, 1e-12, 1e-12), (PREC_F32, 1e-4, 1e-5)])
def test_graph_steps_with_diagnostic_steps_between(precision, tol, ctol):
    """Every fifth step has diagnostics on (it runs outside the graph); the
    continuity residual of those steps is at round-off, and the run still
    matches the ordinary one.  The plasma stays 4 cells inside every face: a
    particle removed at a face before it deposits changes rho without a
    current, as in the reference, and continuity holds away from the faces."""
    cfg, _, ebf = small_3d()
    part = uniform_plasma(cfg, (2, 2, 2), x_start_cell=4, x_end_cell=cfg.nx - 4,
                          y_cells=(4, cfg.ny - 4), z_cells=(4, cfg.nz - 4), seed=5,
                          thermal=0.05)
    print(f"graphs + diagnostics precision={precision}", flush=True)
    eager, graph = _pair(cfg, precision, [part], ebf)
    try:
        for it in range(16):
            diag = it % 5 == 4
            graph.set_diagnostics(diag)
            eager.lpf2_evolve(it, initial_time=(it == 0))
            graph.lpf2_evolve(it, initial_time=(it == 0))
            if diag:
                r, dmax = graph.continuity_residual()
                assert dmax > 0 and r / dmax <= ctol, (it, r, dmax)
        assert rel_maxerr(graph.get_fields(), eager.get_fields()) <= tol
        if precision == PREC_F64:
            assert graph.species_count(0) == eager.species_count(0)
    finally:
        eager.close()
        graph.close()
