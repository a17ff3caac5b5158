"""Moving window (lp_window_xshift, window.f90:276-347).

CPU: the oracle's shift is pinned by what the reference's definition fixes:
the field interior moves nshx cells to the left with zeros entering on the
right and the ghosts untouched (fields_left_xshift,
mpi_field_interface.f90:299-359); the box moves by dx*nshx; the particles at
or behind the new left face leave, the others keep their order
(reset_all_part_dist, mpi_part_interface.f90:494-498); nshx follows
nint(xlapse/dx) of the accumulated w_speed*dt*witr, and init_iter = 0 resets.

GPU: steps, a window move, injection of the plasma entering on the right and
more steps through the C-ABI, against the oracle doing the same; bars as in
test_gpu_parity's ten-step runs (FP64 1e-10, FP32 1e-4).
"""
import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64
from aladyn_amd.inputs import uniform_plasma
from cases import small_2d, small_3d
from conftest import rel_maxerr
from oracle.pyoracle import Oracle, OrcWindow


def test_oracle_window_shift_moves_fields_box_and_particles():
    cfg, part, ebf = small_2d()
    rng = np.random.default_rng(5)
    ebf = np.asfortranarray(rng.normal(size=cfg.field_shape))
    orc = Oracle(cfg)
    win = OrcWindow()
    assert orc.window_xshift(win, 10, 0, 1.0, ebf, part)[0] == 0  # reset
    before = ebf.copy()
    x0 = orc.g.xmn
    nsh, kept = orc.window_xshift(win, 10, 10, 1.0, ebf, part.copy(order="F"))
    assert nsh == int(np.floor(cfg.dt * 10 / cfg.dx + 0.5)) and nsh > 0
    i1, i2 = cfg.ix1 - 1, cfg.ix2 - 1  # 0-based interior
    j = slice(cfg.jy1 - 1, cfg.jy2)
    assert np.array_equal(ebf[i1:i2 + 1 - nsh, j], before[i1 + nsh:i2 + 1, j])
    assert np.all(ebf[i2 + 1 - nsh:i2 + 1, j] == 0.0)
    assert np.array_equal(ebf[:i1], before[:i1]) and np.array_equal(ebf[i2 + 1:], before[i2 + 1:])
    assert np.array_equal(ebf[:, :cfg.jy1 - 1], before[:, :cfg.jy1 - 1])
    assert orc.g.xmn == pytest.approx(x0 + nsh * cfg.dx, abs=1e-15)
    xmn = orc.g.xmn
    inside = part[:, 0] > xmn
    assert kept.shape[0] == int(inside.sum()) < part.shape[0]
    assert np.array_equal(kept, part[inside])
    # the second move: nint of the accumulated lapse minus what has moved
    nsh2, _ = orc.window_xshift(win, 10, 20, 1.0, ebf, kept.copy(order="F"))
    assert nsh + nsh2 == int(np.floor(cfg.dt * 20 / cfg.dx + 0.5))


def _entering_plasma(cfg, nsh, box_xmn, seed):
    """The plasma of the nsh columns that entered on the right, in the moved
    box frame (same sub-cell pattern as the initial plasma)."""
    p = uniform_plasma(cfg, (2, 2) if cfg.ndim == 2 else (2, 2, 2),
                       x_start_cell=cfg.nx - nsh, seed=seed, thermal=0.05)
    p[:, 0] += box_xmn - cfg.xmn
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-10), (PREC_F32, 1e-4)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_window_run_matches_oracle(dim, precision, tol):
    from aladyn_amd import LPStep

    cfg, part, ebf = {"2d": small_2d, "3d": small_3d}[dim]()
    w_sh, w_speed = 4, 1.0
    print(f"window {dim} precision={precision}", flush=True)
    orc = Oracle(cfg)
    win = OrcWindow()
    p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
    with LPStep(cfg, precision, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        assert gpu.lp_window_xshift(w_sh, 0, w_speed) == 0
        orc.window_xshift(win, w_sh, 0, w_speed, e, p)
        for it in range(3 * w_sh):
            p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
            gpu.lpf2_evolve(it, initial_time=(it == 0))
            if (it + 1) % w_sh == 0:
                nsh, p = orc.window_xshift(win, w_sh, it + 1, w_speed, e, p)
                assert gpu.lp_window_xshift(w_sh, it + 1, w_speed) == nsh > 0
                box = gpu.box()
                assert box[0] == pytest.approx(orc.g.xmn, abs=1e-12)
                new = _entering_plasma(cfg, nsh, orc.g.xmn, seed=it)
                p = np.asfortranarray(np.concatenate([p, new]))
                gpu.inject_species(0, new)
        pg, eg, jg = gpu.get_species(0), gpu.get_fields(), gpu.get_current()
    assert pg.shape == p.shape, (pg.shape, p.shape)
    D = cfg.curr_ndim
    for d in range(D):
        span = (cfg.xmx, cfg.ymx, cfg.zmx)[d] - (cfg.xmn, cfg.ymn, cfg.zmn)[d]
        assert np.max(np.abs(pg[:, d] - p[:, d])) / span <= tol
    assert rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) <= tol
    assert rel_maxerr(eg, e) <= tol
    if precision == PREC_F64:  # FP32 currents are compared by the one-step tests
        assert rel_maxerr(jg, jc) <= tol
