"""The restart dump (dumpRestart/, pic_dump.f90:35-795) on one rank.

CPU: the bytes of Comm-data.bin against a file assembled by hand from the
reference's write statements (pic_dump.f90:164-199: Fortran sequential
records with 4-byte markers, in the order rdata, ndata, nptx, sptx_max,
table columns when targ_end > xmax, npt_arr, dist_npy, dist_npz), and a
write/read round trip of every file.

GPU: a run that moves its window, dumps, and goes on, against a fresh
context restored from the dump that does the same after the dump.
"""
import struct

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64
from aladyn_amd.inputs import plasma_table
from aladyn_amd.restart import read_dump, write_dump
from cases import small_2d, small_3d
from conftest import rel_maxerr


def _rec(b):
    return struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))


def test_comm_data_bytes_follow_the_reference_write_order(tmp_path):
    cfg, part, ebf = small_2d()
    table = {"xpt": np.array([2.5, 2.6, 2.7]), "wghpt": np.array([0.5, 0.25, 1.0]),
             "ypt": np.array([0.1, 0.2]), "wghyz": np.ones((2, 1))}
    x0 = 3 * cfg.dx
    write_dump(tmp_path, cfg, iteration=40, time=1.5, parts=[part[:7]], ebf=ebf, x0=x0,
               nptx=[1], tables=[table], targ_in=0.0, targ_end=9.0, j0_norm=1.0)
    rdata = struct.pack("<10d", 1.5, 1.0, cfg.ompe, 0.0, 9.0, 0.0, 0.0, 0.0, x0, 0.0)
    ndata = struct.pack("<10i", 40, cfg.n1p, cfg.n2p, cfg.n3p, 3, cfg.nx + 2, 0, 4, 0, 0)
    expect = (_rec(rdata) + _rec(ndata) + _rec(struct.pack("<i", 1)) + _rec(struct.pack("<i", 3))
              + _rec(struct.pack("<dd", 2.5, 0.5)) + _rec(struct.pack("<dd", 2.6, 0.25))
              + _rec(struct.pack("<dd", 2.7, 1.0)) + _rec(struct.pack("<i", 7))
              + _rec(struct.pack("<i", 2)) + _rec(struct.pack("<i", 1)))
    assert (tmp_path / "dumpRestart" / "Comm-data.bin").read_bytes() == expect
    # Particles00: the species' columns one after the other; EB-fields00: ebf in Fortran order
    raw = np.fromfile(tmp_path / "dumpRestart" / "Particles00.bin", dtype="<f8")
    assert np.array_equal(raw, part[:7].ravel(order="F"))
    raw = np.fromfile(tmp_path / "dumpRestart" / "EB-fields00.bin", dtype="<f8")
    assert np.array_equal(raw, ebf.ravel(order="F"))
    raw = np.fromfile(tmp_path / "dumpRestart" / "Dist-wgyz.bin", dtype="<f8")
    assert np.array_equal(raw, [0.1, 0.2, 0.0, 1.0, 1.0])  # ypt, zpt (2-D: one), wghyz


@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_dump_round_trip(tmp_path, dim):
    cfg, part, ebf = {"2d": small_2d, "3d": small_3d}[dim]()
    ppc = (2, 2) if cfg.ndim == 2 else (2, 2, 2)
    t = plasma_table(cfg, ppc, x_start_cell=cfg.nx, x_end_cell=cfg.nx + 8)
    other = part[::3].copy(order="F")
    write_dump(tmp_path, cfg, iteration=7, time=0.25, parts=[part, other], ebf=ebf,
               x0=2 * cfg.dx, nptx=[0, 0], tables=[t, t], targ_in=0.0, targ_end=100.0)
    d = read_dump(tmp_path, cfg, 2)
    assert d["iteration"] == 7 and d["time"] == 0.25 and d["x0"] == 2 * cfg.dx
    assert np.array_equal(d["parts"][0], part) and np.array_equal(d["parts"][1], other)
    assert np.array_equal(d["ebf"], ebf)
    for k in ("xpt", "wghpt", "ypt", "wghyz") + (("zpt",) if cfg.ndim == 3 else ()):
        assert np.array_equal(np.asarray(d["tables"][1][k]).ravel(), np.asarray(t[k]).ravel()), k


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(PREC_F64, 1e-11), (PREC_F32, 1e-4)])
@pytest.mark.parametrize("dim", ["2d", "3d"])
def test_restart_continues_the_run(tmp_path, dim, precision, tol):
    from aladyn_amd import LPStep
    from aladyn_amd.restart import dump_gpu, restore_gpu

    cfg, part, ebf = {"2d": small_2d, "3d": small_3d}[dim]()
    ppc = (2, 2) if cfg.ndim == 2 else (2, 2, 2)
    table = plasma_table(cfg, ppc, x_start_cell=cfg.nx, x_end_cell=cfg.nx + 60)
    targ = (cfg.xmn, float(table["xpt"][-1]) + cfg.dx)
    w_sh, kdump, nsteps = 4, 8, 14
    print(f"restart {dim} precision={precision}", flush=True)

    def advance(gpu, it):
        gpu.lpf2_evolve(it, initial_time=(it == 0))
        if (it + 1) % w_sh == 0:
            assert gpu.lp_window_xshift(w_sh, it + 1, 1.0) > 0

    with LPStep(cfg, precision, track_ids=True) as g1:
        g1.put_species(0, part)
        g1.put_fields(ebf)
        g1.set_injection(0, table, nptx=0, targ_in=targ[0], targ_end=targ[1])
        for it in range(kdump):
            advance(g1, it)
        assert g1.box()[0] > cfg.xmn and g1.injection_state(0) > 0
        dump_gpu(tmp_path, g1, iteration=kdump, time=kdump * cfg.dt, tables=[table], targ=targ)
        # the restarted run starts with the window's saved state at zero
        g1.lp_window_xshift(w_sh, 0, 1.0)
        for it in range(kdump, nsteps):
            advance(g1, it)
        p1, e1, b1 = g1.get_species(0), g1.get_fields(), g1.box()
    with LPStep(cfg, precision, track_ids=True) as g2:
        it0, _ = restore_gpu(tmp_path, g2)
        assert it0 == kdump
        for it in range(kdump, nsteps):
            advance(g2, it)
        p2, e2, b2 = g2.get_species(0), g2.get_fields(), g2.box()
    # the restored box is x0 + nx*dx, the running one a sum of moves (the
    # reference's restart has the same last-bit difference)
    assert np.allclose(b1, b2, rtol=0.0, atol=1e-12)
    assert p1.shape == p2.shape
    D = cfg.curr_ndim
    for d in range(D):
        span = (cfg.xmx, cfg.ymx, cfg.zmx)[d] - (cfg.xmn, cfg.ymn, cfg.zmn)[d]
        assert np.max(np.abs(p1[:, d] - p2[:, d])) / span <= tol
    assert rel_maxerr(p2[:, D:2 * D], p1[:, D:2 * D]) <= tol
    assert rel_maxerr(e2, e1) <= tol
