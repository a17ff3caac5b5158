"""The field and particle output files (pic_out.f90 fields_out and
part_pdata_out, one rank), written by aladyn_amd.outputs.

* The formatted header records against what a Fortran compiler prints for
  the same edit descriptors (tests/golden/fortran_formats.txt, written by
  make_formats.py from fortran_formats.f90 with amdflang).
* The .bin files byte for byte against files assembled here, loop by loop,
  from the reference's write statements (pic_out.f90:636-754, 1383-1491).
* GPU: the files of a run on the device hold the oracle's fields and
  particles of the same steps (float32 in the files).
"""
import struct
from pathlib import Path

import numpy as np
import pytest

from aladyn_amd.config import LPConfig
from aladyn_amd.inputs import uniform_plasma
from aladyn_amd.outputs import (IPAR, RPAR, data_out_gpu, fields_out, fortran_e, i_end,
                                part_pdata_out, read_fields_bin, read_part_bin)

GOLDEN = Path(__file__).resolve().parent / "golden"
VALUES = np.array([1.0, -2.5e-3, 0.0, 123456.78, 3.0e-30, 0.999995, 9.99995e5, -1.0e10, 7.2e-8,
                   1.0 / 3.0, 0.05, -0.2, 25.6, 1.0e-3, 2.81794033, 0.510998928, 11.1487, -6.4,
                   12345.0, 123445.0], dtype=np.float32)
INTS = [1, 1, 1, 256, 128, 128, 128, 128, 1, 0, 0, 1, 1, 1, 3, 8, 2, 2, 2, 1]


def test_header_formats_match_fortran():
    want = (GOLDEN / "fortran_formats.txt").read_text()
    got = "  Integer parameters\n"
    got += "".join("".join(f"{v:14d}" for v in INTS[i:i + 4]) + "\n" for i in range(0, 20, 4))
    got += "  Real parameters\n"
    got += "".join("".join(fortran_e(float(v), 14, 5) for v in VALUES[i:i + 4]) + "\n"
                   for i in range(0, 20, 4))
    got += "".join(" time =      " + fortran_e(float(v), 11, 4) + "\n" for v in VALUES)
    got += "".join(" npe =      " + f"{v:8d}\n" for v in INTS)
    got += "  Number of particles in the output box\n" + f"{33554432:20d}\n"
    assert got == want


def _cfg3():
    return LPConfig(ndim=3, nx=12, ny=8, nz=6, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.8,
                    n_over_nc=0.005)


def _records(data: bytes):
    out, o = [], 0
    while o < len(data):
        (n,) = struct.unpack_from("<i", data, o)
        out.append(data[o + 4:o + 4 + n])
        o += 8 + n
    return out


@pytest.mark.parametrize("ndim,jump", [(3, 1), (3, 2), (2, 1), (2, 3)])
def test_fields_out_bytes(tmp_path, ndim, jump):
    cfg = _cfg3() if ndim == 3 else LPConfig(ndim=2, nx=10, ny=7, k0=20.0, cfl=0.9,
                                              n_over_nc=0.01)
    rng = np.random.default_rng(1)
    ebf = np.asfortranarray(rng.normal(size=cfg.field_shape))
    run = {"w0_x": 2.0, "w0_y": 10.0, "a0": 3.0, "model_id": 1, "dmodel_id": 0,
           "mp_per_cell": [8], "iform": 0, "lpf_ord": 2}
    f_ind = 2
    path = fields_out(tmp_path, cfg, ebf, f_ind, f_ind, 7, tnow=1.25, xmin=0.5, jump=jump,
                      run=run)
    assert path == tmp_path / "0007" / "Eyfout07.bin"
    # wdata: do iz = kz1, nzp, jump; do iy = jy1, nyp, jump; do ix = ix1, nxp, jump
    ix1, nxp, jy1, nyp = 3, cfg.nx + 2, 3, cfg.ny + 2
    kz1, nzp = (3, cfg.nz + 2) if ndim == 3 else (1, 1)
    w = [np.float32(ebf[ix - 1, iy - 1, iz - 1, f_ind - 1])
         for iz in range(kz1, nzp + 1, jump) for iy in range(jy1, nyp + 1, jump)
         for ix in range(ix1, nxp + 1, jump)]
    gr = [len(range(ix1, nxp + 1, jump)), len(range(jy1, nyp + 1, jump)),
          len(range(kz1, nzp + 1, jump))]
    nz = cfg.nz if ndim == 3 else 1
    int_par = [1, 1, 1, gr[0], gr[1], cfg.ny, gr[2], nz, jump, 0, 0, 1, 0, 1,
               cfg.curr_ndim, 8, 2, 2, 2, i_end()]
    real_par = [1.25, 0.5, 0.5 + (cfg.nx - 1) * cfg.dx, cfg.ymn, cfg.ymx, cfg.zmn, cfg.zmx,
                2.0, 10.0, cfg.n_over_nc, 3.0, 1.0, 0.510998928, cfg.ompe, 0, 0, 0, 0, 0, 0]
    xs = [0.5 + cfg.dx * (i - 1) for i in range(1, cfg.nx + 1, jump)]
    ys = [cfg.dy * (i - 1 - cfg.ny // 2) for i in range(1, cfg.ny + 1, jump)]
    zs = [cfg.dz * (i - 1 - cfg.nz // 2) for i in range(1, cfg.nz + 1, jump)] if ndim == 3 \
        else [0.0]
    rec = lambda b: struct.pack("<i", len(b)) + b + struct.pack("<i", len(b))
    want = (rec(struct.pack("<i", 20)) + rec(struct.pack("<20i", *int_par))
            + rec(struct.pack("<20f", *real_par)) + rec(struct.pack("<3i", *gr))
            + rec(np.array(w, "<f4").tobytes()) + rec(np.array(xs, "<f4").tobytes())
            + rec(np.array(ys, "<f4").tobytes()) + rec(np.array(zs, "<f4").tobytes()))
    assert path.read_bytes() == want
    dat = (path.parent / "Eyfout07.dat").read_text().splitlines()
    assert dat[0] == "  Integer parameters" and dat[6] == "  Real parameters" and len(dat) == 12
    assert dat[1] == "".join(f"{v:14d}" for v in int_par[:4])
    assert dat[7] == "".join(fortran_e(float(np.float32(v)), 14, 5) for v in real_par[:4])
    ip, rp, fld, x, y, z = read_fields_bin(path)
    assert list(ip) == int_par and fld.shape == tuple(gr)
    assert np.array_equal(fld.ravel(order="F"), np.array(w, "<f4"))


@pytest.mark.parametrize("ndim,pjump", [(3, 1), (3, 3), (2, 2)])
def test_part_pdata_out_bytes(tmp_path, ndim, pjump):
    cfg = _cfg3() if ndim == 3 else LPConfig(ndim=2, nx=10, ny=7, k0=20.0, cfl=0.9,
                                              n_over_nc=0.01)
    part = uniform_plasma(cfg, (2, 2, 2)[:cfg.curr_ndim], seed=4, thermal=0.1, charge=1)
    D = cfg.curr_ndim
    xo0, xo1, yo = cfg.xmn + 2 * cfg.dx, cfg.xmn + 8 * cfg.dx, 1.5 * cfg.dy
    path = part_pdata_out(tmp_path, cfg, part, 2, 3, tnow=0.75, xmin=0.0, xmin_out=xo0,
                          xmax_out=xo1, ymax_out=yo, pjump=pjump, run={"mp_per_cell": [8, 8]})
    assert path == tmp_path / "0003" / "H1pout03.bin"
    rows = []
    for p in range(0, part.shape[0], pjump):   # do p = 1, np, jmp
        yy = part[p, 1]
        zz = part[p, 2] if ndim > 2 else 0.0
        if abs(yy) <= yo and abs(zz) <= yo and xo0 <= part[p, 0] <= xo1:
            b = struct.pack("<d", part[p, 2 * D])
            wgh = struct.unpack("<f", b[4:8])[0]
            charge = struct.unpack("<h", b[0:2])[0]
            rows.append([*part[p, :2 * D], wgh, float(charge)])
    want = struct.pack("<i", len(rows)) + np.array(rows, "<f4").tobytes()
    assert len(rows) > 0
    assert path.read_bytes() == want
    dat = (path.parent / "H1pout03.dat").read_text().splitlines()
    assert dat[0] == "  Real parameters" and dat[21] == "  Integer parameters"
    assert dat[1] == RPAR[0] + fortran_e(0.75, 11, 4)
    assert dat[22 + 13] == IPAR[13] + f"{2 * D + 2:8d}"      # ph_sp_nc = ndv
    assert dat[-1] == f"{len(rows):20d}"
    assert np.array_equal(read_part_bin(path, D), np.array(rows, "<f4"))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [64, 32])
def test_data_out_of_a_gpu_run(tmp_path, precision):
    from aladyn_amd import LPStep
    from cases import small_3d
    from oracle.pyoracle import Oracle

    cfg, part, ebf = small_3d()
    orc = Oracle(cfg)
    p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
    for it in range(3):
        p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
    print(f"data_out precision={precision}", flush=True)
    with LPStep(cfg, precision, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        for it in range(3):
            gpu.lpf2_evolve(it, initial_time=(it == 0))
        paths = data_out_gpu(gpu, tmp_path / "gpu", 1, tnow=3 * cfg.dt, nvout=cfg.nfield,
                             npout=1)
    assert len(paths) == cfg.nfield + 1
    for i in range(1, cfg.nfield + 1):
        ref = fields_out(tmp_path / "orc", cfg, e, i, i, 1, tnow=3 * cfg.dt, xmin=cfg.xmn)
        _, _, fg, *_ = read_fields_bin(paths[i - 1])
        _, _, fo, *_ = read_fields_bin(ref)
        tol = 1e-6 if precision == 64 else 1e-4
        assert np.max(np.abs(fg - fo)) <= tol * max(np.max(np.abs(fo)), 1e-30)
    ref = part_pdata_out(tmp_path / "orc", cfg, p, 1, 1, tnow=3 * cfg.dt, xmin=cfg.xmn,
                         xmin_out=cfg.xmn, xmax_out=cfg.xmn + cfg.nx * cfg.dx,
                         ymax_out=max(abs(cfg.ymn), abs(cfg.ymx)))
    pg, po = read_part_bin(paths[-1], cfg.curr_ndim), read_part_bin(ref, cfg.curr_ndim)
    assert pg.shape == po.shape
    assert np.max(np.abs(pg - po)) <= 1e-4 * np.max(np.abs(po))
