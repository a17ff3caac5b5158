"""The reference's field and particle output files (pic_out.f90) for one
rank, written from arrays the library returns (aladyn_get_fields,
aladyn_get_species), so that a run driven from Python emits the files the
reference's post-processing reads.

Output folder ``{iout:04d}/`` (pic_out.f90:627, 1387), one rank
(npe = npe_xloc = npe_yloc = npe_zloc = 1), l_force_singlefile_output (the
default, read_input.f90:163):

The header's xmax is the grid's x(nx) = xmin + (nx-1)*dx
(set_grid_param.f90:105-107), ymin, ymax = y(1), y(ny+1) and zmin, zmax =
z(1), z(nz+1) (0, 0 in 2-D).

* fields_out (pic_out.f90:615-755, file_version 2), called as
  ``fields_out(ebf, i, i)`` for i = 1..nvout (ALaDyn.F90:183-186):
  ``{Name}{iout:02d}.dat`` -- ' Integer parameters', int_par(1:20) in
  (4i14), ' Real parameters', real_par(1:20) in (4e14.5);
  ``{Name}{iout:02d}.bin`` -- Fortran unformatted sequential records (a
  4-byte length before and after each): par_dim; int_par (int32); real_par
  (float32); gr_dim(3); the interior nodes ix1..nxp, jy1..nyp, kz1..nzp
  with step jump, x fastest (float32); x(1:nx:jump), y(1:ny:jump),
  z(1:nz:jump) (float32).
* part_pdata_out (pic_out.f90:1356-1499, file_version 4):
  ``{Elpout|H1pout|Prpout|H2pout}{iout:02d}.dat`` -- the 20 real and 20
  integer parameters with their labels, (a13,e11.4) and (a12,i8), and the
  particle count (4i20); ``.bin`` written with MPI-IO in native order (no
  record markers): the count (int32), then per particle the 2*curr_ndim
  coordinates and momenta, the FP32 weight and the charge (float32 each).
  Particles are taken every pjump-th, in array order, inside
  xmin_out <= x <= xmax_out and |y|, |z| <= ymax_out.

Run parameters the library does not hold (laser and target description,
model ids, ...) come from a dict ``run``; missing keys are written as 0.
"""
from __future__ import annotations

import struct
import sys
from decimal import ROUND_HALF_EVEN, Decimal
from pathlib import Path

import numpy as np

from .inputs import unpack_weight

PAR_DIM = 20
ELECTRON_MASS_MEV = 0.510998928       # phys_param.f90:47 (e0)
PROTON_MASS_NORM = 1836.1527706       # phys_param.f90:37
PART_FILES = ("Elpout", "H1pout", "Prpout", "H2pout")  # pic_out.f90:1359-1360
RPAR = (" time =      ", " xmin =      ", " xmax =      ", " ymin =      ", " ymax =      ",
        " zmin =      ", " zmax =      ", " w0_x =      ", " w0_y =      ", " a0 =        ",
        " lam0 =      ", " mc2(MeV) =  ", " n0(e18) =   ", " np/cell =   ", " weight =    ",
        " mass =      ", " xmin_out =  ", " xmax_out =  ", " ymax_out =  ", " gam_min =   ")
IPAR = (" npe =      ", " nx =       ", " ny =       ", " nz =       ", " model =    ",
        " dmodel =   ", " nsp =      ", " curr_ndim =", " mp/cell =  ", " ion_ch =   ",
        " tsch_ord = ", " der_ord =  ", " iform =    ", " ph_sp_nc = ", " f_version =",
        " i_end =    ", " nx_loc =   ", " ny_loc =   ", " nz_loc =   ", " pjump  =   ")


# --------------------------------------------------------------- formatting
def fortran_e(x: float, w: int, d: int) -> str:
    """The Ew.d edit descriptor: [-]0.ddd...E+xx, right-justified in w,
    the exact binary value rounded to d digits (ties to even)."""
    x = float(x)
    if x == 0.0:
        body, e10, neg = "0." + "0" * d, 0, False
    else:
        v = Decimal(x).copy_abs()
        e10 = v.adjusted() + 1
        m = v.scaleb(-e10).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN)
        if m >= 1:
            m, e10 = m.scaleb(-1).quantize(Decimal(1).scaleb(-d), rounding=ROUND_HALF_EVEN), e10 + 1
        body, neg = "0." + str(m)[2:], x < 0
    s = ("-" if neg else "") + body + ("E-" if e10 < 0 else "E+") + f"{abs(e10):02d}"
    return s.rjust(w)


def _ints(values, per: int, w: int) -> str:
    vals = [int(v) for v in values]
    return "".join("".join(f"{v:{w}d}" for v in vals[i:i + per]) + "\n"
                   for i in range(0, len(vals), per))


def _reals(values, per: int, w: int, d: int) -> str:
    vals = [float(np.float32(v)) for v in values]
    return "".join("".join(fortran_e(v, w, d) for v in vals[i:i + per]) + "\n"
                   for i in range(0, len(vals), per))


def _list_directed(text: str) -> str:
    """write (u, *) 'text': a blank, then the characters."""
    return " " + text + "\n"


def _rec(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def i_end() -> int:
    """endian (pic_out.f90:53-65): 1 on a little-endian host, 2 otherwise."""
    return 1 if sys.byteorder == "little" else 2


# ----------------------------------------------------------------- fields
def field_name(cfg, f_var: int, iout: int) -> str:
    """File name of fields_out(ebf, f_var, f_var) (pic_out.f90:665-684)."""
    names = {0: "Jxfout", 1: "Exfout", 2: "Eyfout", 3: "Bzfout" if cfg.nfield == 3 else "Ezfout",
             4: "Bxfout", 5: "Byfout", 6: "Bzfout"}
    return f"{names[f_var]}{iout:02d}"


def fields_out(root, cfg, ef, f_ind: int, f_var: int, iout: int, *, tnow: float, xmin: float,
               jump: int = 1, run: dict | None = None) -> Path:
    """fields_out(ef, f_ind, f_var) on one rank: ef(n1p, n2p, n3p, :) is ebf
    (or jc for f_var 0), f_ind is 1-based; xmin is the box's x (it moves
    with the window).  Returns the .bin path."""
    run = run or {}
    ef = np.asarray(ef)
    nz = cfg.nz if cfg.ndim == 3 else 1
    # interior nodes ix1..nxp, jy1..nyp, kz1..nzp (0-based from 2)
    k0 = 2 if cfg.ndim == 3 else 0
    block = ef[2:2 + cfg.nx:jump, 2:2 + cfg.ny:jump, k0:k0 + nz:jump, f_ind - 1]
    gr_dim = block.shape
    wdata = np.asarray(block, dtype="<f4").ravel(order="F")
    xmax = xmin + (cfg.nx - 1) * cfg.dx  # x(nx) (set_grid_param.f90:105-107)
    real_par = [tnow, xmin, xmax, cfg.ymn, cfg.ymx, cfg.zmn, cfg.zmx,
                run.get("w0_x", 0.0), run.get("w0_y", 0.0), cfg.n_over_nc, run.get("a0", 0.0),
                cfg.lam0, ELECTRON_MASS_MEV, cfg.ompe, run.get("targ_in", 0.0),
                run.get("targ_end", 0.0), run.get("gam0", 0.0), run.get("nb_over_np", 0.0),
                run.get("b_charge", 0.0), run.get("vbeam", 0.0)]
    mp = run.get("mp_per_cell", [0])
    int_par = [1, 1, 1, gr_dim[0], gr_dim[1], cfg.ny, gr_dim[2], nz, jump, cfg.iby,
               run.get("iform", 0), run.get("model_id", 0), run.get("dmodel_id", 0), cfg.nsp,
               cfg.curr_ndim, mp[0], run.get("lpf_ord", 2), cfg.der_ord, 2, i_end()]
    folder = Path(root) / f"{iout:04d}"
    folder.mkdir(parents=True, exist_ok=True)
    name = field_name(cfg, f_var, iout)
    (folder / f"{name}.dat").write_text(
        _list_directed(" Integer parameters") + _ints(int_par, 4, 14)
        + _list_directed(" Real parameters") + _reals(real_par, 4, 14, 5))
    x = (xmin + cfg.dx * np.arange(cfg.nx))[::jump]
    y = (cfg.dy * (np.arange(cfg.ny) - cfg.ny // 2))[::jump]
    z = (cfg.dz * (np.arange(nz) - nz // 2))[::jump] if cfg.ndim == 3 else np.zeros(1)
    out = folder / f"{name}.bin"
    with open(out, "wb") as f:
        f.write(_rec(struct.pack("<i", PAR_DIM)))
        f.write(_rec(np.asarray(int_par, dtype="<i4").tobytes()))
        f.write(_rec(np.asarray(real_par, dtype="<f4").tobytes()))
        f.write(_rec(np.asarray(gr_dim, dtype="<i4").tobytes()))
        f.write(_rec(wdata.tobytes()))
        for c in (x, y, z):
            f.write(_rec(np.asarray(c, dtype="<f4").tobytes()))
    return out


def read_fields_bin(path):
    """(int_par, real_par, field(gr_dim, order F), x, y, z) of a fields_out .bin."""
    data = Path(path).read_bytes()
    recs, o = [], 0
    while o < len(data):
        (n,) = struct.unpack_from("<i", data, o)
        recs.append(data[o + 4:o + 4 + n])
        (m,) = struct.unpack_from("<i", data, o + 4 + n)
        if m != n:
            raise ValueError("corrupt Fortran record markers")
        o += 8 + n
    int_par = np.frombuffer(recs[1], "<i4")
    real_par = np.frombuffer(recs[2], "<f4")
    gr = tuple(np.frombuffer(recs[3], "<i4"))
    field = np.frombuffer(recs[4], "<f4").reshape(gr, order="F")
    return int_par, real_par, field, *(np.frombuffer(r, "<f4") for r in recs[5:8])


# -------------------------------------------------------------- particles
def part_pdata_out(root, cfg, part, pid: int, iout: int, *, tnow: float, xmin: float,
                   xmin_out: float, xmax_out: float, ymax_out: float, pjump: int = 1,
                   run: dict | None = None) -> Path:
    """part_pdata_out for species pid (1-based) on one rank, from
    part(np, 2D+1) in the reference's layout.  Returns the .bin path."""
    run = run or {}
    D = cfg.curr_ndim
    part = np.asarray(part)
    p = part[::pjump]
    sel = (np.abs(p[:, 1]) <= ymax_out) & (p[:, 0] >= xmin_out) & (p[:, 0] <= xmax_out)
    if cfg.ndim > 2:
        sel &= np.abs(p[:, 2]) <= ymax_out
    p = p[sel]
    ndv = 2 * D + 2
    charge, _, wgh = unpack_weight(p[:, 2 * D]) if p.shape[0] else (np.zeros(0), None, np.zeros(0))
    pdata = np.empty((p.shape[0], ndv), dtype="<f4")
    pdata[:, :2 * D] = p[:, :2 * D]
    pdata[:, 2 * D] = wgh
    pdata[:, 2 * D + 1] = np.asarray(charge, dtype=np.float32)
    nz = cfg.nz if cfg.ndim == 3 else 1
    mp = run.get("mp_per_cell", [0] * 4)
    mass = run.get("mass", [1.0, PROTON_MASS_NORM, PROTON_MASS_NORM, PROTON_MASS_NORM])
    real_par = [tnow, xmin, xmin + (cfg.nx - 1) * cfg.dx, cfg.ymn, cfg.ymx, cfg.zmn, cfg.zmx,
                run.get("w0_x", 0.0), run.get("w0_y", 0.0), run.get("a0", 0.0), cfg.lam0,
                ELECTRON_MASS_MEV, cfg.n0_ref, run.get("np_per_cell", 0.0),
                run.get("j0_norm", 0.0), mass[pid - 1], xmin_out, xmax_out, ymax_out,
                run.get("gam_min", 0.0)]
    int_par = [1, cfg.nx, cfg.ny, nz, run.get("model_id", 0), run.get("dmodel_id", 0), cfg.nsp,
               D, mp[pid - 1], run.get("ion_min", [0])[0], run.get("lpf_ord", 2), cfg.der_ord,
               run.get("iform", 0), ndv, 4, i_end(), cfg.nx, cfg.ny, nz, pjump]
    folder = Path(root) / f"{iout:04d}"
    folder.mkdir(parents=True, exist_ok=True)
    name = f"{PART_FILES[pid - 1]}{iout:02d}"
    text = _list_directed(" Real parameters")
    for lab, v in zip(RPAR, real_par):
        text += lab + fortran_e(float(np.float32(v)), 11, 4) + "\n"
    text += _list_directed(" Integer parameters")
    for lab, v in zip(IPAR, int_par):
        text += f"{lab}{int(v):8d}\n"
    text += _list_directed(" Number of particles in the output box")
    text += f"{p.shape[0]:20d}\n"
    (folder / f"{name}.dat").write_text(text)
    out = folder / f"{name}.bin"
    with open(out, "wb") as f:
        f.write(struct.pack("<i", p.shape[0]))
        f.write(pdata.tobytes())
    return out


def read_part_bin(path, curr_ndim: int) -> np.ndarray:
    """pdata(np, 2*curr_ndim + 2) of a part_pdata_out .bin (one rank)."""
    data = Path(path).read_bytes()
    (n,) = struct.unpack_from("<i", data, 0)
    return np.frombuffer(data, "<f4", offset=4).reshape(n, 2 * curr_ndim + 2)


# ------------------------------------------------------------ from a run
def data_out_gpu(gpu, root, iout: int, *, tnow: float, nvout: int = 0, npout: int = 0,
                 xmin_out: float | None = None, xmax_out: float | None = None,
                 ymax_out: float | None = None, jump: int = 1, pjump: int = 1,
                 run: dict | None = None):
    """Data_out's field and particle files (ALaDyn.F90:165-213) from an
    LPStep: fields_out(ebf, i, i) for i = 1..nvout; part_pdata_out for
    species npout (1-based) or, when npout > nsp, for every species.
    Returns the paths written."""
    cfg = gpu.cfg
    xmin = float(gpu.box()[0])
    paths = []
    if nvout > 0:
        ebf = gpu.get_fields()
        for i in range(1, nvout + 1):
            paths.append(fields_out(root, cfg, ebf, i, i, iout, tnow=tnow, xmin=xmin, jump=jump,
                                    run=run))
    if npout > 0:
        xo0 = xmin if xmin_out is None else xmin_out
        xo1 = xmin + cfg.nx * cfg.dx if xmax_out is None else xmax_out
        yo = max(abs(cfg.ymn), abs(cfg.ymx)) if ymax_out is None else ymax_out
        species = [npout] if npout <= cfg.nsp else range(1, cfg.nsp + 1)
        for pid in species:
            paths.append(part_pdata_out(root, cfg, gpu.get_species(pid - 1), pid, iout,
                                        tnow=tnow, xmin=xmin, xmin_out=xo0, xmax_out=xo1,
                                        ymax_out=yo, pjump=pjump, run=run))
    return paths
