"""The reference's restart dump (dumpRestart/, pic_dump.f90:35-795) for one
rank, so that a run driven from Python can write a dump the reference
resumes from, and resume from one the reference wrote.

Files (single rank: npe = npe_yloc = npe_zloc = 1, imody = imodz = 0):

* ``Comm-data.bin`` -- Fortran unformatted sequential records (a 4-byte
  length before and after each record), written by rank 0
  (pic_dump.f90:164-199): rdata(1:10) f64 = (t, j0_norm, ompe, targ_in,
  targ_end, lp_in(1), xp0_out, xp1_out, x(1), 0); ndata(1:10) int32 = (iter,
  n1p, n2p, n3p, nptx_max, size(x) = nx+2, nxf, nd2, 0, 0); nptx(1:nsp);
  sptx_max(1:nsp); when targ_end > xmax, one record (xpt(j,ic), wghpt(j,ic))
  per column j = 1..nptx_max of each species; npt_arr(1:npe, 1:nsp) (the
  particle counts); dist_npy(1, 1:nsp); dist_npz(1, 1:nsp).
* ``Particles00.bin`` -- raw f64: each species' part(1:np, 1:2D+1), column
  by column, species after species (pic_dump.f90:205-229).
* ``EB-fields00.bin`` -- raw f64: ebf(n1p, n2p, n3p, nfield) in Fortran
  order (pic_dump.f90:231-251).
* ``Dist-wgyz.bin`` -- raw f64: loc_ypt(1:npty(ic), ic) for each species,
  then loc_zpt(1:nptz(ic), ic), then loc_wghyz(1:npty, 1:nptz, ic)
  (pic_dump.f90:343-374).

The envelope, fluid and bunch files of other models are not written.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

FOLDER = "dumpRestart"


def _rec(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _read_rec(f) -> bytes:
    head = f.read(4)
    if len(head) != 4:
        raise EOFError("truncated Fortran record")
    (n,) = struct.unpack("<i", head)
    payload = f.read(n)
    (m,) = struct.unpack("<i", f.read(4))
    if m != n or len(payload) != n:
        raise ValueError("corrupt Fortran record markers")
    return payload


def write_dump(root, cfg, *, iteration: int, time: float, parts, ebf, x0: float = 0.0,
               nptx=None, tables=None, targ_in: float = 0.0, targ_end: float = 0.0,
               j0_norm: float = 1.0, lp_in: float = 0.0, xp0_out: float = 0.0,
               xp1_out: float = 0.0):
    """Writes root/dumpRestart/ as dump_data does on one rank.

    parts: one part(np, 2D+1) array per species; ebf: the fields; x0: the
    box's x(1) (0 before the window moved); nptx and tables (dicts with xpt,
    wghpt, ypt, zpt, wghyz, per species) describe the injection target; the
    table columns are written only when targ_end > xmax, as the reference
    does."""
    d = Path(root) / FOLDER
    d.mkdir(parents=True, exist_ok=True)
    nsp = len(parts)
    D = cfg.curr_ndim
    tables = tables or [None] * nsp
    nptx = [0] * nsp if nptx is None else list(nptx)
    sptx_max = [0 if t is None else len(t["xpt"]) for t in tables]
    nptx_max = max(sptx_max) if sptx_max else 0
    xmax = x0 + cfg.nx * cfg.dx
    rdata = np.zeros(10)
    rdata[:9] = [time, j0_norm, cfg.ompe, targ_in, targ_end, lp_in, xp0_out, xp1_out, x0]
    ndata = np.zeros(10, dtype="<i4")
    ndata[:8] = [iteration, cfg.n1p, cfg.n2p, cfg.n3p, nptx_max, cfg.nx + 2, 0, 2 * D]
    with open(d / "Comm-data.bin", "wb") as f:
        _rec(f, rdata.astype("<f8").tobytes())
        _rec(f, ndata.tobytes())
        _rec(f, np.asarray(nptx, dtype="<i4").tobytes())
        _rec(f, np.asarray(sptx_max, dtype="<i4").tobytes())
        if targ_end > xmax:
            for t in tables:
                xpt = np.zeros(nptx_max) if t is None else np.asarray(t["xpt"], dtype=np.float64)
                wpt = np.zeros(nptx_max) if t is None else np.asarray(t["wghpt"], dtype=np.float64)
                for j in range(nptx_max):
                    x = xpt[j] if j < xpt.size else 0.0
                    w = wpt[j] if j < wpt.size else 0.0
                    _rec(f, struct.pack("<dd", x, w))
        _rec(f, np.asarray([p.shape[0] for p in parts], dtype="<i4").tobytes())
        npy = [0 if t is None else len(t["ypt"]) for t in tables]
        npz = [0 if t is None else (len(t["zpt"]) if D == 3 else 1) for t in tables]
        _rec(f, np.asarray(npy, dtype="<i4").tobytes())
        _rec(f, np.asarray(npz, dtype="<i4").tobytes())
    with open(d / "Particles00.bin", "wb") as f:
        for p in parts:
            f.write(np.asfortranarray(p, dtype="<f8").tobytes(order="F"))
    with open(d / "EB-fields00.bin", "wb") as f:
        f.write(np.asfortranarray(ebf, dtype="<f8").tobytes(order="F"))
    if any(t is not None for t in tables):
        with open(d / "Dist-wgyz.bin", "wb") as f:
            for t in tables:
                if t is not None:
                    f.write(np.asarray(t["ypt"], dtype="<f8").tobytes())
            for t in tables:
                if t is not None:
                    f.write(np.asarray(t["zpt"] if D == 3 else [0.0], dtype="<f8").tobytes())
            for t in tables:
                if t is not None:
                    w = np.asarray(t["wghyz"], dtype="<f8").reshape(len(t["ypt"]), -1, order="F")
                    f.write(w.tobytes(order="F"))


def read_dump(root, cfg, nsp: int) -> dict:
    """Reads root/dumpRestart/ as restart does on one rank.  Returns
    iteration, time, rdata, ndata, nptx, sptx_max, x0 (= x(1)), the
    particles of each species, ebf, and, when the dump holds them, the
    target tables (xpt, wghpt, ypt, zpt, wghyz per species)."""
    d = Path(root) / FOLDER
    D = cfg.curr_ndim
    with open(d / "Comm-data.bin", "rb") as f:
        rdata = np.frombuffer(_read_rec(f), dtype="<f8").copy()
        ndata = np.frombuffer(_read_rec(f), dtype="<i4").copy()
        nptx = np.frombuffer(_read_rec(f), dtype="<i4").copy()
        sptx_max = np.frombuffer(_read_rec(f), dtype="<i4").copy()
        nptx_max = int(ndata[4])
        x0 = float(rdata[8])
        targ_end = float(rdata[4])
        xpt = wghpt = None
        # the reference reads the table when targ_end > xmax + x(1), xmax its
        # own box edge before the shift (pic_dump.f90:477)
        if targ_end > cfg.nx * cfg.dx + x0:
            xpt = np.zeros((nptx_max, nsp))
            wghpt = np.zeros((nptx_max, nsp))
            for i in range(nsp):
                for j in range(nptx_max):
                    xpt[j, i], wghpt[j, i] = struct.unpack("<dd", _read_rec(f))
        npt = np.frombuffer(_read_rec(f), dtype="<i4").copy()
        npy = np.frombuffer(_read_rec(f), dtype="<i4").copy()
        npz = np.frombuffer(_read_rec(f), dtype="<i4").copy()
    if int(ndata[1]) != cfg.n1p or int(ndata[2]) != cfg.n2p or int(ndata[3]) != cfg.n3p:
        raise ValueError("dump grid differs from the configuration")
    ndv = int(ndata[7]) + 1
    raw = np.fromfile(d / "Particles00.bin", dtype="<f8") if npt.sum() > 0 else np.zeros(0)
    parts, k = [], 0
    for ic in range(nsp):
        n = int(npt[ic])
        parts.append(np.asfortranarray(raw[k:k + n * ndv].reshape(n, ndv, order="F")))
        k += n * ndv
    ebf = np.fromfile(d / "EB-fields00.bin", dtype="<f8").reshape(cfg.field_shape, order="F")
    out = dict(iteration=int(ndata[0]), time=float(rdata[0]), rdata=rdata, ndata=ndata,
               nptx=nptx, sptx_max=sptx_max, x0=x0, parts=parts, ebf=np.asfortranarray(ebf),
               tables=None)
    if xpt is not None and (d / "Dist-wgyz.bin").exists():
        raw = np.fromfile(d / "Dist-wgyz.bin", dtype="<f8")
        k, tables = 0, []
        ys, zs = [], []
        for ic in range(nsp):
            ys.append(raw[k:k + npy[ic]]); k += int(npy[ic])
        for ic in range(nsp):
            zs.append(raw[k:k + npz[ic]]); k += int(npz[ic])
        for ic in range(nsp):
            m = int(npy[ic]) * int(npz[ic])
            w = raw[k:k + m].reshape(int(npy[ic]), int(npz[ic]), order="F"); k += m
            t = {"xpt": xpt[:int(sptx_max[ic]), ic].copy(), "wghpt": wghpt[:int(sptx_max[ic]), ic].copy(),
                 "ypt": ys[ic].copy(), "wghyz": np.asfortranarray(w)}
            if D == 3:
                t["zpt"] = zs[ic].copy()
            tables.append(t)
        out["tables"] = tables
    return out


def dump_gpu(root, gpu, *, iteration: int, time: float, tables=None, targ=(0.0, 0.0), **kw):
    """dump_data from a running LPStep: the device state is copied back
    (get_species / get_fields) and written with the box origin and the
    injection state of the context."""
    cfg = gpu.cfg
    parts = [gpu.get_species(ic) for ic in range(cfg.nsp)]
    nptx = [gpu.injection_state(ic) for ic in range(cfg.nsp)]
    write_dump(root, cfg, iteration=iteration, time=time, parts=parts, ebf=gpu.get_fields(),
               x0=float(gpu.box()[0]), nptx=nptx, tables=tables, targ_in=targ[0],
               targ_end=targ[1], **kw)


def restore_gpu(root, gpu):
    """restart into a fresh LPStep: box origin, fields, particles and the
    injection tables.  Returns the dump's (iteration, time)."""
    cfg = gpu.cfg
    d = read_dump(root, cfg, cfg.nsp)
    gpu.set_box_origin(d["x0"])
    gpu.put_fields(d["ebf"])
    for ic, p in enumerate(d["parts"]):
        gpu.put_species(ic, p)
    if d["tables"] is not None:
        targ_in, targ_end = float(d["rdata"][3]), float(d["rdata"][4])
        for ic, t in enumerate(d["tables"]):
            if len(t["xpt"]):
                gpu.set_injection(ic, t, nptx=int(d["nptx"][ic]), targ_in=targ_in,
                                  targ_end=targ_end)
    return d["iteration"], d["time"]
