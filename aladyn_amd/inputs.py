"""Seeded synthetic inputs for the LP step: plasma particles and laser fields.

The reference builds its initial state in part_distribute / lp_pulse
(src/IO/init_part_distrib.f90:2475-2553, src/IO/init_laser_field.f90:34-103),
placing particles at sub-cell centres and drawing thermal momenta from an
unseeded generator.  Parity runs need identical seeded inputs on every path,
so the state is generated here instead (numpy PCG64 from a given seed) in the
reference's layouts:

    part(np, 2*curr_ndim+1) float64, column-major: x, y[, z], px, py[, pz], W
    ebf(n1p, n2p, n3p, nfield) float64, column-major

W packs int16 charge, int16 part_ind and the FP32 weight as in
src/work/precision_def.F90:33-40.
"""
from __future__ import annotations

import math

import numpy as np

from .config import PI, LPConfig


def pack_weight(charge: int, part_ind: int, wgh) -> np.ndarray:
    """Packs (int16 charge, int16 part_ind, float32 wgh) into float64 slots."""
    wgh = np.asarray(wgh, dtype=np.float32)
    rec = np.zeros(wgh.shape, dtype=[("c", "<i2"), ("p", "<i2"), ("w", "<f4")])
    rec["c"] = charge
    rec["p"] = part_ind
    rec["w"] = wgh
    return rec.view("<f8").reshape(wgh.shape)


def unpack_weight(w: np.ndarray):
    rec = np.ascontiguousarray(w, dtype="<f8").view([("c", "<i2"), ("p", "<i2"), ("w", "<f4")])
    return rec["c"].astype(np.int32), rec["p"].astype(np.int32), rec["w"].astype(np.float32)


def uniform_plasma(cfg: LPConfig, ppc, *, x_start_cell: float = 0.0, x_end_cell: float | None = None,
                   y_cells=None, z_cells=None, seed: int = 0, jitter: bool = True,
                   charge: int = -1, n_rel: float = 1.0, thermal: float = 0.0,
                   ramp_cells: float = 0.0) -> np.ndarray:
    """Electrons at n0_ref*n_rel on a ppc sub-grid per cell.

    Each particle sits at its sub-cell centre (init_part_distrib.f90:513-516)
    plus, if jitter, a uniform U(-1/2, 1/2) displacement in sub-cell units.
    The weight is n_rel*n0_ref/ppc (init_part_distrib.f90:501-503); with
    ramp_cells > 0 the weight rises linearly over the first ramp_cells cells.
    thermal > 0 draws momenta from N(0, thermal) per component.
    Returns part(np, 2D+1), column-major.
    """
    rng = np.random.default_rng(seed)
    D = cfg.curr_ndim
    ppc = tuple(int(v) for v in ppc)[:D]
    x1 = cfg.nx if x_end_cell is None else x_end_cell
    ylo, yhi = (0, cfg.ny) if y_cells is None else y_cells
    cells_x = np.arange(int(math.ceil(x_start_cell)), int(x1))
    cells_y = np.arange(ylo, yhi)
    axes_cells = [cells_x, cells_y]
    if D == 3:
        zlo, zhi = (0, cfg.nz) if z_cells is None else z_cells
        axes_cells.append(np.arange(zlo, zhi))
    # sub-cell centres in cell units along each axis, local coordinate
    axes = []
    for d in range(D):
        sub = (np.arange(ppc[d]) + 0.5) / ppc[d]
        axes.append((axes_cells[d][:, None] + sub[None, :]).ravel())
    grids = np.meshgrid(*axes, indexing="ij")
    loc = [g.ravel(order="F") for g in grids]  # x fastest: cell-ordered
    n = loc[0].size
    mins = (cfg.xmn, cfg.ymn, cfg.zmn)
    sizes = (cfg.dx, cfg.dy, cfg.dz)
    part = np.zeros((n, cfg.ncol), dtype=np.float64, order="F")
    for d in range(D):
        u = loc[d]
        if jitter:
            u = u + (rng.random(n) - 0.5) / ppc[d]
        # keep strictly inside (0, n] in local units
        top = (cfg.nx, cfg.ny, cfg.nz)[d]
        u = np.clip(u, 1e-6, top - 1e-6)
        part[:, d] = mins[d] + u * sizes[d]
    if thermal > 0:
        for d in range(D):
            part[:, D + d] = rng.normal(0.0, thermal, n)
    wgh = np.full(n, n_rel * cfg.n0_ref / float(np.prod(ppc)), dtype=np.float64)
    if ramp_cells > 0:
        xl = (part[:, 0] - cfg.xmn) / cfg.dx - x_start_cell
        wgh *= np.clip(xl / ramp_cells, 0.0, 1.0)
    part[:, 2 * D] = pack_weight(charge, 0, wgh.astype(np.float32))
    return part


def plasma_table(cfg: LPConfig, ppc, *, x_start_cell: float = 0.0, x_end_cell: float | None = None,
                 n_rel: float = 1.0) -> dict:
    """Target table of a uniform plasma for injection at window moves: the
    arrays particles_inject / add_particles read (window.f90:39-139), with
    particles at sub-cell centres (init_part_distrib.f90:513-516).

    xpt: x of the plasma columns from cell x_start_cell to x_end_cell (cells
    of the initial box, which may lie beyond it), ppc[0] per cell; ypt, zpt:
    the transverse sub-cell centres of the box; wghpt = n_rel*n0_ref/ppc per
    column and wghyz = 1 (the weight of add_particles is their product).
    The particles of the columns inside the initial box are those of
    uniform_plasma(..., jitter=False) with the same arguments.
    """
    D = cfg.curr_ndim
    ppc = tuple(int(v) for v in ppc)[:D]
    x1 = cfg.nx if x_end_cell is None else x_end_cell
    cells_x = np.arange(int(math.ceil(x_start_cell)), int(math.ceil(x1)))
    sub = lambda n: (np.arange(n) + 0.5) / n
    xpt = cfg.xmn + (cells_x[:, None] + sub(ppc[0])[None, :]).ravel() * cfg.dx
    ypt = cfg.ymn + (np.arange(cfg.ny)[:, None] + sub(ppc[1])[None, :]).ravel() * cfg.dy
    table = {"xpt": xpt, "ypt": ypt,
             "wghpt": np.full(xpt.size, n_rel * cfg.n0_ref / float(np.prod(ppc)))}
    if D == 3:
        table["zpt"] = cfg.zmn + (np.arange(cfg.nz)[:, None] + sub(ppc[2])[None, :]).ravel() * cfg.dz
        table["wghyz"] = np.ones((ypt.size, table["zpt"].size), order="F")
    else:
        table["wghyz"] = np.ones((ypt.size, 1), order="F")
    return table


def gaussian_laser(cfg: LPConfig, *, a0: float, w0: float, tau_fwhm: float, xc: float,
                   lam0: float | None = None, polarization: str = "y") -> np.ndarray:
    """Linearly polarised Gaussian pulse, paraxial leading order, at its focus.

    E_y = B_z = a0*omega0 * exp(-r^2/w0^2) * exp(-(x-xc)^2/L^2) * cos(k0 (x-xc))
    with the envelope length L from the intensity FWHM tau_fwhm (laser
    periods), plus the longitudinal E_x that keeps div E = 0 to leading order.
    Staggering follows the Yee layout of init_grid_fields.f90:1025-1029:
    Ex(i+1/2,j,k), Ey(i,j+1/2,k), Ez(i,j,k+1/2), Bx(i,j+1/2,k+1/2),
    By(i+1/2,j,k+1/2), Bz(i+1/2,j+1/2,k).  xc, w0 in micron.
    """
    lam = cfg.lam0 if lam0 is None else lam0
    om = 2.0 * PI / lam
    amp = a0 * om
    tau = tau_fwhm * lam  # FWHM of the intensity in micron (c = 1)
    sig = tau / math.sqrt(2.0 * math.log(2.0))  # exp(-(x/sig)^2) envelope of E
    ebf = np.zeros(cfg.field_shape, dtype=np.float64, order="F")

    def coord(n, mn, d, half):
        idx = np.arange(1, n + 1, dtype=np.float64)
        return mn + (idx - 3.0 + (0.5 if half else 0.0)) * d

    x_i = coord(cfg.n1p, cfg.xmn, cfg.dx, False)
    x_h = coord(cfg.n1p, cfg.xmn, cfg.dx, True)
    y_i = coord(cfg.n2p, cfg.ymn, cfg.dy, False)
    y_h = coord(cfg.n2p, cfg.ymn, cfg.dy, True)
    if cfg.ndim == 3:
        z_i = coord(cfg.n3p, cfg.zmn, cfg.dz, False)
        z_h = coord(cfg.n3p, cfg.zmn, cfg.dz, True)
    else:
        z_i = z_h = np.zeros(1)

    def envelope(x, y, z):
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        r2 = Y * Y + Z * Z
        xi = X - xc
        return X, Y, Z, np.exp(-r2 / (w0 * w0)) * np.exp(-(xi / sig) ** 2), xi

    if cfg.ndim == 2:
        _, _, _, env, xi = envelope(x_i, y_h, z_i)   # Ey(i, j+1/2)
        ebf[:, :, :, 1] = amp * env * np.cos(om * xi)
        _, _, _, env, xi = envelope(x_h, y_h, z_i)   # Bz(i+1/2, j+1/2)
        ebf[:, :, :, 2] = amp * env * np.cos(om * xi)
        _, Y, _, env, xi = envelope(x_h, y_i, z_i)   # Ex(i+1/2, j)
        ebf[:, :, :, 0] = amp * env * (2.0 * Y / (w0 * w0)) * np.sin(om * xi) / om
    else:
        _, _, _, env, xi = envelope(x_i, y_h, z_i)   # Ey(i, j+1/2, k)
        ebf[:, :, :, 1] = amp * env * np.cos(om * xi)
        _, _, _, env, xi = envelope(x_h, y_h, z_i)   # Bz(i+1/2, j+1/2, k)
        ebf[:, :, :, 5] = amp * env * np.cos(om * xi)
        _, Y, _, env, xi = envelope(x_h, y_i, z_i)   # Ex(i+1/2, j, k)
        ebf[:, :, :, 0] = amp * env * (2.0 * Y / (w0 * w0)) * np.sin(om * xi) / om
        _, _, Z, env, xi = envelope(x_i, y_h, z_h)   # Bx(i, j+1/2, k+1/2)
        ebf[:, :, :, 3] = amp * env * (2.0 * Z / (w0 * w0)) * np.sin(om * xi) / om
    if polarization != "y":
        raise ValueError("only y polarisation is generated")
    return ebf


def lwfa_case(cfg: LPConfig, *, ppc, a0: float, w0: float, tau_fwhm: float, xc: float,
              x_start_cell: float = 0.0, seed: int = 0, thermal: float = 0.0,
              ramp_cells: float = 0.0):
    """Plasma (x > x_start_cell) plus a Gaussian pulse: (part, ebf)."""
    part = uniform_plasma(cfg, ppc, x_start_cell=x_start_cell, seed=seed, thermal=thermal,
                          ramp_cells=ramp_cells)
    ebf = gaussian_laser(cfg, a0=a0, w0=w0, tau_fwhm=tau_fwhm, xc=xc)
    return part, ebf


# ----------------------------------------------------------- named workloads
def lwfa_2d_config(nx=512, ny=256) -> LPConfig:
    """2-D Cartesian LWFA: dx = dy = 0.05, n/nc = 0.01, cfl 0.9."""
    return LPConfig(ndim=2, nx=nx, ny=ny, k0=20.0, yx_rat=1.0, cfl=0.9, n_over_nc=0.01)


def lwfa_3d_config(nx=256, ny=128, nz=128, n_over_nc=0.005) -> LPConfig:
    """3-D LWFA: dx = 0.05, dy = dz = 0.2, cfl 0.8.

    cfl 0.8 is the value of the reference's example inputs
    (examples/input_lwfa.nml:122).  With open boundaries the first-order Mur
    conditions meet at the box corners, and there the 3-D Yee update grows
    exponentially at cfl 0.9 (a property of the scheme, reproduced by the
    oracle; tests/test_oracle_math.py checks 0.8 is stable)."""
    return LPConfig(ndim=3, nx=nx, ny=ny, nz=nz, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.8,
                    n_over_nc=n_over_nc)


def lwfa_2d_case(nx=512, ny=256, seed=0):
    """512x256, 2x2 ppc for x > 64 cells, a0 = 2, w0 = 8, tau 6 periods at x = 32 cells."""
    cfg = lwfa_2d_config(nx, ny)
    part, ebf = lwfa_case(cfg, ppc=(2, 2), a0=2.0, w0=8.0, tau_fwhm=6.0, xc=32 * cfg.dx,
                          x_start_cell=nx / 8, seed=seed)
    return cfg, part, ebf


def lwfa_3d_case(nx=256, ny=128, nz=128, seed=0, ppc=(2, 2, 2)):
    """256x128x128, n/nc = 0.005, 8 ppc, a0 = 3, w0 = 10 (pulse centred in x/4)."""
    cfg = lwfa_3d_config(nx, ny, nz)
    part, ebf = lwfa_case(cfg, ppc=ppc, a0=3.0, w0=10.0, tau_fwhm=6.0, xc=nx * cfg.dx / 4,
                          x_start_cell=0.0, seed=seed)
    return cfg, part, ebf


# ------------------------------------------------- batched 2-D parameter scan
def scan_2d_params(seed: int = 7, nruns: int = 64) -> list[dict]:
    """Parameters of the batched 2-D scan: nruns independent LP runs.

    Drawn from numpy's PCG64 generator (np.random.Generator(PCG64(seed)),
    the generator's published algorithm, so the table is the same on every
    machine with numpy >= 1.17), one run after another, in this order:
    nx from {128, 160, ..., 1024}, ny from {64, 96, ..., 512}, n/nc
    log-uniform on [1e-3, 2e-2], a0 uniform on [1, 4], particles per cell
    per axis from {1, 2, 3, 4} (x then y).  Run i is seeded by i.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    runs = []
    for i in range(nruns):
        nx = 128 + 32 * int(rng.integers(0, 29))
        ny = 64 + 32 * int(rng.integers(0, 15))
        nnc = float(math.exp(rng.uniform(math.log(1e-3), math.log(2e-2))))
        a0 = float(rng.uniform(1.0, 4.0))
        ppc = [int(v) for v in rng.integers(1, 5, size=2)]
        runs.append({"run": i, "nx": nx, "ny": ny, "n_over_nc": nnc, "a0": a0, "ppc": ppc,
                     "seed": i})
    return runs


def scan_2d_case(run: dict):
    """One run of the batched 2-D scan: (cfg, part, ebf).

    The cells are those of the 2-D LWFA workload (dx = dy = 0.05, cfl 0.9).
    The pulse (FWHM tau = min(6, nx*dx/8) periods, waist ny*dy/4) is centred
    at 1.5 tau, fully inside the box; the plasma starts 1.5 tau ahead of the
    pulse centre, at x = 3 tau, with jittered particles seeded by the run.
    """
    nx, ny = run["nx"], run["ny"]
    cfg = LPConfig(ndim=2, nx=nx, ny=ny, k0=20.0, yx_rat=1.0, cfl=0.9,
                   n_over_nc=run["n_over_nc"])
    tau = min(6.0, nx * cfg.dx / 8.0)
    part = uniform_plasma(cfg, tuple(run["ppc"]), x_start_cell=3.0 * tau / cfg.dx,
                          seed=run["seed"])
    ebf = gaussian_laser(cfg, a0=run["a0"], w0=ny * cfg.dy / 4.0, tau_fwhm=tau, xc=1.5 * tau)
    return cfg, part, ebf


# ------------------------------------------- plasma of the large 3-D study
def uniform_plasma_slabs(cfg: LPConfig, ppc, *, seed: int = 0, thermal: float = 0.0,
                         ramp_cells: float = 0.0, charge: int = -1, jitter: bool = True,
                         slab_cells: int = 32, workers: int = 8) -> np.ndarray:
    """The plasma of uniform_plasma over the whole box, made slab by slab in
    threads for sizes where one numpy pass is too slow (hundreds of millions
    of particles): the same sub-cell centres, U(-1/2, 1/2) sub-cell jitter,
    clipping, N(0, thermal) momenta and linear weight ramp.  The random
    numbers come from one PCG64 stream per slab of slab_cells x-cells,
    seeded by SeedSequence([seed, slab]), so the particles are not those of
    uniform_plasma with the same seed.  Particles are x-fastest within a
    slab.  Returns part(np, 2D+1), column-major.
    """
    from concurrent.futures import ThreadPoolExecutor

    D = cfg.curr_ndim
    ppc = tuple(int(v) for v in ppc)[:D]
    dims = (cfg.nx, cfg.ny, cfg.nz)[:D]
    mins = (cfg.xmn, cfg.ymn, cfg.zmn)
    sizes = (cfg.dx, cfg.dy, cfg.dz)
    per_x = int(np.prod([dims[d] * ppc[d] for d in range(1, D)]))
    nslab = (cfg.nx + slab_cells - 1) // slab_cells
    n = cfg.nx * ppc[0] * per_x
    part = np.empty((n, 2 * D + 1), dtype=np.float64, order="F")
    w0 = cfg.n0_ref / float(np.prod(ppc))

    def sub_axis(c0, c1, d):
        return ((np.arange(c0, c1)[:, None] + (np.arange(ppc[d]) + 0.5)[None, :] / ppc[d])
                .ravel())

    def one(k):
        c0, c1 = k * slab_cells, min(cfg.nx, (k + 1) * slab_cells)
        o0 = c0 * ppc[0] * per_x
        m = (c1 - c0) * ppc[0] * per_x
        rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, k])))
        shape = (dims[2] * ppc[2], dims[1] * ppc[1], (c1 - c0) * ppc[0]) if D == 3 else \
            (dims[1] * ppc[1], (c1 - c0) * ppc[0])
        axes = [sub_axis(c0, c1, 0)] + [sub_axis(0, dims[d], d) for d in range(1, D)]
        for d in range(D):
            col = part[o0:o0 + m, d]
            view = col.reshape(shape)
            bshape = [1] * D
            bshape[D - 1 - d] = axes[d].size
            np.copyto(view, axes[d].reshape(bshape))
            if jitter:
                col += (rng.random(m) - 0.5) / ppc[d]
            np.clip(col, 1e-6, dims[d] - 1e-6, out=col)
            col *= sizes[d]
            col += mins[d]
        for d in range(D):
            col = part[o0:o0 + m, D + d]
            if thermal > 0:
                rng.standard_normal(out=col)
                col *= thermal
            else:
                col[:] = 0.0
        if ramp_cells > 0:
            xl = (part[o0:o0 + m, 0] - cfg.xmn) / cfg.dx
            wgh = (w0 * np.clip(xl / ramp_cells, 0.0, 1.0)).astype(np.float32)
        else:
            wgh = np.full(m, w0, dtype=np.float32)
        words = part[o0:o0 + m, 2 * D].view(np.uint32).reshape(m, 2)
        words[:, 0] = np.uint32(charge & 0xFFFF)
        words[:, 1] = wgh.view(np.uint32)
        return m

    with ThreadPoolExecutor(max_workers=workers) as ex:
        assert sum(ex.map(one, range(nslab))) == n
    return part


def lwfa_3d_large_case(nx=1024, ny=256, nz=256, seed=0, workers=8):
    """The large FP32 study: 1024x256x256 cells of the 3-D LWFA workload,
    n/nc = 0.002 with a linear weight up-ramp over the first 128 cells, 2x2x2
    ppc with Maxwellian momenta N(0, 1e-3), a0 = 4, w0 = 10, FWHM 6 periods,
    pulse centred at nx*dx/4."""
    cfg = lwfa_3d_config(nx, ny, nz, n_over_nc=0.002)
    part = uniform_plasma_slabs(cfg, (2, 2, 2), seed=seed, thermal=1e-3,
                                ramp_cells=128.0 * nx / 1024, workers=workers)
    ebf = gaussian_laser(cfg, a0=4.0, w0=10.0, tau_fwhm=6.0, xc=nx * cfg.dx / 4)
    return cfg, part, ebf


# ------------------------------------------------ PWFA-like bunch in plasma
def pwfa_3d_config(nx=512, ny=192, nz=192, n0_ref=0.1, cfl=0.8) -> LPConfig:
    """3-D box in plasma skin depths: dx = 0.1, dy = dz = 0.2 kp^-1 at the
    density n0_ref (1e18 cm^-3 units; 0.1 = 1e17 cm^-3), lengths in micron
    as everywhere (kp^2 = ompe*n0_ref, so kp^-1 = 16.8 um at 1e17 cm^-3).
    Two electron species: the plasma (0) and the bunch (1)."""
    ncrit = PI / (2.81794033 * 1.0)
    probe = LPConfig(ndim=3, nx=8, ny=8, nz=8)
    skin = 1.0 / math.sqrt(probe.ompe * n0_ref)
    return LPConfig(ndim=3, nx=nx, ny=ny, nz=nz, k0=1.0 / (0.1 * skin), yx_rat=2.0, zx_rat=2.0,
                    cfl=cfl, n_over_nc=1.0e-3 * n0_ref / ncrit, nsp=2,
                    lorentz_fact=(1.0, 1.0, 1.0, 1.0))


def skin_depth(cfg: LPConfig) -> float:
    """kp^-1 in micron at the configuration's n0_ref."""
    return 1.0 / math.sqrt(cfg.ompe * cfg.n0_ref)


def gaussian_bunch(cfg: LPConfig, n: int, *, xc: float, sigma, gamma: float, spread: float,
                   eps_n: float, nb_over_n0: float, seed: int, charge: int = -1,
                   cut: float = 4.5) -> np.ndarray:
    """Electron bunch: x, y, z ~ N((xc, 0, 0), sigma) cut at `cut` sigma
    (redrawn), p_x = gamma*(1 + spread*N(0,1)), p_y, p_z ~ N(0, eps_n/sigma_y,z)
    (normalised emittance eps_n = sigma_y*sigma_py, micron).  Equal weights
    that give a peak density nb_over_n0*n0_ref: w = nb/n0 * n0_ref *
    (2 pi)^(3/2) sx sy sz / (n dx dy dz), as plasma weights n0_ref/ppc give
    the density n0_ref.  Returns part(n, 7), column-major."""
    rng = np.random.Generator(np.random.PCG64(seed))

    def cut_normal(size):
        v = rng.standard_normal(size)
        bad = np.abs(v) > cut
        while bad.any():
            v[bad] = rng.standard_normal(int(bad.sum()))
            bad = np.abs(v) > cut
        return v

    sx, sy, sz = sigma
    part = np.zeros((n, 7), dtype=np.float64, order="F")
    part[:, 0] = xc + sx * cut_normal(n)
    part[:, 1] = sy * cut_normal(n)
    part[:, 2] = sz * cut_normal(n)
    part[:, 3] = gamma * (1.0 + spread * rng.standard_normal(n))
    part[:, 4] = (eps_n / sy) * rng.standard_normal(n)
    part[:, 5] = (eps_n / sz) * rng.standard_normal(n)
    w = nb_over_n0 * cfg.n0_ref * (2.0 * PI) ** 1.5 * sx * sy * sz / (n * cfg.dx * cfg.dy * cfg.dz)
    part[:, 6] = pack_weight(charge, 0, np.full(n, w, dtype=np.float32))
    return part


def grid_charge_host(cfg: LPConfig, part: np.ndarray) -> np.ndarray:
    """rho(n1p, n2p, n3p) = sum charge*wgh*W_x W_y W_z with the quadratic
    node weights of set_grid_charge (grid_part_util.f90:240-335; qden_*,
    grid_part_lib.f90:337-358, 388-417), for initial conditions."""
    D = cfg.curr_ndim
    c, _, wgh = unpack_weight(part[:, 2 * D])
    q = c.astype(np.float64) * wgh.astype(np.float64)
    dims = (cfg.n1p, cfg.n2p, cfg.n3p)
    idx, wts = [], []
    for d in range(D):
        mn, h = (cfg.xmn, cfg.ymn, cfg.zmn)[d], (cfg.dx, cfg.dy, cfg.dz)[d]
        xx = 3.0 + (part[:, d] - mn) / h
        ix = np.floor(xx + 0.5).astype(np.int64)
        s = xx - ix
        a1 = 0.75 - s * s
        a2 = 0.5 * (0.25 + s * s + s)
        a0 = 1.0 - a1 - a2
        idx.append(ix - 2)  # 0-based index of node ix-1
        wts.append((a0, a1, a2))
    rho = np.zeros(int(np.prod(dims)), dtype=np.float64)
    offs = [(i, j, k) for k in range(3 if D == 3 else 1) for j in range(3) for i in range(3)]
    for i, j, k in offs:
        flat = (idx[0] + i) + dims[0] * ((idx[1] + j) + dims[1] * ((idx[2] + k) if D == 3 else 0))
        w = q * wts[0][i] * wts[1][j] * (wts[2][k] if D == 3 else 1.0)
        rho += np.bincount(flat, weights=w, minlength=rho.size)
    return rho.reshape(dims, order="F")


def bunch_fields(cfg: LPConfig, bunch: np.ndarray, ebf: np.ndarray | None = None) -> np.ndarray:
    """Space-charge field of an ultra-relativistic bunch, added to ebf.

    Per x node plane, the discrete transverse Poisson problem
    D_y^2 phi + D_z^2 phi = -ompe*rho (the Yee grid's 5-point operator,
    phi = 0 on the outermost nodes) is solved with a type-I sine transform;
    E_y = -D_y phi, E_z = -D_z phi at their Yee points, so that the discrete
    div E equals ompe*rho (the Gauss law the step conserves), and
    B_z = beta*E_y, B_y = -beta*E_z averaged to their half-integer x.
    E_x, of order 1/gamma^2, is left at zero.  beta from the mean p_x."""
    from scipy.fft import dstn, idstn

    if cfg.ndim != 3:
        raise ValueError("bunch_fields is 3-D")
    rho = grid_charge_host(cfg, bunch)
    n1, n2, n3 = rho.shape
    M2, M3 = n2 - 2, n3 - 2
    lam_y = -(4.0 / cfg.dy ** 2) * np.sin(PI * np.arange(1, M2 + 1) / (2.0 * (M2 + 1))) ** 2
    lam_z = -(4.0 / cfg.dz ** 2) * np.sin(PI * np.arange(1, M3 + 1) / (2.0 * (M3 + 1))) ** 2
    lam = lam_y[:, None] + lam_z[None, :]
    src = -cfg.ompe * rho[:, 1:n2 - 1, 1:n3 - 1]
    phi = np.zeros_like(rho)
    phi[:, 1:n2 - 1, 1:n3 - 1] = idstn(dstn(src, type=1, axes=(1, 2)) / lam[None], type=1,
                                       axes=(1, 2))
    ey = np.zeros_like(rho)
    ez = np.zeros_like(rho)
    ey[:, :-1, :] = -(phi[:, 1:, :] - phi[:, :-1, :]) / cfg.dy
    ez[:, :, :-1] = -(phi[:, :, 1:] - phi[:, :, :-1]) / cfg.dz
    pxm = float(np.mean(bunch[:, 3]))
    beta = pxm / math.sqrt(1.0 + pxm * pxm)
    out = np.zeros(cfg.field_shape, dtype=np.float64, order="F") if ebf is None else ebf
    out[:, :, :, 1] += ey
    out[:, :, :, 2] += ez
    out[:-1, :, :, 5] += beta * 0.5 * (ey[:-1] + ey[1:])
    out[:-1, :, :, 4] -= beta * 0.5 * (ez[:-1] + ez[1:])
    return out


def pwfa_3d_case(nx=512, ny=192, nz=192, *, n_bunch=2_000_000, sigma_kp=(2.0, 1.5, 1.5),
                 xc_kp=12.0, gamma=2000.0, spread=0.01, eps_n=2.0, nb_over_n0=1.0,
                 ppc=(1, 2, 2), seed=3, workers=8):
    """Electron bunch (sigma in kp^-1, gamma 2000, 1 % spread, eps_n 2 mm mrad
    = 2 micron) at xc in a uniform plasma of 1e17 cm^-3 filling the box,
    4 particles per cell, with the bunch's space-charge field.
    Returns (cfg, plasma, bunch, ebf)."""
    cfg = pwfa_3d_config(nx, ny, nz)
    kp = skin_depth(cfg)
    plasma = uniform_plasma_slabs(cfg, ppc, seed=seed + 1, workers=workers)
    bunch = gaussian_bunch(cfg, n_bunch, xc=xc_kp * kp, sigma=tuple(s * kp for s in sigma_kp),
                           gamma=gamma, spread=spread, eps_n=eps_n, nb_over_n0=nb_over_n0,
                           seed=seed)
    return cfg, plasma, bunch, bunch_fields(cfg, bunch)
