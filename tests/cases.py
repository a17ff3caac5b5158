"""Small seeded parity cases shared by the CPU and GPU tests."""
from __future__ import annotations

import numpy as np

from aladyn_amd.config import LPConfig
from aladyn_amd.inputs import gaussian_laser, pack_weight, uniform_plasma


def small_2d(nx=48, ny=32, seed=1, a0=2.0):
    cfg = LPConfig(ndim=2, nx=nx, ny=ny, k0=20.0, yx_rat=1.0, cfl=0.9, n_over_nc=0.01)
    part = uniform_plasma(cfg, (2, 2), x_start_cell=4, seed=seed, thermal=0.05)
    ebf = gaussian_laser(cfg, a0=a0, w0=0.5, tau_fwhm=0.6, xc=nx * cfg.dx * 0.5)
    return cfg, part, ebf


def small_3d(nx=24, ny=16, nz=16, seed=2, a0=2.0):
    # cfl 0.8 as the reference's example inputs (examples/input_lwfa.nml:122);
    # at 0.9 the first-order Mur corners of the 3-D Yee scheme are unstable
    cfg = LPConfig(ndim=3, nx=nx, ny=ny, nz=nz, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.8,
                   n_over_nc=0.005)
    part = uniform_plasma(cfg, (2, 2, 2), x_start_cell=2, seed=seed, thermal=0.05)
    ebf = gaussian_laser(cfg, a0=a0, w0=1.2, tau_fwhm=0.6, xc=nx * cfg.dx * 0.5)
    return cfg, part, ebf


def random_fields(cfg, seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.normal(0.0, scale, cfg.field_shape))


def random_particles(cfg, n, seed=4, pmax=2.0, margin=3.0, charge=-1):
    """Particles uniformly inside the box, at least `margin` cells from faces."""
    rng = np.random.default_rng(seed)
    D = cfg.curr_ndim
    part = np.zeros((n, cfg.ncol), order="F")
    mins = (cfg.xmn, cfg.ymn, cfg.zmn)
    sizes = (cfg.dx, cfg.dy, cfg.dz)
    ns = (cfg.nx, cfg.ny, cfg.nz)
    for d in range(D):
        part[:, d] = mins[d] + rng.uniform(margin, ns[d] - margin, n) * sizes[d]
        part[:, D + d] = rng.uniform(-pmax, pmax, n)
    part[:, 2 * D] = pack_weight(charge, 0, rng.uniform(0.1, 1.0, n).astype(np.float32))
    return part
