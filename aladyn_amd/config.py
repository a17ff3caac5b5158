"""Run constants of the LP step, derived the way ALaDyn derives them.

Mirrors the parts of set_grid (src/grid/set_grid_param.f90:52-225),
set_initial_param (src/start/set_init_param.f90:76-102, 151-162, 254-279)
and set_field_param (src/fields/grid_field_param.f90:36-99) that the hot path
reads: cell sizes, dt, ompe, djc, se_coeff, array extents and the particle box.
Units are the reference's: lengths in micron, time in micron/c, fields in
m c^2 / (e micron), momenta in m c.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

# phys_param.f90:27-49
PI = 3.141592653589793
RC0 = 2.81794033  # classical electron radius, 1e-13 cm


@dataclass
class LPConfig:
    ndim: int
    nx: int
    ny: int
    nz: int = 1
    k0: float = 20.0          # cells per micron along x (dx = 1/k0)
    yx_rat: float = 1.0       # dy/dx
    zx_rat: float = 1.0       # dz/dx
    cfl: float = 0.9
    der_ord: int = 2
    n_over_nc: float = 0.01   # electron density over critical density
    lam0: float = 1.0         # laser wavelength, micron
    nsp: int = 1
    lorentz_fact: tuple = (1.0, 1.0, 1.0, 1.0)  # 1/m per species (electrons: 1)
    dt: float | None = None   # dt_loc; default cfl/sqrt(sum dx_d^-2)
    ibx: int = 0
    iby: int = 0
    ibz: int = 0
    extra: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.ndim not in (2, 3):
            raise ValueError("ndim must be 2 or 3")
        if self.ndim == 2:
            self.nz = 1
        # set_init_param.f90:76 and set_grid: dx = 1/k0, dy = yx_rat*dx
        self.dx = 1.0 / self.k0
        self.dy = self.yx_rat * self.dx
        self.dz = self.zx_rat * self.dx if self.ndim == 3 else 1.0
        self.dx_inv = 1.0 / self.dx
        self.dy_inv = 1.0 / self.dy
        self.dz_inv = 1.0 / self.dz if self.ndim == 3 else 1.0
        if self.dt is None:
            s = self.dx_inv ** 2 + self.dy_inv ** 2
            if self.ndim == 3:
                s += self.dz_inv ** 2
            self.dt = self.cfl / math.sqrt(s)  # set_init_param.f90:79-90
        # ncrit = pi/(rc0*lam0^2) in 1e21 cm^-3 (set_init_param.f90:272);
        # n0_ref in 1e18 cm^-3 so that n_over_nc = 1e-3*n0_ref/ncrit (:279).
        ncrit = PI / (RC0 * self.lam0 * self.lam0)
        # laser frequency in units c/l0 (set_init_param.f90:291): the
        # envelope solver's om0
        self.oml = 2.0 * PI / self.lam0
        self.n0_ref = self.n_over_nc * ncrit * 1.0e3
        # ompe = omega_p^2 at the reference density 1e18 cm^-3 = 1e6 um^-3
        # (set_init_param.f90:254-255, phys_param.f90:57); the plasma density
        # enters through the particle weights n0_ref/ppc.
        reference_density = 1.0e6
        self.ompe = 1.0e-3 * (4.0 * PI * RC0 * (reference_density * 1.0e-6))
        self.nfield = 6 if self.ndim == 3 else 3
        self.curr_ndim = 3 if self.ndim == 3 else 2
        self.ncol = 2 * self.curr_ndim + 1
        self.n1p = self.nx + 5
        self.n2p = self.ny + 5
        self.n3p = self.nz + 5 if self.ndim == 3 else 1
        # interior (1-based Fortran) bounds
        self.ix1, self.ix2 = 3, self.nx + 2
        self.jy1, self.jy2 = 3, self.ny + 2
        self.kz1, self.kz2 = (3, self.nz + 2) if self.ndim == 3 else (1, 1)
        # particle box (loc_*grid%gmin, gmax on one rank)
        self.xmn, self.xmx = 0.0, self.dx * self.nx
        self.ymn, self.ymx = self.dy * -(self.ny // 2), self.dy * (self.ny - self.ny // 2)
        if self.ndim == 3:
            self.zmn, self.zmx = self.dz * -(self.nz // 2), self.dz * (self.nz - self.nz // 2)
        else:
            self.zmn = self.zmx = 0.0

    @property
    def field_shape(self):
        return (self.n1p, self.n2p, self.n3p, self.nfield)

    @property
    def current_shape(self):
        return (self.n1p, self.n2p, self.n3p, self.curr_ndim)

    @property
    def ncell(self) -> int:
        return self.n1p * self.n2p * self.n3p

    def algorithmic_bytes(self, n_particles: int, precision: int) -> float:
        """Bytes per step counted once (particle state + field/current grids)."""
        cells = self.nx * self.ny * (self.nz if self.ndim == 3 else 1)
        if precision == 32:
            bp, bc = (44, 52) if self.ndim == 2 else (60, 96)
        else:
            bp, bc = (72, 104) if self.ndim == 2 else (104, 192)
        return float(n_particles) * bp + float(cells) * bc
