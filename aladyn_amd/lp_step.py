"""Host-side mirror of the reference's LP-step interface over the C-ABI.

Method names follow the Fortran routines they replace
(src/dynamics/pic_evolve_in_time.f90, src/dynamics/curr_and_fields_util.f90):
``lpf2_evolve``, ``pfields_prepare``, ``particle_stage`` (the species-loop
body), ``curr_mpi_collect``, ``advance_lpf_fields``; diagnostics mirror
``energy_momenta`` / ``Envar`` / ``set_grid_charge``.  Arrays use the
reference's Fortran layouts as numpy column-major arrays:
part(np, 2*curr_ndim+1), ebf(n1p, n2p, n3p, nfield), jc(n1p, n2p, n3p, curr_ndim).
Every call goes to the HIP library; errors raise AladynError.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from ._lib import PREC_F32, PREC_F64, AladynConfig, AladynError, load
from .config import LPConfig


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class LPStep:
    """One device context holding the fields, currents and species of a run."""

    def __init__(self, cfg: LPConfig, precision: int = PREC_F64, *, tile=None, block: int = 0,
                 sort_every: int = 0, track_ids: bool = False):
        if precision not in (PREC_F64, PREC_F32):
            raise ValueError("precision must be 64 or 32")
        self.lib = load()
        self.cfg = cfg
        self.precision = precision
        c = AladynConfig()
        c.ndim = cfg.ndim
        c.nx, c.ny, c.nz = cfg.nx, cfg.ny, cfg.nz
        c.ibx, c.iby, c.ibz = cfg.ibx, cfg.iby, cfg.ibz
        c.der_ord = cfg.der_ord
        c.precision = precision
        c.nsp = cfg.nsp
        c.dx, c.dy, c.dz = cfg.dx, cfg.dy, cfg.dz
        c.dt = cfg.dt
        c.cfl = cfg.cfl
        c.ompe = cfg.ompe
        for i in range(4):
            c.lorentz_fact[i] = cfg.lorentz_fact[i]
        if tile is not None:
            for i in range(3):
                c.tile[i] = int(tile[i]) if i < len(tile) else 0
        c.block = block
        c.sort_every = sort_every
        c.reserved[0] = 1 if track_ids else 0
        self._c = c
        h = C.c_void_p()
        rc = self.lib.aladyn_create(C.byref(c), C.byref(h))
        self._h = h
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise AladynError(rc, msg)

    # ------------------------------------------------------------ plumbing
    def last_error(self) -> str:
        if not self._h:
            return "no context"
        s = self.lib.aladyn_last_error(self._h)
        return s.decode() if s else ""

    def _ck(self, rc: int):
        if rc != 0:
            raise AladynError(rc, self.last_error())

    def close(self):
        if getattr(self, "_h", None):
            self.lib.aladyn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ transfers
    def put_species(self, ic: int, part: np.ndarray):
        part = np.asfortranarray(part, dtype=np.float64)
        if part.ndim != 2 or part.shape[1] != self.cfg.ncol:
            raise ValueError(f"part must be (np, {self.cfg.ncol})")
        n = part.shape[0]
        self._ck(self.lib.aladyn_put_species(self._h, ic, _dptr(part), max(n, 1), n))

    def species_count(self, ic: int = 0) -> int:
        n = C.c_int64()
        self._ck(self.lib.aladyn_species_count(self._h, ic, C.byref(n)))
        return n.value

    def get_species(self, ic: int = 0) -> np.ndarray:
        n = self.species_count(ic)
        out = np.zeros((max(n, 1), self.cfg.ncol), dtype=np.float64, order="F")
        got = C.c_int64()
        self._ck(self.lib.aladyn_get_species(self._h, ic, _dptr(out), max(n, 1), C.byref(got)))
        return out[: got.value]

    def put_fields(self, ebf: np.ndarray):
        ebf = np.asfortranarray(ebf, dtype=np.float64)
        if ebf.shape != self.cfg.field_shape:
            raise ValueError(f"ebf must have shape {self.cfg.field_shape}")
        self._ck(self.lib.aladyn_put_fields(self._h, _dptr(ebf)))

    def get_fields(self) -> np.ndarray:
        out = np.zeros(self.cfg.field_shape, dtype=np.float64, order="F")
        self._ck(self.lib.aladyn_get_fields(self._h, _dptr(out)))
        return out

    def get_current(self) -> np.ndarray:
        out = np.zeros(self.cfg.current_shape, dtype=np.float64, order="F")
        self._ck(self.lib.aladyn_get_current(self._h, _dptr(out)))
        return out

    # ------------------------------------------------------------- the step
    def lpf2_evolve(self, iter_loc: int = 0, initial_time: bool = False):
        self._ck(self.lib.aladyn_lpf2_evolve(self._h, iter_loc, 1 if initial_time else 0))

    def set_graphs(self, enable: bool = True):
        """Replays lpf2_evolve steps as captured HIP graphs (aladyn_set_graphs)."""
        self._ck(self.lib.aladyn_set_graphs(self._h, 1 if enable else 0))

    def pfields_prepare(self):
        self._ck(self.lib.aladyn_pfields_prepare(self._h))

    def curr_clean(self):
        self._ck(self.lib.aladyn_curr_clean(self._h))

    def particle_stage(self, ic: int = 0, initial_time: bool = False):
        self._ck(self.lib.aladyn_particle_stage(self._h, ic, 1 if initial_time else 0))

    def curr_mpi_collect(self):
        self._ck(self.lib.aladyn_curr_mpi_collect(self._h))

    def advance_lpf_fields(self):
        self._ck(self.lib.aladyn_advance_lpf_fields(self._h))

    def set_dt(self, dt: float):
        """Lp_cycle's trim of the last step, dt_loc = tmax - tnow
        (ALaDyn.F90:101): the following steps use dt."""
        self._ck(self.lib.aladyn_set_dt(self._h, float(dt)))
        self.cfg = dataclasses.replace(self.cfg, dt=float(dt))

    # ------------------------------------------------------ envelope model
    def put_envelope(self, env: np.ndarray, oml: float):
        """The envelope env(n1p, n2p, n3p, 4) = A_R^n, A_I^n, A_R^{n-1},
        A_I^{n-1} (array_alloc.f90:108-111) and the laser frequency oml."""
        c = self.cfg
        a = np.asfortranarray(env, dtype=np.float64)
        if a.shape != (c.n1p, c.n2p, c.n3p, 4):
            raise ValueError(f"env shape {a.shape}, expected {(c.n1p, c.n2p, c.n3p, 4)}")
        self._ck(self.lib.aladyn_put_envelope(self._h, float(oml), _dptr(a)))

    def get_envelope(self) -> np.ndarray:
        c = self.cfg
        out = np.zeros((c.n1p, c.n2p, c.n3p, 4), dtype=np.float64, order="F")
        self._ck(self.lib.aladyn_get_envelope(self._h, _dptr(out)))
        return out

    def env_kernel_timing(self, reset: bool = False):
        """(push_ms, move_ms, launches): mean durations of the envelope step's
        two particle kernels (HIP events on the context stream)."""
        a, b, n = C.c_double(), C.c_double(), C.c_int64()
        self._ck(self.lib.aladyn_env_kernel_timing(self._h, 1 if reset else 0, C.byref(a),
                                                   C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def env_lpf2_evolve(self, it: int = 0):
        """One env_lpf2_evolve (env_evolve_in_time.f90:196-361) for species 0."""
        self._ck(self.lib.aladyn_env_lpf2_evolve(self._h, int(it)))

    def synchronize(self):
        self._ck(self.lib.aladyn_synchronize(self._h))

    # ---------------------------------------------------------- diagnostics
    def energy_momenta(self, ic: int = 0):
        ek = np.zeros(7)
        emax = np.zeros(1)
        self._ck(self.lib.aladyn_energy_momenta(self._h, ic, _dptr(ek), _dptr(emax)))
        return ek, float(emax[0])

    def field_energy(self) -> np.ndarray:
        out = np.zeros(6)
        self._ck(self.lib.aladyn_field_energy(self._h, _dptr(out)))
        return out[: self.cfg.nfield]

    def beam_fields(self, ic: int, gam2: float, bet0: float):
        """Adds the bunch fields of species ic to the fields
        (init_beam_part_distrib.f90:538-602: set_grid_charge, the
        sine-transform Poisson solve of fft_3d_psolv / fft_2d_psolv,
        initial_beam_fields); gam2 = gamma^2, bet0 = beta of the bunch."""
        self._ck(self.lib.aladyn_beam_fields(self._h, int(ic), float(gam2), float(bet0)))

    def grid_charge(self, ic: int = 0) -> np.ndarray:
        rho = np.zeros((self.cfg.n1p, self.cfg.n2p, self.cfg.n3p), dtype=np.float64, order="F")
        self._ck(self.lib.aladyn_grid_charge(self._h, ic, _dptr(rho)))
        return rho

    def set_diagnostics(self, enable: bool = True):
        self._ck(self.lib.aladyn_set_diagnostics(self._h, 1 if enable else 0))

    def continuity_residual(self):
        out = np.zeros(2)
        self._ck(self.lib.aladyn_continuity_residual(self._h, _dptr(out)))
        return float(out[0]), float(out[1])

    def lp_window_xshift(self, witr: int, init_iter: int, w_speed: float) -> int:
        """lp_window_xshift (window.f90:276-347): moves the box; returns nshx."""
        n = C.c_int32()
        self._ck(self.lib.aladyn_lp_window_xshift(self._h, int(witr), int(init_iter),
                                                  float(w_speed), C.byref(n)))
        return int(n.value)

    def inject_species(self, ic: int, part: np.ndarray):
        """Appends particles (add_particles, window.f90:39-111)."""
        p = np.asfortranarray(part, dtype=np.float64)
        self._ck(self.lib.aladyn_inject_species(self._h, ic, _dptr(p), p.shape[0], p.shape[0]))

    def set_injection(self, ic: int, table: dict, nptx: int = 0, targ_in: float | None = None,
                      targ_end: float | None = None):
        """Registers species ic's target table (particles_inject /
        add_particles, window.f90:39-139): a dict with xpt, ypt, zpt (3-D),
        wghpt and wghyz (nyp x nzp).  targ_in/targ_end default to the table's
        first and last column plus one cell."""
        xpt = np.ascontiguousarray(table["xpt"], dtype=np.float64)
        ypt = np.ascontiguousarray(table["ypt"], dtype=np.float64)
        zpt = np.ascontiguousarray(table.get("zpt", np.ones(1)), dtype=np.float64)
        wghpt = np.ascontiguousarray(table["wghpt"], dtype=np.float64)
        wghyz = np.asfortranarray(table["wghyz"], dtype=np.float64).ravel(order="F")
        t_in = float(xpt[0]) if targ_in is None else float(targ_in)
        t_end = float(xpt[-1] + self.cfg.dx) if targ_end is None else float(targ_end)
        self._ck(self.lib.aladyn_set_injection(self._h, ic, _dptr(xpt), xpt.size, _dptr(ypt),
                                               ypt.size, _dptr(zpt), zpt.size, _dptr(wghpt),
                                               _dptr(wghyz), int(nptx), t_in, t_end))

    def injection_state(self, ic: int) -> int:
        """Columns of species ic's target table consumed so far (nptx)."""
        n = C.c_int64()
        self._ck(self.lib.aladyn_injection_state(self._h, ic, C.byref(n)))
        return int(n.value)

    def set_box_origin(self, xmin: float):
        """Places the box at (xmin, xmin + nx*dx] before any put_species
        (restart of a moved window, pic_dump.f90:520-531)."""
        self._ck(self.lib.aladyn_set_box_origin(self._h, float(xmin)))

    def box(self) -> np.ndarray:
        """The current particle box (xmin, xmax, ymin, ymax, zmin, zmax)."""
        out = np.zeros(6)
        self._ck(self.lib.aladyn_box(self._h, _dptr(out)))
        return out

    def kernel_timing(self, reset: bool = False):
        pm = np.zeros(1)
        fm = np.zeros(1)
        n = C.c_int64()
        self._ck(self.lib.aladyn_kernel_timing(self._h, 1 if reset else 0, _dptr(pm), _dptr(fm),
                                               C.byref(n)))
        return float(pm[0]), float(fm[0])


class LPBatch:
    """Independent LPStep contexts advanced together, each kernel of the step
    launched once for all of them (aladyn_batch_*, include/aladyn_gpu.h)."""

    def __init__(self, steppers, graphs: bool = True):
        self.lib = load()
        self.steppers = list(steppers)
        arr = (C.c_void_p * len(self.steppers))(*[g._h.value for g in self.steppers])
        h = C.c_void_p()
        rc = self.lib.aladyn_batch_create(arr, len(self.steppers), 1 if graphs else 0,
                                          C.byref(h))
        self._h = h
        if rc != 0:
            msg = self.last_error()
            self.close()
            raise AladynError(rc, msg)

    def last_error(self) -> str:
        if not self._h:
            return "no batch"
        s = self.lib.aladyn_batch_last_error(self._h)
        return s.decode() if s else ""

    def _ck(self, rc: int):
        if rc != 0:
            raise AladynError(rc, self.last_error())

    def lpf2_evolve(self, it: int = 0, initial_time: bool = False):
        self._ck(self.lib.aladyn_batch_lpf2_evolve(self._h, int(it), 1 if initial_time else 0))

    def synchronize(self):
        self._ck(self.lib.aladyn_batch_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self.lib.aladyn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.synchronize()
        self.close()
