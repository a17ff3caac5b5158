"""ctypes wrappers of the CPU checkers -- TEST INFRASTRUCTURE ONLY.

* Oracle: the C restatement (oracle/liboracle.so, pic_oracle.c).
* RefHarness: the reference's own Fortran routines compiled with
  oracle/ref_harness.f90 into oracle/_ref/libaladyn_ref.so (built only where
  the reference sources exist; the built library travels with the tree).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use this.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libaladyn_ref.so"


class OrcGrid(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("curr_ndim", C.c_int), ("nfield", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("n1p", C.c_int), ("n2p", C.c_int), ("n3p", C.c_int),
        ("ix1", C.c_int), ("ix2", C.c_int), ("jy1", C.c_int), ("jy2", C.c_int),
        ("kz1", C.c_int), ("kz2", C.c_int),
        ("ibx", C.c_int), ("iby", C.c_int), ("ibz", C.c_int),
        ("der_ord", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dx_inv", C.c_double), ("dy_inv", C.c_double), ("dz_inv", C.c_double),
        ("xmn", C.c_double), ("ymn", C.c_double), ("zmn", C.c_double),
        ("xmx", C.c_double), ("ymx", C.c_double), ("zmx", C.c_double),
        ("dt", C.c_double), ("ompe", C.c_double),
        ("djc", C.c_double * 3), ("se_coeff", C.c_double * 2),
        ("lorentz_fact", C.c_double),
    ]


_D = C.POINTER(C.c_double)
_G = C.POINTER(OrcGrid)
_I64 = C.c_int64


def _p(a):
    return a.ctypes.data_as(_D)


def _f(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need Fortran-order f64"
    return a


class OrcWindow(C.Structure):
    """Saved state of lp_window_xshift (window.f90:276-288)."""

    _fields_ = [("xlapse", C.c_double), ("wi1", C.c_int)]


class OrcEnv(C.Structure):
    """The envelope model's constants (om0 and set_field_param's hord_der2,
    opt_der1)."""

    _fields_ = [("om0", C.c_double), ("hord_der2", C.c_double), ("opt_der1", C.c_double)]


_E = C.POINTER(OrcEnv)


class Oracle:
    """The C restatement of the LP step on the reference layouts."""

    def __init__(self, cfg, lorentz_fact: float | None = None, lib: Path | None = None):
        p = Path(lib) if lib else ORACLE_LIB
        if not p.exists():
            raise ImportError(f"{p} missing: run `make -C oracle`")
        L = C.CDLL(str(p))
        sig = {
            "orc_grid_setup": [_G, C.c_double],
            "orc_gather": [_G, _I64, _D, _I64, _D, _D],
            "orc_charge_multiply": [_G, _I64, _D, _I64, _D],
            "orc_init_momenta": [_G, _I64, _D, _I64, _D],
            "orc_push": [_G, _I64, _D, _I64, _D],
            "orc_part_dist": [_G, _I64, _D, _I64, _D],
            "orc_esirkepov": [_G, _I64, _D, _I64, _D, _D],
            "orc_esirkepov_skipped": [],
            "orc_field_xyzbd": [_G, _D, C.c_int],
            "orc_curr_collect": [_G, _D],
            "orc_jc_xyzbd": [_G, _D],
            "orc_ef_bds": [_G, _D, C.c_double],
            "orc_bf_bds": [_G, _D, C.c_double],
            "orc_rote": [_G, _D, C.c_double],
            "orc_rotb": [_G, _D, C.c_double],
            "orc_advance_fields": [_G, _D, _D],
            "orc_particle_stage": [_G, _I64, _D, _I64, _D, _D, _D, C.c_int],
            "orc_lpf2_evolve": [_G, _I64, _D, _I64, _D, _D, _D, C.c_int],
            "orc_grid_charge": [_G, _I64, _D, _I64, _D],
            "orc_energy_momenta": [_G, _I64, _D, _I64, _D, _D],
            "orc_field_energy": [_G, _D, _D],
            "orc_window_xshift": [_G, C.POINTER(OrcWindow), C.c_int, C.c_int, C.c_double, _D,
                                  C.POINTER(C.c_int64), _D, _I64],
            "orc_fields_left_xshift": [_G, _D, C.c_int, C.c_int],
            "orc_sine_transform": [_D, C.c_int],
            "orc_beam_potential": [_G, C.c_double, _D],
            "orc_initial_beam_fields": [_G, _D, _D, C.c_double, C.c_double],
            "orc_beam_fields": [_G, C.c_double, C.c_double, _D, _D],
            "orc_inject_select": [_D, _I64, C.POINTER(C.c_int64), C.c_double, C.c_double,
                                  C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
            "orc_add_particles": [_G, _D, _I64, _I64, _D, C.c_int, _D, C.c_int, _D, _D, C.c_int,
                                  _D, _I64, _I64],
            "orc_env_acc": [_G, _I64, _D, _I64, _D, _D, C.c_double, _D],
            "orc_env_momenta": [_G, _I64, _D, _I64, _D],
            "orc_env_density": [_G, _I64, _D, _I64, _D],
            "orc_env_grad_interp": [_G, _I64, _D, _I64, _D, _D],
            "orc_env_positions": [_G, _I64, _D, _I64, _D],
            "orc_env_setup": [_G, C.c_double, C.c_double, _E],
            "orc_env_grad": [_G, _D],
            "orc_env_amp_prepare": [_G, _D, _D],
            "orc_env_fields_average": [_G, _D, _D],
            "orc_env_bds": [_G, _D, C.c_int, C.c_int, C.c_int, C.c_int],
            "orc_potential_lapl": [_G, _E, _D, _D, C.c_int, C.c_int],
            "orc_env_maxw_solve": [_G, _E, _D, _D, C.c_double],
            "orc_advance_lpf_envelope": [_G, _E, _D, _D],
            "orc_den_zyxbd": [_G, _D, C.c_int],
            "orc_env_lpf2_evolve": [_G, _E, _I64, _D, _I64, _D, _D, _D, _D],
        }
        for name, at in sig.items():
            fn = getattr(L, name)
            fn.argtypes = at
            fn.restype = _I64 if name in ("orc_part_dist", "orc_particle_stage", "orc_lpf2_evolve",
                                          "orc_esirkepov_skipped", "orc_env_lpf2_evolve") else None
            if name == "orc_window_xshift":
                fn.restype = C.c_int
            if name in ("orc_inject_select", "orc_add_particles"):
                fn.restype = _I64
        self.L = L
        self.cfg = cfg
        g = OrcGrid()
        g.ndim, g.nx, g.ny, g.nz = cfg.ndim, cfg.nx, cfg.ny, cfg.nz
        g.ibx, g.iby, g.ibz = cfg.ibx, cfg.iby, cfg.ibz
        g.der_ord = cfg.der_ord
        g.dx, g.dy, g.dz = cfg.dx, cfg.dy, cfg.dz
        g.dt = cfg.dt
        g.ompe = cfg.ompe
        g.lorentz_fact = cfg.lorentz_fact[0] if lorentz_fact is None else lorentz_fact
        L.orc_grid_setup(C.byref(g), cfg.cfl)
        self.g = g
        self.env = OrcEnv()
        L.orc_env_setup(C.byref(g), cfg.cfl, getattr(cfg, "oml", 0.0), C.byref(self.env))

    def set_om0(self, om0: float):
        """The envelope's laser frequency (oml) for the envelope solver."""
        self.L.orc_env_setup(C.byref(self.g), self.cfg.cfl, om0, C.byref(self.env))

    # array factories in the reference layouts
    def zeros_fields(self):
        return np.zeros(self.cfg.field_shape, dtype=np.float64, order="F")

    def zeros_current(self):
        return np.zeros(self.cfg.current_shape, dtype=np.float64, order="F")

    def zeros_stage(self, n):
        return np.zeros((max(n, 1), self.cfg.ncol), dtype=np.float64, order="F")

    def _ld(self, part):
        return max(part.shape[0], 1)

    def gather(self, part, ebf):
        ebfp = self.zeros_stage(part.shape[0])
        self.L.orc_gather(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(_f(ebf)), _p(ebfp))
        return ebfp

    def charge_multiply(self, part, ebfp):
        self.L.orc_charge_multiply(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(ebfp))

    def init_momenta(self, part, ebfp):
        self.L.orc_init_momenta(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(ebfp))

    def push(self, part, ebfp):
        self.L.orc_push(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(ebfp))

    def part_dist(self, part, ebfp) -> int:
        return int(self.L.orc_part_dist(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(ebfp)))

    def esirkepov(self, part, ebfp, jc, np_=None):
        n = part.shape[0] if np_ is None else np_
        self.L.orc_esirkepov(C.byref(self.g), n, _p(_f(part)), self._ld(part), _p(ebfp), _p(_f(jc)))

    def esirkepov_skipped(self) -> int:
        return int(self.L.orc_esirkepov_skipped())

    def field_xyzbd(self, ebf, nc=None):
        self.L.orc_field_xyzbd(C.byref(self.g), _p(_f(ebf)), self.cfg.nfield if nc is None else nc)

    def curr_collect(self, jc):
        self.L.orc_curr_collect(C.byref(self.g), _p(_f(jc)))

    def jc_xyzbd(self, jc):
        self.L.orc_jc_xyzbd(C.byref(self.g), _p(_f(jc)))

    def ef_bds(self, ebf, dtl):
        self.L.orc_ef_bds(C.byref(self.g), _p(_f(ebf)), dtl)

    def bf_bds(self, ebf, dtl):
        self.L.orc_bf_bds(C.byref(self.g), _p(_f(ebf)), dtl)

    def rote(self, ebf, dtf):
        self.L.orc_rote(C.byref(self.g), _p(_f(ebf)), dtf)

    def rotb(self, ebf, dtf):
        self.L.orc_rotb(C.byref(self.g), _p(_f(ebf)), dtf)

    def advance_fields(self, ebf, jc):
        self.L.orc_advance_fields(C.byref(self.g), _p(_f(ebf)), _p(_f(jc)))

    def particle_stage(self, part, ebf, jc, initial_time=False):
        """Returns (part_new[:np], ebfp); part is modified in place."""
        ebfp = self.zeros_stage(part.shape[0])
        n = self.L.orc_particle_stage(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                                      _p(ebfp), _p(_f(ebf)), _p(_f(jc)), 1 if initial_time else 0)
        return np.asfortranarray(part[:n]), ebfp

    def lpf2_evolve(self, part, ebf, jc, initial_time=False):
        """One full step; returns the surviving particles (a new array)."""
        ebfp = self.zeros_stage(part.shape[0])
        n = self.L.orc_lpf2_evolve(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                                   _p(ebfp), _p(_f(ebf)), _p(_f(jc)), 1 if initial_time else 0)
        return np.asfortranarray(part[:n])

    def lp_cycle(self, part, ebf, jc, *, tmax, tstart=0.0, w_speed=0.0, wi_time=0.0,
                 wf_time=float("inf"), w_sh=1, win=None):
        """Lp_cycle (ALaDyn.F90:89-107) with lp_run (pic_evolve_in_time.f90:
        113-131) and timing (run_data_info.f90:68-71): initial_time = tnow <
        dt_loc (ALaDyn.F90:66-67); while tnow < tmax: lpf2_evolve(iter), the
        window call when w_speed > 0, wi_time <= tnow < wf_time and mod(iter,
        w_sh) == 0; tnow += dt_loc, iter += 1; if tnow + dt_loc >= tmax,
        dt_loc = tmax - tnow (this oracle's grid dt is changed in place).
        Returns (particles, iter, tnow, the dt of every step)."""
        tnow, it = tstart, 0
        initial = tnow < self.g.dt
        win = OrcWindow() if win is None else win
        dts = []
        while tnow < tmax:
            part = self.lpf2_evolve(part, ebf, jc, initial_time=initial)
            dts.append(self.g.dt)
            if w_speed > 0.0 and wi_time <= tnow < wf_time and it % w_sh == 0:
                _, part = self.window_xshift(win, w_sh, it, w_speed, ebf, part)
            if tnow < tmax:
                tnow += self.g.dt
                it += 1
            if tnow + self.g.dt >= tmax:
                self.g.dt = tmax - tnow
            initial = False
        return part, it, tnow, dts

    def window_xshift(self, win, witr, init_iter, w_speed, ebf, part):
        """lp_window_xshift: returns (nshx, surviving particles as a new array);
        moves this oracle's box (xmn, xmx) with the window."""
        n = C.c_int64(part.shape[0])
        nshx = self.L.orc_window_xshift(C.byref(self.g), C.byref(win), witr, init_iter, w_speed,
                                        _p(_f(ebf)), C.byref(n), _p(_f(part)), self._ld(part))
        return int(nshx), np.asfortranarray(part[:n.value])

    def sine_transform(self, v):
        """FFTW_RODFT00 of the 1-D array v (a new array)."""
        out = np.array(v, dtype=np.float64)
        self.L.orc_sine_transform(_p(out), out.size)
        return out

    def beam_potential(self, rho, gam2):
        """fft_3d_psolv / fft_2d_psolv (sine transforms) of rho(n1p, n2p,
        n3p): a copy with the potential in the interior, ghosts kept."""
        jc = np.array(rho, dtype=np.float64, order="F")
        self.L.orc_beam_potential(C.byref(self.g), gam2, _p(jc))
        return jc

    def initial_beam_fields(self, pot, gam2, bet0):
        """initial_beam_fields from jc(1) = pot, jc(2) = bet0*pot on every
        node: returns efb(n1p, n2p, n3p, 6), zero where it sets nothing."""
        cfg = self.cfg
        jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 2), order="F")
        jc[..., 0] = pot
        jc[..., 1] = bet0 * jc[..., 0]
        efb = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 6), order="F")
        self.L.orc_initial_beam_fields(C.byref(self.g), _p(jc), _p(efb), gam2, bet0)
        return efb

    def beam_fields(self, rho, ebf, gam2, bet0):
        """The bunch fields of charge density rho added to ebf in place
        (init_beam_part_distrib.f90:538-602)."""
        cfg = self.cfg
        jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 2), order="F")
        jc[..., 0] = rho
        assert ebf.flags.f_contiguous and ebf.dtype == np.float64
        self.L.orc_beam_fields(C.byref(self.g), gam2, bet0, _p(jc), _p(ebf))

    def fields_left_xshift(self, fld, nshx):
        """fields_left_xshift of every component of fld (n1p, n2p, n3p, nc),
        in place: what lp_window_xshift does to the envelope arrays."""
        assert fld.flags.f_contiguous and fld.dtype == np.float64
        self.L.orc_fields_left_xshift(C.byref(self.g), _p(fld), fld.shape[3], nshx)

    def inject_select(self, xpt, nptx, xmx, targ_in, targ_end):
        """particles_inject's column selection: returns (ncols, c0, c1, new nptx)."""
        xpt = np.ascontiguousarray(xpt, dtype=np.float64)
        n = C.c_int64(nptx)
        c0, c1 = C.c_int64(0), C.c_int64(0)
        k = self.L.orc_inject_select(_p(xpt), xpt.size, C.byref(n), xmx, targ_in, targ_end,
                                     C.byref(c0), C.byref(c1))
        return int(k), int(c0.value), int(c1.value), int(n.value)

    def add_particles(self, part, table, c0, c1):
        """add_particles of columns c0..c1 of table (a dict: xpt, ypt, zpt,
        wghpt, wghyz, charge); returns part with the new rows appended."""
        ypt = np.ascontiguousarray(table["ypt"], dtype=np.float64)
        zpt = np.ascontiguousarray(table.get("zpt", np.zeros(1)), dtype=np.float64)
        xpt = np.ascontiguousarray(table["xpt"], dtype=np.float64)
        wghpt = np.ascontiguousarray(table["wghpt"], dtype=np.float64)
        wghyz = np.ascontiguousarray(table["wghyz"], dtype=np.float64).ravel(order="F")
        nyp, nzp = ypt.size, zpt.size if self.cfg.curr_ndim == 3 else 1
        add = max(c1 - c0 + 1, 0) * nyp * nzp
        np0 = part.shape[0]
        out = np.zeros((np0 + add, part.shape[1]), order="F")
        out[:np0] = part
        n = self.L.orc_add_particles(C.byref(self.g), _p(xpt), c0, c1, _p(ypt), nyp, _p(zpt), nzp,
                                     _p(wghpt), _p(wghyz), int(table["charge"]), _p(out),
                                     out.shape[0], np0)
        assert n == np0 + add
        return out

    # envelope model: av(n1p, n2p, n3p, 1 + curr_ndim) = F, grad F
    def zeros_envelope(self):
        c = self.cfg
        return np.zeros((c.n1p, c.n2p, c.n3p, c.curr_ndim + 1), dtype=np.float64, order="F")

    def env_acc(self, part, ebf, av, ebfp=None):
        """set_env_acc: returns the stage buffer (q E - q^2 grad F/2gamp,
        q B/gamp, wgh/gamp)."""
        ebfp = self.zeros_stage(part.shape[0]) if ebfp is None else ebfp
        self.L.orc_env_acc(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                           _p(_f(ebf)), _p(_f(av)), self.cfg.dt, _p(ebfp))
        return ebfp

    def env_momenta(self, part, ebfp):
        self.L.orc_env_momenta(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                               _p(ebfp))

    def env_density(self, ebfp, np_, den=None):
        """set_env_density: adds wgh/gamp shapes to den (n1p, n2p, n3p)."""
        c = self.cfg
        den = np.zeros((c.n1p, c.n2p, c.n3p), order="F") if den is None else den
        self.L.orc_env_density(C.byref(self.g), np_, _p(ebfp), max(ebfp.shape[0], 1), _p(_f(den)))
        return den

    def env_grad_interp(self, part, av, ebfp):
        self.L.orc_env_grad_interp(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                                   _p(_f(av)), _p(ebfp))

    def env_positions(self, part, ebfp):
        self.L.orc_env_positions(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part),
                                 _p(ebfp))

    def env_particle_step(self, part, ebf, av0, av1, jc):
        """The particle side of env_lpf2_evolve (env_evolve_in_time.f90:196-
        361) for one species: set_env_acc with (F, grad F) av0 at t^n,
        lpf_env_momenta, set_env_density, then set_env_grad_interp with av1 at
        t^{n+1/2}, lpf_env_positions, cell_part_dist and the Esirkepov
        deposit into jc.  Returns (surviving particles, density)."""
        p = part.copy(order="F")
        ebfp = self.env_acc(p, ebf, av0)
        self.env_momenta(p, ebfp)
        den = self.env_density(ebfp.copy(order="F"), p.shape[0])
        self.env_grad_interp(p, av1, ebfp)
        self.env_positions(p, ebfp)
        n = self.part_dist(p, ebfp)
        p = np.asfortranarray(p[:n])
        e = np.asfortranarray(ebfp[:max(n, 1)])
        self.esirkepov(p, e, jc, n)
        return p, den

    # envelope model, grid side: env(n1p, n2p, n3p, 4) = A_R^n, A_I^n,
    # A_R^{n-1}, A_I^{n-1}; jc in envelope mode has curr_ndim + 1 components
    def zeros_env_fields(self):
        c = self.cfg
        return np.zeros((c.n1p, c.n2p, c.n3p, 4), dtype=np.float64, order="F")

    def env_grad(self, av):
        self.L.orc_env_grad(C.byref(self.g), _p(_f(av)))

    def env_amp_prepare(self, env, av):
        self.L.orc_env_amp_prepare(C.byref(self.g), _p(_f(env)), _p(_f(av)))

    def env_fields_average(self, env, av):
        self.L.orc_env_fields_average(C.byref(self.g), _p(_f(env)), _p(_f(av)))

    def env_bds(self, ef, ptrght=2, ptlft=2, c1=1, c2=2):
        self.L.orc_env_bds(C.byref(self.g), _p(_f(ef)), ptrght, ptlft, c1, c2)

    def potential_lapl(self, apf, curr, ic1=1, ic2=2):
        self.L.orc_potential_lapl(C.byref(self.g), C.byref(self.env), _p(_f(apf)), _p(_f(curr)),
                                  ic1, ic2)

    def env_maxw_solve(self, curr, env, dtl=None):
        self.L.orc_env_maxw_solve(C.byref(self.g), C.byref(self.env), _p(_f(curr)), _p(_f(env)),
                                  self.cfg.dt if dtl is None else dtl)

    def advance_lpf_envelope(self, curr, env):
        self.L.orc_advance_lpf_envelope(C.byref(self.g), C.byref(self.env), _p(_f(curr)),
                                        _p(_f(env)))

    def den_zyxbd(self, rho, ik=1):
        self.L.orc_den_zyxbd(C.byref(self.g), _p(_f(rho)), ik)

    def env_lpf2_evolve(self, part, ebf, env, jc):
        """One env_lpf2_evolve; returns the surviving particles (new array)."""
        p = part.copy(order="F")
        ebfp = self.zeros_stage(p.shape[0])
        n = self.L.orc_env_lpf2_evolve(C.byref(self.g), C.byref(self.env), p.shape[0], _p(p),
                                       self._ld(p), _p(ebfp), _p(_f(ebf)), _p(_f(env)),
                                       _p(_f(jc)))
        return np.asfortranarray(p[:n])

    def grid_charge(self, part):
        rho = np.zeros((self.cfg.n1p, self.cfg.n2p, self.cfg.n3p), dtype=np.float64, order="F")
        self.L.orc_grid_charge(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(rho))
        return rho

    def energy_momenta(self, part):
        ek = np.zeros(7)
        emax = np.zeros(1)
        self.L.orc_energy_momenta(C.byref(self.g), part.shape[0], _p(_f(part)), self._ld(part), _p(ek), _p(emax))
        return ek, float(emax[0])

    def field_energy(self, ebf):
        out = np.zeros(6)
        self.L.orc_field_energy(C.byref(self.g), _p(_f(ebf)), _p(out))
        return out[: self.cfg.nfield]


class RefHarness:
    """The reference's own gather / push / Esirkepov / grid-charge routines."""

    def __init__(self, cfg, lorentz_fact: float | None = None, lib: Path | None = None):
        p = Path(lib) if lib else REF_LIB
        if not p.exists():
            raise ImportError(f"{p} missing: run `make -C oracle ref` where the reference exists")
        L = C.CDLL(str(p))
        L.ref_setup.argtypes = [C.c_int] * 4 + [C.c_double] * 5
        L.ref_gather.argtypes = [_I64, _D, _I64, _D, _D]
        L.ref_push.argtypes = [_I64, _D, _I64, _D, C.c_int]
        L.ref_esirkepov.argtypes = [_I64, _D, _I64, _D, _D]
        L.ref_grid_charge.argtypes = [_I64, _D, _I64, _D]
        L.ref_gather_push_deposit.argtypes = [_I64, _D, _I64, _D, _D, C.c_int]
        for f in (L.ref_setup, L.ref_gather, L.ref_push, L.ref_esirkepov, L.ref_grid_charge,
                  L.ref_gather_push_deposit):
            f.restype = None
        self.L = L
        self.cfg = cfg
        lf = cfg.lorentz_fact[0] if lorentz_fact is None else lorentz_fact
        L.ref_setup(cfg.ndim, cfg.nx, cfg.ny, cfg.nz, cfg.dx, cfg.dy, cfg.dz, cfg.dt, lf)

    def gather(self, part, ebf):
        ebfp = np.zeros((max(part.shape[0], 1), self.cfg.ncol), order="F")
        self.L.ref_gather(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(_f(ebf)), _p(ebfp))
        return ebfp

    def push(self, part, ebfp, initial=False):
        self.L.ref_push(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(ebfp), 1 if initial else 0)

    def esirkepov(self, part, ebfp, jc):
        self.L.ref_esirkepov(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(ebfp), _p(_f(jc)))

    def window_xshift(self, win, witr, init_iter, w_speed, ebf, part):
        """lp_window_xshift: returns (nshx, surviving particles as a new array);
        moves this oracle's box (xmn, xmx) with the window."""
        n = C.c_int64(part.shape[0])
        nshx = self.L.orc_window_xshift(C.byref(self.g), C.byref(win), witr, init_iter, w_speed,
                                        _p(_f(ebf)), C.byref(n), _p(_f(part)), self._ld(part))
        return int(nshx), np.asfortranarray(part[:n.value])

    def grid_charge(self, part):
        rho = np.zeros((self.cfg.n1p, self.cfg.n2p, self.cfg.n3p), order="F")
        self.L.ref_grid_charge(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(rho))
        return rho

    def _env_sigs(self):
        L = self.L
        L.ref_env_acc.argtypes = [_I64, _D, _I64, _D, _D, C.c_double, _D]
        L.ref_env_momenta.argtypes = [_I64, _D, _I64, _D]
        L.ref_env_density.argtypes = [_I64, _I64, _D, _D]
        L.ref_env_grad_interp.argtypes = [_I64, _D, _I64, _D, _D]
        L.ref_env_positions.argtypes = [_I64, _D, _I64, _D]
        for f in (L.ref_env_acc, L.ref_env_momenta, L.ref_env_density, L.ref_env_grad_interp,
                  L.ref_env_positions):
            f.restype = None

    def env_acc(self, part, ebf, av, ebfp=None):
        self._env_sigs()
        n = part.shape[0]
        ebfp = np.zeros((max(n, 1), self.cfg.ncol), order="F") if ebfp is None else ebfp
        self.L.ref_env_acc(n, _p(_f(part)), max(n, 1), _p(_f(ebf)), _p(_f(av)), self.cfg.dt,
                           _p(ebfp))
        return ebfp

    def env_momenta(self, part, ebfp):
        self._env_sigs()
        self.L.ref_env_momenta(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(ebfp))

    def env_density(self, ebfp, np_, den=None):
        self._env_sigs()
        c = self.cfg
        den = np.zeros((c.n1p, c.n2p, c.n3p), order="F") if den is None else den
        self.L.ref_env_density(np_, max(ebfp.shape[0], 1), _p(ebfp), _p(_f(den)))
        return den

    def env_grad_interp(self, part, av, ebfp):
        self._env_sigs()
        self.L.ref_env_grad_interp(part.shape[0], _p(_f(part)), max(part.shape[0], 1),
                                   _p(_f(av)), _p(ebfp))

    def env_positions(self, part, ebfp):
        self._env_sigs()
        self.L.ref_env_positions(part.shape[0], _p(_f(part)), max(part.shape[0], 1), _p(ebfp))

    @staticmethod
    def se_coeff(ndim, der_ord, cfl, yx_rat, lib: Path | None = None):
        """set_field_param's se_coeff(1:2) (grid_field_param.f90:36-99)."""
        p = Path(lib) if lib else REF_LIB
        if not p.exists():
            raise ImportError(f"{p} missing: run `make -C oracle ref` where the reference exists")
        L = C.CDLL(str(p))
        L.ref_se_coeff.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _D]
        L.ref_se_coeff.restype = None
        out = np.zeros(2)
        L.ref_se_coeff(ndim, der_ord, cfl, yx_rat, _p(out))
        return out

    @staticmethod
    def field_param(ndim, der_ord, cfl, yx_rat, lib: Path | None = None):
        """set_field_param's (hord_der2, opt_der1) (grid_field_param.f90:36-99)."""
        p = Path(lib) if lib else REF_LIB
        if not p.exists():
            raise ImportError(f"{p} missing: run `make -C oracle ref` where the reference exists")
        L = C.CDLL(str(p))
        L.ref_field_param.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _D]
        L.ref_field_param.restype = None
        out = np.zeros(2)
        L.ref_field_param(ndim, der_ord, cfl, yx_rat, _p(out))
        return out

    def gather_push_deposit(self, part, ebf, nsteps=1):
        jc = np.zeros(self.cfg.current_shape, order="F")
        self.L.ref_gather_push_deposit(part.shape[0], _p(_f(part)), max(part.shape[0], 1),
                                       _p(_f(ebf)), _p(jc), nsteps)
        return jc
