"""The oracle's envelope-model grid routines, pinned by what the mathematics
fixes (the reference modules that hold them, grid_fields.f90,
mpi_curr_interface.f90 and curr_and_fields_util.f90, need an MPI library to
compile, so they are not built here):

* env_grad (grid_fields.f90:426-586): the staggered gradient of a linear F
  is its slope everywhere, the one-sided forms at the faces included; for a
  quadratic F the interior differences are exact at the half-nodes and the
  one-sided right-face form 2F(i)-3F(i-1)+F(i-2) is exact at i+1/2 (taken
  at i = ix2-1 for x, at j = jy2 for y, as the reference writes them); F
  is extrapolated linearly into the z ghost planes.
* env_bds (grid_fields.f90:864-1067): x ghosts repeat the face value; y and
  z ghosts hold the quadratic extrapolation of the first three layers (exact
  for a quadratic profile) in both ghost layers, over the x (and y) ghosts
  written before.
* den_zyxbd (mpi_curr_interface.f90:646-720): per face, hand-worked: a unit
  in the first ghost layer lands in the second interior layer, which is
  then copied onto the first.
* env_maxw_solve / advance_lpf_envelope (grid_fields.f90:588-694,
  curr_and_fields_util.f90:262-299): one step maps a lattice plane wave
  A^{n-1} = e^{i theta}, A^n = g e^{i theta} onto
  A^{n+1} = (dt^2 sigma g + 2 g - (1 + i kappa)) / (1 - i kappa) e^{i theta}
  with kappa = om0 dt and sigma the scheme's Fourier symbol (Laplacian,
  -ompe chi and the centred 2 om0 d/dx term), derived here independently
  of the stencils.
* env_lpf2_evolve (env_evolve_in_time.f90:196-361): the step's Esirkepov
  current keeps the discrete continuity equation with the grid charge of
  the particles before and after it.
"""
import numpy as np
import pytest

from aladyn_amd.config import LPConfig
from aladyn_amd.inputs import pack_weight
from cases import random_particles
from oracle.pyoracle import Oracle


def _cfg(ndim, der_ord=2, nx=24, ny=16, nz=12):
    if ndim == 2:
        return LPConfig(ndim=2, nx=nx, ny=ny, k0=20.0, yx_rat=2.0, cfl=0.8, der_ord=der_ord)
    return LPConfig(ndim=3, nx=nx, ny=ny, nz=nz, k0=20.0, yx_rat=3.0, zx_rat=2.0, cfl=0.8,
                    der_ord=der_ord)


def _node_coords(cfg):
    """Node coordinates x(i) = dx*(i-3) etc. on the whole (n1p, n2p, n3p) array."""
    i = np.arange(1, cfg.n1p + 1) - 3.0
    j = np.arange(1, cfg.n2p + 1) - 3.0
    k = np.arange(1, cfg.n3p + 1) - 3.0 if cfg.ndim == 3 else np.zeros(1)
    return np.meshgrid(i * cfg.dx, j * cfg.dy, k * cfg.dz, indexing="ij")


def _interior(cfg):
    s = (slice(cfg.ix1 - 1, cfg.ix2), slice(cfg.jy1 - 1, cfg.jy2),
         slice(cfg.kz1 - 1, cfg.kz2))
    return s


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("der_ord", [2, 4])
def test_env_grad_of_a_linear_profile_is_its_slope(ndim, der_ord):
    cfg = _cfg(ndim, der_ord)
    orc = Oracle(cfg)
    X, Y, Z = _node_coords(cfg)
    a, b, c = 0.7, -1.3, 0.4
    av = orc.zeros_envelope()
    av[..., 0] = 2.0 + a * X + b * Y + (c * Z if ndim == 3 else 0.0)
    orc.env_grad(av)
    s = _interior(cfg)
    for d, slope in enumerate((a, b, c)[: cfg.curr_ndim]):
        np.testing.assert_allclose(av[s + (d + 1,)], slope, rtol=1e-11, atol=1e-11)
    if ndim == 3:
        # the ghost planes kz1-1 and kz2+1 hold the linear continuation
        si, sj = s[0], s[1]
        want = 2.0 + a * X + b * Y + c * Z
        for k in (cfg.kz1 - 2, cfg.kz2):
            np.testing.assert_allclose(av[si, sj, k, 0], want[si, sj, k], rtol=1e-13)


@pytest.mark.parametrize("ndim", [2, 3])
def test_env_grad_of_a_quadratic_profile_at_the_half_nodes(ndim):
    cfg = _cfg(ndim, 2)
    orc = Oracle(cfg)
    X, Y, Z = _node_coords(cfg)
    av = orc.zeros_envelope()
    av[..., 0] = X * X + 3.0 * Y * Y
    orc.env_grad(av)
    s = _interior(cfg)
    # d/dx of x^2 at x(i) + dx/2 on the interior nodes up to ix2 - 1; at ix2
    # the reference's one-sided form 2F(i)-3F(i-1)+F(i-2) is taken at
    # i = ix2 - 1 (grid_fields.f90:478-480), so it repeats d/dx at ix2 - 1/2
    want = 2.0 * (X + 0.5 * cfg.dx)
    want[cfg.ix2 - 1] = want[cfg.ix2 - 2]
    np.testing.assert_allclose(av[s + (1,)], want[s], rtol=1e-11, atol=1e-11)
    # d/dy of 3 y^2 at y(j) + dy/2, the right face included
    np.testing.assert_allclose(av[s + (2,)], 6.0 * (Y + 0.5 * cfg.dy)[s], rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("ndim", [2, 3])
def test_env_bds_extends_each_face(ndim):
    cfg = _cfg(ndim)
    orc = Oracle(cfg)
    X, Y, Z = _node_coords(cfg)
    rng = np.random.default_rng(3)
    env = orc.zeros_env_fields()
    # components 1..2: x-random, quadratic in y and z (the extrapolation is
    # exact for quadratics); components 3..4 must be left alone
    fx = rng.normal(size=cfg.n1p)[:, None, None]
    prof = fx * (1.0 + 0.3 * Y - 0.2 * Y * Y) * ((1.0 - 0.5 * Z + 0.25 * Z * Z) if ndim == 3 else 1.0)
    env[..., 0] = prof
    env[..., 1] = -2.0 * prof
    env[..., 2:] = rng.normal(size=env[..., 2:].shape)
    before = env.copy()
    orc.env_bds(env, 2, 2, 1, 2)
    assert np.array_equal(env[..., 2:], before[..., 2:])
    i1, i2, j1, j2 = cfg.ix1 - 1, cfg.ix2 - 1, cfg.jy1 - 1, cfg.jy2 - 1
    k1, k2 = cfg.kz1 - 1, cfg.kz2 - 1
    ks = slice(k1, k2 + 1)
    for c in (0, 1):
        # x ghosts (interior y, z) repeat the face value
        for g in (i1 - 1, i1 - 2):
            assert np.array_equal(env[g, j1:j2 + 1, ks, c], env[i1, j1:j2 + 1, ks, c])
        for g in (i2 + 1, i2 + 2):
            assert np.array_equal(env[g, j1:j2 + 1, ks, c], env[i2, j1:j2 + 1, ks, c])
        # y ghosts over the whole x range (the x ghosts included): the value
        # of the quadratic one layer out, in both layers
        xs = slice(i1 - 2, i2 + 3)
        ext = env[:, :, :, c].copy()
        ext[xs] = np.broadcast_to(ext[np.clip(np.arange(cfg.n1p), i1, i2)][xs], ext[xs].shape)
        sgn = 1.0 if c == 0 else -2.0
        exact_lo = (sgn * fx[np.clip(np.arange(cfg.n1p), i1, i2)] *
                    (1.0 + 0.3 * (Y - cfg.dy) - 0.2 * (Y - cfg.dy) ** 2) *
                    ((1.0 - 0.5 * Z + 0.25 * Z * Z) if ndim == 3 else 1.0))
        for g in (j1 - 1, j1 - 2):
            np.testing.assert_allclose(env[xs, g, ks, c], exact_lo[xs, j1, ks], rtol=1e-12,
                                       atol=1e-12)
        exact_hi = (sgn * fx[np.clip(np.arange(cfg.n1p), i1, i2)] *
                    (1.0 + 0.3 * (Y + cfg.dy) - 0.2 * (Y + cfg.dy) ** 2) *
                    ((1.0 - 0.5 * Z + 0.25 * Z * Z) if ndim == 3 else 1.0))
        for g in (j2 + 1, j2 + 2):
            np.testing.assert_allclose(env[xs, g, ks, c], exact_hi[xs, j2, ks], rtol=1e-12,
                                       atol=1e-12)
        if ndim == 3:
            # z ghosts over the x and y ghosts too
            ys = slice(j1 - 2, j2 + 3)
            for g in (k1 - 1, k1 - 2):
                want = 3 * env[xs, ys, k1, c] - 3 * env[xs, ys, k1 + 1, c] + env[xs, ys, k1 + 2, c]
                np.testing.assert_allclose(env[xs, ys, g, c], want, rtol=1e-12, atol=1e-12)
            # exact for the quadratic z profile on the interior x, y
            zq = lambda z: 1.0 - 0.5 * z + 0.25 * z * z  # noqa: E731
            zk = (k1 - 3) * cfg.dz  # the ghost layer kz1-1 (0-based k1-1)
            ratio = env[i1:i2 + 1, j1:j2 + 1, k1 - 1, c] / env[i1:i2 + 1, j1:j2 + 1, k1, c]
            np.testing.assert_allclose(ratio, zq(zk) / zq(zk + cfg.dz), rtol=1e-10)


@pytest.mark.parametrize("ndim", [2, 3])
def test_den_zyxbd_per_face(ndim):
    cfg = _cfg(ndim, nx=10, ny=8, nz=8)
    orc = Oracle(cfg)
    i1, i2, j1, j2 = cfg.ix1 - 1, cfg.ix2 - 1, cfg.jy1 - 1, cfg.jy2 - 1
    k1, k2 = cfg.kz1 - 1, cfg.kz2 - 1
    km = (k1 + k2) // 2 if ndim == 3 else 0
    jm, im = (j1 + j2) // 2, (i1 + i2) // 2
    cases = [
        ((i1 - 1, jm, km), [(i1 + 1, jm, km), (i1, jm, km)]),
        ((i2 + 1, jm, km), [(i2 - 1, jm, km), (i2, jm, km)]),
        ((im, j1 - 1, km), [(im, j1 + 1, km), (im, j1, km)]),
        ((im, j2 + 1, km), [(im, j2 - 1, km), (im, j2, km)]),
    ]
    if ndim == 3:
        cases += [((im, jm, k1 - 1), [(im, jm, k1 + 1), (im, jm, k1)]),
                  ((im, jm, k2 + 1), [(im, jm, k2 - 1), (im, jm, k2)])]
    for src, dst in cases:
        rho = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 1), order="F")
        rho[src + (0,)] = 1.0
        orc.den_zyxbd(rho, 1)
        for d in dst:
            assert rho[d + (0,)] == 1.0, (src, d)
        # the ghost keeps its value; nothing else changes
        rest = rho.copy()
        for d in dst:
            rest[d + (0,)] = 0.0
        rest[src + (0,)] = 0.0
        assert not rest.any(), src
        assert rho[src + (0,)] == 1.0
    # the second interior layer's own value is added to, and then copied
    rho = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, 1), order="F")
    rho[i1 + 1, jm, km, 0] = 2.0
    rho[i1, jm, km, 0] = 5.0
    rho[i1 - 1, jm, km, 0] = 0.5
    orc.den_zyxbd(rho, 1)
    assert rho[i1 + 1, jm, km, 0] == 2.5 and rho[i1, jm, km, 0] == 2.5


def _symbol(cfg, orc, kx, ky, kz, c0):
    """sigma: Fourier symbol of S(A) = c0 A + Laplacian A + first_ader."""
    e = orc.env
    der = cfg.der_ord
    cfx = (1.0, 0.0) if der <= 2 else (1.0 - 4.0 * e.hord_der2, e.hord_der2)
    cft = (4.0 / 3.0, -1.0 / 12.0) if der == 4 else (1.0, 0.0)

    def lap(k, h, cf):
        return (cf[0] * (2 * np.cos(k * h) - 2) + cf[1] * (2 * np.cos(2 * k * h) - 2)) / h ** 2

    s = c0 + lap(kx, cfg.dx, cfx) + lap(ky, cfg.dy, cft)
    if cfg.ndim == 3:
        s = s + lap(kz, cfg.dz, cft)
    dx1 = e.om0 / cfg.dx
    if der < 3:
        a1, a2 = dx1, 0.0
    elif der == 3:
        a1, a2 = dx1 * e.opt_der1, dx1 * 0.5 * (1 - e.opt_der1)
    else:
        a1, a2 = 1.0, 0.0  # the reference leaves aph_opt = (1, 0) for der_ord 4
    # i a1 (e^{ikh} - e^{-ikh}) + i a2 (e^{2ikh} - e^{-2ikh})
    return s - 2 * a1 * np.sin(kx * cfg.dx) - 2 * a2 * np.sin(2 * kx * cfg.dx)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("der_ord", [2, 3, 4])
def test_env_maxw_solve_maps_a_plane_wave_by_its_symbol(ndim, der_ord):
    cfg = _cfg(ndim, der_ord)
    orc = Oracle(cfg)
    X, Y, Z = _node_coords(cfg)
    kx, ky, kz = 2 * np.pi / (9.3 * cfg.dx), 2 * np.pi / (7.1 * cfg.dy), 2 * np.pi / (6.4 * cfg.dz)
    th = kx * X + ky * Y + (kz * Z if ndim == 3 else 0.0)
    g = 0.9 * np.exp(0.37j)
    chi = 0.8
    c0 = -cfg.ompe * chi
    a_old, a_now = np.exp(1j * th), g * np.exp(1j * th)
    env = orc.zeros_env_fields()
    env[..., 0], env[..., 1] = a_now.real, a_now.imag
    env[..., 2], env[..., 3] = a_old.real, a_old.imag
    D = cfg.curr_ndim
    jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, D + 1), order="F")
    jc[..., 2] = chi
    # (a) the solver alone, exact ghosts: every interior node
    e1, j1 = env.copy(order="F"), jc.copy(order="F")
    s = _interior(cfg)
    j1[s + (0,)] = c0 * a_now.real[s]
    j1[s + (1,)] = c0 * a_now.imag[s]
    orc.env_maxw_solve(j1, e1)
    kap = orc.env.om0 * cfg.dt
    sig = _symbol(cfg, orc, kx, ky, kz, c0)
    a_new = (cfg.dt ** 2 * sig * g + 2 * g - (1 + 1j * kap)) / (1 - 1j * kap) * np.exp(1j * th)
    np.testing.assert_allclose(e1[s + (0,)], a_new.real[s], rtol=0, atol=1e-12)
    np.testing.assert_allclose(e1[s + (1,)], a_new.imag[s], rtol=0, atol=1e-12)
    np.testing.assert_allclose(e1[s + (2,)], a_now.real[s], rtol=0, atol=0)
    np.testing.assert_allclose(e1[s + (3,)], a_now.imag[s], rtol=0, atol=0)
    # (b) advance_lpf_envelope: env_bds and the -ompe chi A source first;
    # nodes whose stencil reaches no ghost are the same
    e2, j2 = env.copy(order="F"), jc.copy(order="F")
    orc.advance_lpf_envelope(j2, e2)
    w = 2
    inner = (slice(cfg.ix1 - 1 + w, cfg.ix2 - w), slice(cfg.jy1 - 1 + w, cfg.jy2 - w),
             slice(cfg.kz1 - 1 + w, cfg.kz2 - w) if ndim == 3 else slice(0, 1))
    np.testing.assert_allclose(e2[inner + (0,)], a_new.real[inner], rtol=0, atol=1e-12)
    np.testing.assert_allclose(e2[inner + (1,)], a_new.imag[inner], rtol=0, atol=1e-12)


def _gaussian_envelope(cfg, orc, a0=0.8, phase=0.3):
    X, Y, Z = _node_coords(cfg)
    xc, yc = 0.5 * cfg.nx * cfg.dx, 0.0
    w = 0.35 * cfg.ny * cfg.dy
    r2 = (Y - yc) ** 2 + (Z ** 2 if cfg.ndim == 3 else 0.0)
    prof = a0 * np.exp(-((X - xc) / (0.2 * cfg.nx * cfg.dx)) ** 2 - r2 / w ** 2)
    env = orc.zeros_env_fields()
    env[..., 0] = prof * np.cos(phase)
    env[..., 1] = prof * np.sin(phase)
    env[..., 2] = prof
    return env


@pytest.mark.parametrize("ndim", [2, 3])
def test_env_lpf2_evolve_keeps_continuity(ndim):
    cfg = _cfg(ndim, 2, nx=24, ny=12, nz=10)
    orc = Oracle(cfg)
    part = random_particles(cfg, 4000, seed=11, pmax=0.5, margin=3.0)
    part[:, -1] = pack_weight(-1, 1, np.full(part.shape[0], 0.37, np.float32))
    ebf = np.zeros(cfg.field_shape, order="F")
    env = _gaussian_envelope(cfg, orc)
    jc = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p, cfg.curr_ndim + 1), order="F")
    rho0 = orc.grid_charge(part)
    p1 = orc.env_lpf2_evolve(part, ebf, env, jc)
    assert p1.shape[0] == part.shape[0]  # particles 3 cells in stay in the box
    rho1 = orc.grid_charge(p1)
    # jc now holds the normalised current (djc applied by curr_mpi_collect):
    # undo it to get the raw Esirkepov fluxes
    D = cfg.curr_ndim
    raw = [jc[..., d] / orc.g.djc[d] for d in range(D)]
    res = rho1 - rho0
    for d in range(D):
        res = res + (raw[d] - np.roll(raw[d], 1, axis=d)) / D
    s = (slice(cfg.ix1 + 1, cfg.ix2 - 2), slice(cfg.jy1 + 1, cfg.jy2 - 2),
         slice(cfg.kz1 + 1, cfg.kz2 - 2) if ndim == 3 else slice(0, 1))
    scale = np.max(np.abs(rho1 - rho0)[s])
    assert scale > 0
    assert np.max(np.abs(res[s])) <= 1e-12 * scale
    # the envelope advanced: A^{n-1} now holds the old A^n on the interior
    env0 = _gaussian_envelope(cfg, orc)
    si = _interior(cfg)
    assert np.array_equal(env[si + (2,)], env0[si + (0,)])
    assert np.array_equal(env[si + (3,)], env0[si + (1,)])
    assert not np.array_equal(env[si + (0,)], env0[si + (0,)])
