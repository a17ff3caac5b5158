"""The oracle's bunch-field initialisation (init_beam_part_distrib.f90:538-602
with fft_3d_psolv / fft_2d_psolv, psolve.f90:117-247, and
initial_beam_fields, init_grid_fields.f90:32-131), pinned by what the
mathematics fixes.  psolve.f90 needs FFTW and the MPI modules, so the
reference is not built here.

* The sine transform is FFTW_RODFT00: scipy.fft.dst(type=1) has the same
  definition, y_k = 2 sum_j x_j sin(pi (j+1)(k+1)/(n+1)).
* The Poisson solve maps one separable sine mode
  rho = sin(pi a i/(nx+1)) sin(pi b j/(ny+1)) sin(pi c k/(nz+1)) onto
  ompe (nx+1)/(nx+2) rho / (sky(b)^2 + skz(c)^2 + skx(a)^2/gam2): the forward
  RODFT00 of the mode is (n+1) at index a-1, the inverse of that is
  2 (n+1) times the mode, so y and z come back with 1/(2(n+1)) each and x
  with the reference's 1/(2(nx+2)) (its x call passes n1+1,
  modern_fft_lib.F90:689).  The wavenumbers are 2 sin(h k/2)/h with
  k = pi m/L, L = (nx-1)dx, ny dy, nz dz (set_grid_param.f90:76-78, 104-108,
  959-978).  In 2-D only y is transformed and rho/sky(b)^2 is kept per x.
* initial_beam_fields of a quadratic potential with a cross term: E is
  minus the gradient at the staggered points (exact for a quadratic:
  Ex = -dphi/dx at i+1/2 over gam2, Ey at j+1/2, Ez at k+1/2), Bz = bet Ey
  and By = -bet Ez averaged in x; the right x face copies phi into its
  ghost (so Ex = 0 at ix2) and leaves Ez unset at ix2+1 (so By at ix2
  averages Ez with 0), and the y and z ghosts are extrapolated (quadratically
  in y, linearly in z, so Ez at kz2 is the linear extrapolation's).
"""
import math

import numpy as np
import pytest
import scipy.fft

from aladyn_amd.config import LPConfig
from oracle.pyoracle import Oracle


def _cfg(ndim):
    if ndim == 3:
        return LPConfig(ndim=3, nx=24, ny=16, nz=12, k0=20.0, yx_rat=3.0, zx_rat=2.0, cfl=0.8,
                        n_over_nc=0.01)
    return LPConfig(ndim=2, nx=40, ny=24, k0=20.0, yx_rat=2.0, cfl=0.8, n_over_nc=0.01)


@pytest.mark.parametrize("n", [1, 2, 7, 64])
def test_sine_transform_is_rodft00(n):
    v = np.random.default_rng(n).normal(size=n)
    orc = Oracle(_cfg(2))
    assert np.max(np.abs(orc.sine_transform(v) - scipy.fft.dst(v, type=1))) <= 1e-13 * n


def _sk(m, h, length):
    return 2.0 * math.sin(0.5 * h * math.pi * m / length) / h


@pytest.mark.parametrize("ndim,modes,gam2", [(3, (1, 1, 1), 4.0), (3, (3, 2, 5), 1.0e6),
                                             (3, (7, 1, 2), 2.25), (2, (2, 3, 0), 9.0)])
def test_poisson_solve_of_one_sine_mode(ndim, modes, gam2):
    cfg = _cfg(ndim)
    orc = Oracle(cfg)
    a, b, c = modes
    i = np.arange(1, cfg.nx + 1)[:, None, None]
    j = np.arange(1, cfg.ny + 1)[None, :, None]
    k = np.arange(1, (cfg.nz if ndim == 3 else 1) + 1)[None, None, :]
    mode = np.sin(math.pi * a * i / (cfg.nx + 1)) * np.sin(math.pi * b * j / (cfg.ny + 1))
    if ndim == 3:
        mode = mode * np.sin(math.pi * c * k / (cfg.nz + 1))
    rho = np.random.default_rng(3).normal(size=(cfg.n1p, cfg.n2p, cfg.n3p))  # ghosts kept
    inner = (slice(2, 2 + cfg.nx), slice(2, 2 + cfg.ny), slice(2, 2 + cfg.nz) if ndim == 3
             else slice(0, 1))
    rho[inner] = mode
    pot = orc.beam_potential(rho, gam2)
    ky = _sk(b, cfg.dy, cfg.ny * cfg.dy)
    if ndim == 3:
        kx = _sk(a, cfg.dx, (cfg.nx - 1) * cfg.dx)
        kz = _sk(c, cfg.dz, cfg.nz * cfg.dz)
        want = cfg.ompe * (cfg.nx + 1) / (cfg.nx + 2) / (ky * ky + kz * kz + kx * kx / gam2) * mode
    else:
        want = cfg.ompe / (ky * ky) * mode
    assert np.max(np.abs(pot[inner] - want)) <= 1e-12 * np.max(np.abs(want))
    outside = np.ones(rho.shape, bool)
    outside[inner] = False
    assert np.array_equal(pot[outside], rho[outside])


def test_initial_beam_fields_of_a_quadratic_potential_3d():
    cfg = _cfg(3)
    orc = Oracle(cfg)
    gam2, bet = 16.0, math.sqrt(1.0 - 1.0 / 16.0)
    # node (i, j, k) at x = (i-3) dx, y = (j-3) dy, z = (k-3) dz
    x = (np.arange(cfg.n1p) - 2.0)[:, None, None] * cfg.dx
    y = (np.arange(cfg.n2p) - 2.0)[None, :, None] * cfg.dy
    z = (np.arange(cfg.n3p) - 2.0)[None, None, :] * cfg.dz
    A, B, Cz, D, E = 0.3, -0.7, 0.45, 0.2, -0.35
    phi = A * x * x + B * y * y + Cz * z * z + D * x * y + E * y * z + 0.1 * x - 0.2 * z
    efb = orc.initial_beam_fields(phi, gam2, bet)
    i1, i2, j1, j2, k1, k2 = 2, 2 + cfg.nx - 1, 2, 2 + cfg.ny - 1, 2, 2 + cfg.nz - 1
    I, J, K = slice(i1, i2), slice(j1, j2 + 1), slice(k1, k2 + 1)  # i up to ix2-1
    xh, yh, zh = x + 0.5 * cfg.dx, y + 0.5 * cfg.dy, z + 0.5 * cfg.dz
    ex = -(2 * A * xh + D * y + 0.1) / gam2 + 0 * z
    ey = -(2 * B * yh + D * x + E * z) + 0 * x
    ez = -(2 * Cz * zh + E * y - 0.2) + 0 * x
    tol = 1e-12 * np.max(np.abs(phi))
    assert np.max(np.abs(efb[I, J, K, 0] - ex[I, J, K])) <= tol
    # Ey is the staggered difference wherever both nodes hold phi; at
    # i = ix2+1 the x ghost holds phi(ix2) (and the corner j = jy2+1 the
    # node's own value)
    assert np.max(np.abs(efb[i1:i2 + 1, J, K, 1] - ey[i1:i2 + 1, J, K])) <= tol
    # Ez at k < kz2 is the staggered difference; at kz2 the linear
    # extrapolation of the z ghost gives phi(kz2) - phi(kz2-1)
    assert np.max(np.abs(efb[i1:i2 + 1, J, k1:k2, 2] - ez[i1:i2 + 1, J, k1:k2])) <= tol
    ez_k2 = -(phi[i1:i2 + 1, J, k2] - phi[i1:i2 + 1, J, k2 - 1]) / cfg.dz
    assert np.max(np.abs(efb[i1:i2 + 1, J, k2, 2] - ez_k2)) <= tol
    # the right x face: Ex = 0 (phi copied into the ghost), By averages Ez
    # with the unset 0 at ix2+1
    assert np.max(np.abs(efb[i2, J, K, 0])) == 0.0
    assert np.max(np.abs(efb[i2, J, K, 4] + bet * 0.5 * efb[i2, J, K, 2])) <= tol
    # Bz = bet <Ey>_x, By = -bet <Ez>_x inside
    bz = 0.5 * bet * (efb[I, J, K, 1] + efb[i1 + 1:i2 + 1, J, K, 1])
    assert np.max(np.abs(efb[I, J, K, 5] - bz)) <= tol
    by = -0.5 * bet * (efb[I, J, K, 2] + efb[i1 + 1:i2 + 1, J, K, 2])
    assert np.max(np.abs(efb[I, J, K, 4] - by)) <= tol
    # Bz at ix2 averages with Ey(ix2+1) = the difference of the x ghost
    ey_g = -(phi[i2, j1 + 1:j2 + 1, K] - phi[i2, j1:j2, K]) / cfg.dy
    assert np.max(np.abs(efb[i2 + 1, j1:j2, K, 1] - ey_g)) <= tol


def test_initial_beam_fields_of_a_quadratic_potential_2d():
    cfg = _cfg(2)
    orc = Oracle(cfg)
    gam2, bet = 9.0, math.sqrt(1.0 - 1.0 / 9.0)
    x = (np.arange(cfg.n1p) - 2.0)[:, None, None] * cfg.dx
    y = (np.arange(cfg.n2p) - 2.0)[None, :, None] * cfg.dy
    A, B, D = -0.4, 0.6, 0.25
    phi = A * x * x + B * y * y + D * x * y + 0.3 * y
    efb = orc.initial_beam_fields(phi, gam2, bet)
    i1, i2, j1, j2 = 2, 2 + cfg.nx - 1, 2, 2 + cfg.ny - 1
    I, J = slice(i1, i2), slice(j1, j2 + 1)
    xh, yh = x + 0.5 * cfg.dx, y + 0.5 * cfg.dy
    ex = -(2 * A * xh + D * y) / gam2
    ey = -(2 * B * yh + D * x + 0.3)
    tol = 1e-12 * np.max(np.abs(phi))
    assert np.max(np.abs(efb[I, J, 0, 0] - ex[I, J, 0])) <= tol
    assert np.max(np.abs(efb[i1:i2 + 1, J, 0, 1] - ey[i1:i2 + 1, J, 0])) <= tol
    assert np.max(np.abs(efb[i2, J, 0, 0])) == 0.0
    bz = 0.5 * bet * (efb[I, J, 0, 1] + efb[i1 + 1:i2 + 1, J, 0, 1])
    assert np.max(np.abs(efb[I, J, 0, 2] - bz)) <= tol


@pytest.mark.parametrize("ndim", [2, 3])
def test_beam_fields_add_the_bunch_fields_with_bx_zero(ndim):
    cfg = _cfg(ndim)
    orc = Oracle(cfg)
    gam2 = 25.0
    bet = math.sqrt(1.0 - 1.0 / gam2)
    rng = np.random.default_rng(9)
    rho = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p))
    rho[4:-4, 4:-4, (slice(4, -4) if ndim == 3 else 0)] = rng.normal(
        size=rho[4:-4, 4:-4, (slice(4, -4) if ndim == 3 else 0)].shape)
    ebf = np.asfortranarray(rng.normal(size=cfg.field_shape))
    ebf0 = ebf.copy()
    orc.beam_fields(rho, ebf, gam2, bet)
    efb = orc.initial_beam_fields(orc.beam_potential(rho, gam2), gam2, bet)
    comps = [0, 1, 2] if ndim == 2 else [0, 1, 2, 4, 5]
    for c_ebf, c_efb in zip(range(cfg.nfield) if ndim == 2 else [0, 1, 2, 4, 5], comps):
        assert np.max(np.abs(ebf[..., c_ebf] - ebf0[..., c_ebf] - efb[..., c_efb])) <= 1e-14 * (
            1 + np.max(np.abs(efb[..., c_efb])))
    if ndim == 3:
        assert np.array_equal(ebf[..., 3], ebf0[..., 3])
