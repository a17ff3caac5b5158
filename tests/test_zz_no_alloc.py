"""aladyn_grid_charge_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against aladyn_grid_charge, through the
C-ABI.  Both deposit set_grid_charge (grid_part_util.f90:240-335) with FP64
device atomics, so they agree to the rounding of a different summation order:
<= 1e-13 of max|rho|.  Two calls in a row, the second on a smaller grid, share
the library's one device array; a grid beyond that array is refused with
ALADYN_EINVAL and a message."""
import ctypes as C

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from cases import random_particles, small_2d, small_3d
from conftest import rel_maxerr

pytestmark = pytest.mark.gpu


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _both(cfg, part, prec, seed):
    """(aladyn_grid_charge, aladyn_grid_charge_noalloc) of the same store,
    each added to the same random starting rho."""
    rng = np.random.default_rng(seed)
    rho0 = np.asfortranarray(rng.normal(0.0, 1e-3, (cfg.n1p, cfg.n2p, cfg.n3p)))
    with LPStep(cfg, prec) as gpu:
        gpu.put_species(0, part)
        ref = rho0.copy(order="F")
        assert gpu.lib.aladyn_grid_charge(gpu._h, 0, _dptr(ref)) == 0
        got = rho0.copy(order="F")
        rc = gpu.lib.aladyn_grid_charge_noalloc(gpu._h, 0, _dptr(got))
        assert rc == 0, gpu.lib.aladyn_last_error(gpu._h)
    return ref, got


@pytest.mark.parametrize("prec", [PREC_F64, PREC_F32])
def test_noalloc_grid_charge_matches_allocating_call(prec):
    print(f"noalloc grid charge, precision {prec}: 3-D 24x16x16 then 2-D 48x32", flush=True)
    cfg3, part3, _ = small_3d()
    ref, got = _both(cfg3, part3, prec, seed=5)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert rel_maxerr(got.sum(), ref.sum()) <= 1e-13
    # second call: a smaller grid on the same device array, whose first
    # nodes still hold the 3-D call's values
    cfg2, part2, _ = small_2d()
    assert cfg2.n1p * cfg2.n2p * cfg2.n3p < cfg3.n1p * cfg3.n2p * cfg3.n3p
    ref, got = _both(cfg2, part2, prec, seed=6)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_noalloc_grid_charge_refuses_a_grid_beyond_its_array():
    from aladyn_amd.config import LPConfig

    print("noalloc grid charge: 2-D 4096x2048 grid (8.4M nodes > 2^23)", flush=True)
    cfg = LPConfig(ndim=2, nx=4096, ny=2048, k0=20.0, cfl=0.9, n_over_nc=0.01)
    assert cfg.n1p * cfg.n2p * cfg.n3p > 1 << 23
    part = random_particles(cfg, 1000, seed=7)
    with LPStep(cfg, PREC_F32) as gpu:
        gpu.put_species(0, part)
        rho = np.zeros((cfg.n1p, cfg.n2p, cfg.n3p), order="F")
        rc = gpu.lib.aladyn_grid_charge_noalloc(gpu._h, 0, _dptr(rho))
        assert rc == 1  # ALADYN_EINVAL
        msg = gpu.lib.aladyn_last_error(gpu._h).decode()
        assert "exceed" in msg, msg
        assert not rho.any()
