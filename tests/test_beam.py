"""The bunch fields at the start of a run on the device (aladyn_beam_fields:
set_grid_charge of the bunch, the sine-transform Poisson solve of
fft_3d_psolv / fft_2d_psolv and initial_beam_fields,
init_beam_part_distrib.f90:538-602) against the oracle's restatement
(orc_beam_fields, pinned in test_oracle_beam.py).

A bunch of random particles (charge -1, FP32 weights) inside the box, random
fields already on the grid; the fields the call adds are compared in
max-norm relative error: FP64 <= 1e-12; FP32 (the FP64 sum rounded into the
FP32 fields once) <= 1e-6 of max|ebf|.
"""
import math

import numpy as np
import pytest

from aladyn_amd import PREC_F32, PREC_F64, LPStep
from aladyn_amd.config import LPConfig
from cases import random_fields, random_particles
from conftest import rel_maxerr
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def _cfg(ndim):
    if ndim == 3:
        return LPConfig(ndim=3, nx=40, ny=24, nz=20, k0=20.0, yx_rat=3.0, zx_rat=2.0, cfl=0.8,
                        n_over_nc=0.01)
    return LPConfig(ndim=2, nx=64, ny=40, k0=20.0, yx_rat=2.0, cfl=0.8, n_over_nc=0.01)


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("precision", [PREC_F64, PREC_F32])
def test_beam_fields_match_oracle(ndim, precision):
    print(f"beam fields {ndim}-D precision={precision}", flush=True)
    cfg = _cfg(ndim)
    bunch = random_particles(cfg, 3000, seed=21, pmax=0.0, margin=6.0)
    gam2 = 400.0
    bet0 = math.sqrt(1.0 - 1.0 / gam2)
    ebf = random_fields(cfg, seed=22, scale=0.01)
    orc = Oracle(cfg)
    want = ebf.copy(order="F")
    orc.beam_fields(orc.grid_charge(bunch), want, gam2, bet0)
    with LPStep(cfg, precision) as gpu:
        gpu.put_species(0, bunch)
        gpu.put_fields(ebf)
        before = gpu.get_fields()
        gpu.beam_fields(0, gam2, bet0)
        got = gpu.get_fields()
    added_want = want - ebf
    assert np.max(np.abs(added_want)) > 0.0
    if precision == PREC_F64:
        err = rel_maxerr(got - before, added_want)
        print(f"added fields: {err:.2e}", flush=True)
        assert err <= 1e-12
    else:
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        print(f"fields: {err:.2e}", flush=True)
        assert err <= 1e-6
    if ndim == 3:  # Bx is left as it was
        assert np.array_equal(got[..., 3], before[..., 3])


def test_beam_fields_reject_bad_arguments():
    from aladyn_amd._lib import AladynError

    cfg = _cfg(2)
    with LPStep(cfg, PREC_F64) as gpu:
        gpu.put_species(0, random_particles(cfg, 100, seed=23))
        with pytest.raises(AladynError, match="ESTATE"):
            gpu.beam_fields(0, 4.0, 0.5)
        gpu.put_fields(random_fields(cfg))
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.beam_fields(0, 0.5, 0.5)
        with pytest.raises(AladynError, match="EINVAL"):
            gpu.beam_fields(0, 4.0, 1.0)
