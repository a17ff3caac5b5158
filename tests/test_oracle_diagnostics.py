"""Pins for the oracle functions the reference's own routines cannot check here
(their modules need an MPI library): energy_momenta, the Envar field energy and
the open-boundary removal of cell_part_dist.

Each case is built so the expected value is exact in binary floating point
and so a plausible slip fails it: weighting the 3-D momentum sums (the
reference sums unweighted p in 3-D and weighted p in 2-D,
diag_part_and_fields.f90:193 vs 211), a sign of the angular momentum, a ghost
layer counted in the field energy, the 2-D volume dx*dy*dy (Envar,
diag_part_and_fields.f90:506-507), or the (gmin, gmax] interval of the
removal (mpi_part_interface.f90:494-498).
"""
import numpy as np

from aladyn_amd.config import LPConfig
from aladyn_amd.inputs import pack_weight
from oracle.pyoracle import Oracle


def _cfg2():
    return LPConfig(ndim=2, nx=16, ny=12, k0=20.0, yx_rat=1.0, cfl=0.9, n_over_nc=0.01)


def _cfg3():
    return LPConfig(ndim=3, nx=12, ny=8, nz=8, k0=20.0, yx_rat=4.0, zx_rat=4.0, cfl=0.9,
                    n_over_nc=0.005)


def test_energy_momenta_2d_weighted_sums():
    cfg = _cfg2()
    orc = Oracle(cfg)
    part = np.zeros((3, 5), order="F")
    part[:, 0] = cfg.xmn + 4.5 * cfg.dx
    part[:, 1] = cfg.ymn + 5.5 * cfg.dy
    # gamma - 1 = 1/4, 2/3, 0 exactly (3-4-5 and 4/3-1-5/3 triangles)
    part[:, 2] = [0.75, 0.0, 0.0]
    part[:, 3] = [0.0, -4.0 / 3.0, 0.0]
    w = np.array([0.5, 0.25, 0.125], dtype=np.float32)
    part[:, 4] = pack_weight(-1, 0, w)
    ek, ekmax = orc.energy_momenta(part)
    assert ek[0] == 0.5 * 0.75                      # sum w px
    assert ek[1] == 0.25 * (-4.0 / 3.0)             # sum w py
    assert ek[5] == -(0.5 + 0.25 + 0.125)           # sum charge * w
    assert abs(ek[6] - (0.5 * 0.25 + 0.25 * (2.0 / 3.0))) < 1e-15   # sum w (gamma - 1)
    assert abs(ekmax - 2.0 / 3.0) < 1e-15
    assert ek[2] == 0.0 and ek[3] == 0.0 and ek[4] == 0.0


def test_energy_momenta_3d_unweighted_momenta_and_angular_momentum():
    cfg = _cfg3()
    orc = Oracle(cfg)
    part = np.zeros((2, 7), order="F")
    part[:, 0] = cfg.xmn + 3.0 * cfg.dx
    part[:, 1] = [0.5, -1.0]          # y
    part[:, 2] = [0.25, 2.0]          # z
    part[:, 3] = [0.75, 0.0]          # px
    part[:, 4] = [0.0, 0.0]           # py
    part[:, 5] = [0.0, 0.75]          # pz
    w = np.array([0.5, 0.25], dtype=np.float32)
    part[:, 6] = pack_weight(-1, 0, w)
    ek, ekmax = orc.energy_momenta(part)
    assert ek[0] == 0.75 and ek[1] == 0.0 and ek[2] == 0.75   # unweighted in 3-D
    # sum w (y pz - z py) = 0.25 * (-1.0 * 0.75)
    assert ek[3] == 0.25 * (-0.75)
    assert ek[5] == -0.75
    assert abs(ek[6] - (0.5 + 0.25) * 0.25) < 1e-15
    assert abs(ekmax - 0.25) < 1e-15


def test_field_energy_counts_the_interior_only():
    for cfg, dgvol in ((_cfg2(), None), (_cfg3(), None)):
        orc = Oracle(cfg)
        ebf = np.full(cfg.field_shape, 1.0e3, order="F")  # ghosts: large, must not count
        vals = 0.5 * (1.0 + np.arange(cfg.nfield))
        if cfg.ndim == 2:
            inner = (slice(2, cfg.nx + 2), slice(2, cfg.ny + 2), slice(0, 1))
            dgvol = cfg.dx * cfg.dy * cfg.dy
            ncell = cfg.nx * cfg.ny
        else:
            inner = (slice(2, cfg.nx + 2), slice(2, cfg.ny + 2), slice(2, cfg.nz + 2))
            dgvol = cfg.dx * cfg.dy * cfg.dz
            ncell = cfg.nx * cfg.ny * cfg.nz
        for c in range(cfg.nfield):
            ebf[inner + (c,)] = vals[c]
        out = orc.field_energy(ebf)
        expect = dgvol * ncell * vals ** 2
        assert np.allclose(out, expect, rtol=1e-14, atol=0.0), (out, expect)


def test_removal_keeps_the_half_open_box_and_the_order():
    cfg = _cfg2()
    orc = Oracle(cfg)
    xs = [cfg.xmn, cfg.xmn + 1e-12, cfg.xmx, cfg.xmx + 1e-12, 0.5 * (cfg.xmn + cfg.xmx)]
    ys = [0.0, 0.0, 0.0, 0.0, cfg.ymx + 1e-12]
    n = len(xs)
    part = np.zeros((n, 5), order="F")
    part[:, 0] = xs
    part[:, 1] = ys
    part[:, 2] = np.arange(n)          # tags
    part[:, 4] = pack_weight(-1, 0, np.full(n, 0.5, np.float32))
    ebfp = np.zeros((n, 5), order="F")
    ebfp[:, 0] = 10.0 + np.arange(n)
    m = orc.part_dist(part, ebfp)
    # x = xmn is outside (x > gmin), x = xmx inside (x <= gmax); y above ymx outside
    assert m == 2
    assert list(part[:m, 2]) == [1.0, 2.0]
    assert list(ebfp[:m, 0]) == [11.0, 12.0]    # the stage buffer moves with its particle
