"""The batched 2-D scan: 64 independent 2-D LP runs whose sizes, densities,
pulse amplitudes and particles per cell are drawn from one seeded generator,
300 steps each.

* The parameter table is pinned by a committed fixture
  (tests/golden/scan_2d_params.json, written by make_scan_params.py).
* The 64 runs advance eight at a time on one GPU, one context per run,
  their steps interleaved (aladyn_amd.driver.run_interleaved).  Each run's FP32 moments
  after 300 steps are compared with the same run in FP64: kinetic energy,
  charge and field energy per component within 1e-3 relative, momentum
  sums within 1e-3 of sum w|p|.
* The two smallest runs go through the FP64 oracle for the same 300 steps:
  FP64 moments within 1e-9, FP32 within 1e-3, as the 200-step runs.
"""
import json
import time
from pathlib import Path

import numpy as np
import pytest

from aladyn_amd.inputs import scan_2d_case, scan_2d_params, unpack_weight

GOLDEN = Path(__file__).resolve().parent / "golden" / "scan_2d_params.json"
NSTEPS = 300


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_scan_params_match_fixture():
    runs = scan_2d_params()
    assert runs == json.loads(GOLDEN.read_text())["runs"]
    assert len(runs) == 64 and [r["run"] for r in runs] == list(range(64))
    for r in runs:
        assert 128 <= r["nx"] <= 1024 and r["nx"] % 32 == 0
        assert 64 <= r["ny"] <= 512 and r["ny"] % 32 == 0
        assert 1e-3 <= r["n_over_nc"] <= 2e-2
        assert 1.0 <= r["a0"] <= 4.0
        assert all(1 <= p <= 4 for p in r["ppc"])
        assert r["nx"] * r["ny"] * r["ppc"][0] * r["ppc"][1] <= 8_388_608


def test_scan_case_layout():
    """Pulse inside the box, plasma ahead of it, weights n0_ref*n/nc/ppc."""
    r = scan_2d_params()[13]
    cfg, part, ebf = scan_2d_case(r)
    tau = min(6.0, r["nx"] * cfg.dx / 8.0)
    assert part[:, 0].min() >= cfg.xmn + 3.0 * tau - cfg.dx
    assert part.shape[0] == (r["nx"] - int(np.ceil(3.0 * tau / cfg.dx))) * r["ny"] * \
        r["ppc"][0] * r["ppc"][1]
    ey = ebf[:, :, 0, 1]
    assert np.argmax(np.abs(ey).max(axis=1)) in range(int(tau / cfg.dx), int(2 * tau / cfg.dx) + 3)


def _abs_momentum(part):
    """sum_p w |p|: the scale of the 2-D momentum sums, which weight p
    (diag_part_and_fields.f90:193-211)."""
    _, _, w = unpack_weight(part[:, 4])
    return float(np.sum(w.astype(np.float64) * np.hypot(part[:, 2], part[:, 3])))


def _moments(gpu):
    ek, _ = gpu.energy_momenta(0)
    return gpu.species_count(0), ek, gpu.field_energy(), _abs_momentum(gpu.get_species(0))


def _check_moments(m, m_ref, tol, n_slack):
    """Counts, kinetic energy, charge and field energy per component, and the
    signed momentum sums relative to sum w|p|: in the a0 ~ 4 runs single
    orbits part ways under FP32 round-off, and a signed sum that cancels
    to a small value carries that scatter in absolute terms."""
    n1, ek1, fe1, _ = m
    n0, ek0, fe0, pabs = m_ref
    assert abs(n1 - n0) <= n_slack, ("count", n1, n0)
    assert ek0[6] > 0
    assert _rel(ek1[6], ek0[6]) <= tol, ("kinetic", ek1[6], ek0[6])
    assert _rel(ek1[5], ek0[5]) <= tol, ("charge", ek1[5], ek0[5])
    assert max(abs(ek1[i] - ek0[i]) for i in range(3)) / pabs <= tol, ("momentum", ek1, ek0, pabs)
    assert np.max(np.abs(fe1 - fe0)) / np.sum(fe0) <= tol, ("field", fe1, fe0)


@pytest.mark.gpu
def test_scan_fp32_against_fp64_interleaved():
    from aladyn_amd import PREC_F32, PREC_F64, LPStep
    from aladyn_amd.driver import run_interleaved

    runs = scan_2d_params()
    group = 8  # runs advanced together; HIP maps the streams onto 4 hardware queues
    npart, wall, worst, worst_p = 0, {PREC_F64: 0.0, PREC_F32: 0.0}, 0.0, 0.0
    for g0 in range(0, len(runs), group):
        ctx = {PREC_F64: [], PREC_F32: []}
        try:
            for r in runs[g0:g0 + group]:
                cfg, part, ebf = scan_2d_case(r)
                npart += part.shape[0]
                for prec in (PREC_F64, PREC_F32):
                    gpu = LPStep(cfg, prec)
                    ctx[prec].append(gpu)
                    gpu.put_species(0, part)
                    gpu.put_fields(ebf)
            print(f"scan runs {g0}..{g0 + group - 1}", flush=True)
            m = {}
            for prec in (PREC_F64, PREC_F32):
                t = time.perf_counter()
                run_interleaved(ctx[prec], NSTEPS)
                for g in ctx[prec]:
                    g.synchronize()
                wall[prec] += time.perf_counter() - t
                m[prec] = [_moments(g) for g in ctx[prec]]
        finally:
            for prec in (PREC_F64, PREC_F32):
                for g in ctx[prec]:
                    g.close()
        for r, m32, m64 in zip(runs[g0:g0 + group], m[PREC_F32], m[PREC_F64]):
            try:
                _check_moments(m32, m64, 1e-3, max(2, m64[0] // 1000))
            except AssertionError as e:
                raise AssertionError(f"run {r}: {e}") from None
            worst = max(worst, _rel(m32[1][6], m64[1][6]))
            worst_p = max(worst_p, max(abs(m32[1][i] - m64[1][i]) for i in range(3)) / m64[3])
    for prec in (PREC_F64, PREC_F32):
        print(f"scan precision={prec}: {npart} particles x {NSTEPS} steps in {wall[prec]:.2f} s, "
              f"{npart * NSTEPS / wall[prec]:.3e} particle-steps/s", flush=True)
    print(f"scan: worst FP32/FP64 difference: kinetic energy {worst:.2e}, momentum sums "
          f"{worst_p:.2e} of sum w|p|", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("run", [13, 24])
def test_scan_small_runs_against_oracle(run):
    from aladyn_amd import PREC_F32, PREC_F64, LPStep
    from oracle.pyoracle import Oracle

    cfg, part, ebf = scan_2d_case(scan_2d_params()[run])
    orc = Oracle(cfg)
    p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
    for it in range(NSTEPS):
        p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
    ek, _ = orc.energy_momenta(p)
    m_ref = (p.shape[0], ek, orc.field_energy(e), _abs_momentum(p))
    for prec, tol in ((PREC_F64, 1e-9), (PREC_F32, 1e-3)):
        print(f"scan run {run} np={part.shape[0]} precision={prec}", flush=True)
        with LPStep(cfg, prec) as gpu:
            gpu.put_species(0, part)
            gpu.put_fields(ebf)
            for it in range(NSTEPS):
                gpu.lpf2_evolve(it, initial_time=(it == 0))
            m = _moments(gpu)
        _check_moments(m, m_ref, tol, 0 if prec == PREC_F64 else max(2, m_ref[0] // 1000))
