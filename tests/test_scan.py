"""The batched 2-D scan: 64 independent 2-D LP runs whose sizes, densities,
pulse amplitudes and particles per cell are drawn from one seeded generator,
300 steps each.

* The parameter table is pinned by a committed fixture
  (tests/golden/scan_2d_params.json, written by make_scan_params.py).
* The 64 runs advance eight at a time on one GPU, one context per run,
  their steps interleaved (aladyn_amd.driver.run_interleaved), in FP64, in
  FP32 and in FP32 replayed as captured HIP graphs.  Each run's FP32 moments
  after 300 steps are compared with the same run in FP64: kinetic energy,
  charge and field energy per component within 1e-3 relative, momentum
  sums within 1e-3 of sum w|p|.
* The two smallest runs go through the FP64 oracle for the same 300 steps:
  FP64 moments within 1e-9, FP32 within 1e-3, as the 200-step runs.
* All 64 runs in one batch (aladyn_batch_lpf2_evolve: each kernel of a step
  launched once for all runs, the particle kernel over the work list of all
  their tiles, the step replayed as one HIP graph), FP64 and FP32: the same
  moment checks, and the throughput beside the interleaved runs'.  Runs
  stepped in a batch match the same runs stepped alone (FP64 <= 1e-12 after
  20 steps, re-layouts included), and the two oracle runs pass through a
  batch of their own.
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
    group = 8  # runs advanced together
    # three passes over each group: FP64, FP32, FP32 replayed as graphs.  HIP
    # maps a process's streams onto its hardware queues (4) in turn, so the
    # contexts of one pass are made one after another: the eight runs of a
    # pass then spread over all the queues.
    passes = (("f64", PREC_F64, False), ("f32", PREC_F32, False), ("f32 graphs", PREC_F32, True))
    npart, worst, worst_p = 0, 0.0, 0.0
    wall = {name: 0.0 for name, _, _ in passes}
    for g0 in range(0, len(runs), group):
        cases = [scan_2d_case(r) for r in runs[g0:g0 + group]]
        npart += sum(c[1].shape[0] for c in cases)
        print(f"scan runs {g0}..{g0 + group - 1}", flush=True)
        m = {}
        for name, prec, graphs in passes:
            ctx = []
            try:
                for cfg, part, ebf in cases:
                    gpu = LPStep(cfg, prec)
                    ctx.append(gpu)
                    gpu.put_species(0, part)
                    gpu.put_fields(ebf)
                for g in ctx:
                    g.synchronize()
                t = time.perf_counter()
                run_interleaved(ctx, NSTEPS, graphs=graphs)
                for g in ctx:
                    g.synchronize()
                wall[name] += time.perf_counter() - t
                m[name] = [_moments(g) for g in ctx]
            finally:
                for g in ctx:
                    g.close()
        for r, m64, m32, m32g in zip(runs[g0:g0 + group], m["f64"], m["f32"], m["f32 graphs"]):
            try:
                _check_moments(m32, m64, 1e-3, max(2, m64[0] // 1000))
                _check_moments(m32g, m64, 1e-3, max(2, m64[0] // 1000))
            except AssertionError as e:
                raise AssertionError(f"run {r}: {e}") from None
            for tag, mm in (("f32", m32), ("f32 graphs", m32g)):
                dk = _rel(mm[1][6], m64[1][6])
                if dk > 1e-5:
                    print(f"scan run {r['run']} {tag}: kinetic energy differs by {dk:.2e}",
                          flush=True)
                worst = max(worst, dk)
                worst_p = max(worst_p, max(abs(mm[1][i] - m64[1][i]) for i in range(3)) / m64[3])
    for name, _, _ in passes:
        print(f"scan {name}: {npart} particles x {NSTEPS} steps in {wall[name]:.2f} s, "
              f"{npart * NSTEPS / wall[name]:.3e} particle-steps/s", flush=True)
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


def _make(cases, prec, **kw):
    from aladyn_amd import LPStep

    ctx = []
    for cfg, part, ebf in cases:
        gpu = LPStep(cfg, prec, **kw)
        ctx.append(gpu)
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("graphs,groups", [(False, 1), (True, 1), (True, 2)])
def test_batched_runs_match_runs_alone(graphs, groups):
    from aladyn_amd import PREC_F64
    from aladyn_amd.driver import run_batched
    from conftest import rel_maxerr

    runs = [scan_2d_params()[i] for i in (13, 24, 3, 40)]
    cases = [scan_2d_case(r) for r in runs]
    print(f"batch of {len(cases)} runs, graphs={graphs}, {groups} batches", flush=True)
    alone = _make(cases, PREC_F64, sort_every=4, track_ids=True)
    batch = _make(cases, PREC_F64, sort_every=4, track_ids=True)
    try:
        for g in alone:
            for it in range(20):
                g.lpf2_evolve(it, initial_time=(it == 0))
        run_batched(batch, 20, graphs=graphs, groups=groups)
        for r, a, b in zip(runs, alone, batch):
            pa, pb = a.get_species(0), b.get_species(0)
            assert pa.shape == pb.shape, (r["run"], pa.shape, pb.shape)
            assert rel_maxerr(pb[:, 2:4], pa[:, 2:4]) <= 1e-12, r["run"]
            assert rel_maxerr(b.get_fields(), a.get_fields()) <= 1e-12, r["run"]
            assert rel_maxerr(b.get_current(), a.get_current()) <= 1e-12, r["run"]
    finally:
        for g in alone + batch:
            g.close()


@pytest.mark.gpu
def test_scan_fp32_against_fp64_batched():
    from aladyn_amd import PREC_F32, PREC_F64
    from aladyn_amd.driver import run_batched

    runs = scan_2d_params()
    cases = [scan_2d_case(r) for r in runs]
    npart = sum(c[1].shape[0] for c in cases)
    m, wall = {}, {}
    # FP64 as one batch, FP32 as two batches of 32 stepped in turn
    for name, prec, groups in (("f64", PREC_F64, 1), ("f32", PREC_F32, 2)):
        print(f"scan batch {name}: {len(cases)} runs, {npart} particles, {groups} batches",
              flush=True)
        ctx = _make(cases, prec)
        try:
            for g in ctx:
                g.synchronize()
            t = time.perf_counter()
            run_batched(ctx, NSTEPS, graphs=True, groups=groups)
            wall[name] = time.perf_counter() - t
            m[name] = [_moments(g) for g in ctx]
        finally:
            for g in ctx:
                g.close()
        print(f"scan batch {name}: {npart} particles x {NSTEPS} steps in {wall[name]:.2f} s, "
              f"{npart * NSTEPS / wall[name]:.3e} particle-steps/s", flush=True)
    worst = 0.0
    for r, m64, m32 in zip(runs, m["f64"], m["f32"]):
        try:
            _check_moments(m32, m64, 1e-3, max(2, m64[0] // 1000))
        except AssertionError as e:
            raise AssertionError(f"run {r}: {e}") from None
        worst = max(worst, _rel(m32[1][6], m64[1][6]))
    print(f"scan batch: worst FP32/FP64 kinetic energy difference {worst:.2e}", flush=True)


@pytest.mark.gpu
def test_scan_small_runs_against_oracle_batched():
    from aladyn_amd import PREC_F32, PREC_F64
    from aladyn_amd.driver import run_batched
    from oracle.pyoracle import Oracle

    cases = [scan_2d_case(scan_2d_params()[i]) for i in (13, 24)]
    refs = []
    for cfg, part, ebf in cases:
        orc = Oracle(cfg)
        p, e, jc = part.copy(order="F"), ebf.copy(order="F"), orc.zeros_current()
        for it in range(NSTEPS):
            p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
        ek, _ = orc.energy_momenta(p)
        refs.append((p.shape[0], ek, orc.field_energy(e), _abs_momentum(p)))
    for prec, tol in ((PREC_F64, 1e-9), (PREC_F32, 1e-3)):
        print(f"scan runs 13, 24 in one batch, precision={prec}", flush=True)
        ctx = _make(cases, prec)
        try:
            run_batched(ctx, NSTEPS)
            for g, m_ref in zip(ctx, refs):
                _check_moments(_moments(g), m_ref, tol,
                               0 if prec == PREC_F64 else max(2, m_ref[0] // 1000))
        finally:
            for g in ctx:
                g.close()
