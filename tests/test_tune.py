"""The autotuner (aladyn_amd.tune): it times every tile shape, threads per
tile and cadence that fits, returns the fastest, records the pick for the
benchmark, and the configuration it picks gives the oracle's step; each
threads-per-tile setting gives the oracle's step."""
import pytest

from aladyn_amd.tune import candidates


def test_candidates():
    from cases import small_2d, small_3d

    c3 = candidates(small_3d()[0])
    c2 = candidates(small_2d()[0])
    c3_64 = candidates(small_3d()[0], 64)
    assert ((8, 8, 4), 64, 0) in c3 and ((8, 8, 4), 64, 512) in c3 and len(c3) == 12
    assert ((16, 16), 64, 0) in c2 and ((16, 16), 64, 512) in c2 and len(c2) == 9
    # two threads per node-cell of the 3-D tile is compiled for FP32 only
    assert ((8, 8, 4), 64, 512) not in c3_64 and len(c3_64) == 9


def test_pick_is_recorded_per_workload(tmp_path):
    from aladyn_amd.tune import load_pick, save_pick

    f = tmp_path / "tuned.json"
    res = {((8, 8, 4), 64, 0): 2.0, ((8, 8, 4), 64, 512): 1.5, ((8, 8, 8), 32, 0): None}
    save_pick("lwfa3d", ((8, 8, 4), 64, 512), res, path=f)
    save_pick("lwfa2d", ((16, 16), 32, 0), {((16, 16), 32, 0): 0.1}, path=f)
    p = load_pick("lwfa3d", path=f)
    assert (p["tile"], p["sort_every"], p["block"], p["ms"]) == ([8, 8, 4], 64, 512, 1.5)
    assert load_pick("lwfa2d", path=f)["tile"] == [16, 16]
    assert load_pick("other", path=f) is None


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 1e-5)])
def test_autotune_picks_a_working_configuration(precision, tol):
    from aladyn_amd import LPStep
    from aladyn_amd.tune import autotune
    from cases import small_3d
    from conftest import rel_maxerr
    from oracle.pyoracle import Oracle

    cfg, part, ebf = small_3d()
    print(f"autotune precision={precision}", flush=True)
    best, res = autotune(cfg, [part], ebf, precision, warmup=3, steps=5, log=print)
    assert best in res and res[best] is not None and res[best] > 0
    assert all(v is None or v >= res[best] for v in res.values())
    assert res[((8, 8, 4), 64, 0)] is not None
    tile, k, b = best
    orc = Oracle(cfg)
    e, jc = ebf.copy(order="F"), orc.zeros_current()
    p = orc.lpf2_evolve(part.copy(order="F"), e, jc, initial_time=True)
    with LPStep(cfg, precision, tile=tile, sort_every=k, block=b, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.lpf2_evolve(0, initial_time=True)
        assert rel_maxerr(gpu.get_fields(), e) <= tol
        assert rel_maxerr(gpu.get_species(0)[:, 3:6], p[:, 3:6]) <= tol


@pytest.mark.gpu
@pytest.mark.parametrize("dim,block", [("3d", 512), ("2d", 512)])
def test_two_threads_per_node_cell_match_oracle(dim, block):
    """The particle kernel with twice the threads per tile: ten FP32 steps
    and a re-layout every step against the oracle (1e-5)."""
    from aladyn_amd import LPStep
    from cases import small_2d, small_3d
    from conftest import rel_maxerr
    from oracle.pyoracle import Oracle

    cfg, part, ebf = (small_3d if dim == "3d" else small_2d)()
    print(f"threads per tile {block}, {dim}", flush=True)
    orc = Oracle(cfg)
    e, jc = ebf.copy(order="F"), orc.zeros_current()
    p = part.copy(order="F")
    with LPStep(cfg, 32, block=block, sort_every=1, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        for it in range(10):
            p = orc.lpf2_evolve(p, e, jc, initial_time=(it == 0))
            gpu.lpf2_evolve(it, initial_time=(it == 0))
        D = cfg.curr_ndim
        pg = gpu.get_species(0)
        assert pg.shape == p.shape
        assert rel_maxerr(pg[:, D:2 * D], p[:, D:2 * D]) <= 1e-5
        assert rel_maxerr(gpu.get_fields(), e) <= 1e-5
        assert rel_maxerr(gpu.get_current(), jc) <= 1e-4
