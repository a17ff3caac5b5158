"""The autotuner (aladyn_amd.tune): it times every tile shape and cadence
that fits, returns the fastest, and the configuration it picks gives the
oracle's step."""
import pytest

from aladyn_amd.tune import candidates


def test_candidates():
    from cases import small_2d, small_3d

    c3 = candidates(small_3d()[0])
    c2 = candidates(small_2d()[0])
    assert ((8, 8, 4), 64) in c3 and len(c3) == 9
    assert ((16, 16), 64) in c2 and len(c2) == 6


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
    assert res[((8, 8, 4), 64)] is not None
    tile, k = best
    orc = Oracle(cfg)
    e, jc = ebf.copy(order="F"), orc.zeros_current()
    p = orc.lpf2_evolve(part.copy(order="F"), e, jc, initial_time=True)
    with LPStep(cfg, precision, tile=tile, sort_every=k, track_ids=True) as gpu:
        gpu.put_species(0, part)
        gpu.put_fields(ebf)
        gpu.lpf2_evolve(0, initial_time=True)
        assert rel_maxerr(gpu.get_fields(), e) <= tol
        assert rel_maxerr(gpu.get_species(0)[:, 3:6], p[:, 3:6]) <= tol
