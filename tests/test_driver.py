"""The LP time loop (ALaDyn.F90:66-107, pic_evolve_in_time.f90:113-140,
run_data_info.f90:68-71, start_all.F90:219-226) on a recording stand-in for
the device, with expectations worked out from those lines."""


from aladyn_amd.driver import dt_loc, lp_cycle, with_dt_loc
from cases import small_2d


def test_dt_loc_divides_tmax():
    assert dt_loc(0.03, 0.3) == 0.3 / 10
    assert dt_loc(0.03, 0.31) == 0.31 / 10      # int(10.33) = 10 steps
    assert dt_loc(0.03, 0.0) == 0.03             # tmax = 0: dt unchanged
    cfg, _, _ = small_2d()
    c2 = with_dt_loc(cfg, 100 * cfg.dt + 0.5 * cfg.dt)
    assert c2.dt == (100.5 * cfg.dt) / 100 and c2.nx == cfg.nx


class _Rec:
    def __init__(self, dt):
        self.cfg = type("C", (), {"dt": dt})()
        self.calls = []

    def lpf2_evolve(self, it, initial_time=False):
        self.calls.append(("step", it, bool(initial_time)))

    def lp_window_xshift(self, witr, it, w_speed):
        self.calls.append(("window", witr, it))
        return 0


def test_cycle_steps_and_window_calls():
    dt = 0.25
    g = _Rec(dt)
    it, tnow = lp_cycle(g, tmax=2.0, w_speed=1.0, wi_time=0.5, wf_time=1.5, w_sh=2)
    assert (it, tnow) == (8, 2.0)
    steps = [c for c in g.calls if c[0] == "step"]
    assert [c[1] for c in steps] == list(range(8))
    assert [c[2] for c in steps] == [True] + [False] * 7
    # window after lpf2_evolve(iter) when 0.5 <= tnow < 1.5 (iter 2..5) and iter even
    assert [c for c in g.calls if c[0] == "window"] == [("window", 2, 2), ("window", 2, 4)]
    assert g.calls.index(("window", 2, 2)) == g.calls.index(("step", 2, False)) + 1


def test_cycle_restart_time_has_no_initial_step():
    g = _Rec(0.25)
    lp_cycle(g, tmax=1.0, tstart=0.5)
    assert [c for c in g.calls] == [("step", 0, False), ("step", 1, False)]
