"""Host logic of the launch-shape autotuner (protomol3_amd/tune.py) on a
scripted stand-in for a System: which candidate each knob keeps, and that the
shapes left set on the system are the chosen ones."""
from protomol3_amd import pmx, tune


class ScriptedSystem:
    """Times of md_run follow a table keyed by the knob values currently set."""

    def __init__(self, cost):
        self.cost = cost  # (knob, value) -> extra ms per step
        self.shape = {}
        self.runs = 0

    def set_tuning(self, knob, value):
        self.shape[knob] = value

    def md_set_state(self, x, v):
        pass

    def md_run(self, steps):
        self.runs += 1

    def ms_per_step(self):
        return 1.0 + sum(self.cost.get(kv, 0.0) for kv in self.shape.items())


def run_tuner(monkeypatch, cost, **kw):
    sy = ScriptedSystem(cost)
    clock = {"t": 0.0}

    def md_run(steps):
        clock["t"] += steps * sy.ms_per_step() * 1e-3

    sy.md_run = md_run
    monkeypatch.setattr(tune.time, "perf_counter", lambda: clock["t"])
    best, times = tune.tune_launch(sy, None, None, **kw)
    return sy, best, times


def test_defaults_kept_without_a_clear_gain(monkeypatch):
    # FFT lines 8 is 0.5 % faster: under the 1 % margin, the default stays
    sy, best, times = run_tuner(monkeypatch, {(pmx.TUNE_FFT_LINES, 8): -0.005})
    for knob, vals in tune.LAUNCH_CANDIDATES:
        assert best[knob] == vals[0]
        assert sy.shape[knob] == vals[0]
        assert set(times[knob]) == set(vals)


def test_a_clear_gain_is_taken_per_knob(monkeypatch):
    cost = {(pmx.TUNE_PAIR_SCHEDULE, 1): -0.2, (pmx.TUNE_SPREAD_STREAM, 1): -0.05,
            (pmx.TUNE_PAIR_BLOCKS_PER_CU, -2): +0.3}
    sy, best, times = run_tuner(monkeypatch, cost)
    assert best[pmx.TUNE_PAIR_SCHEDULE] == 1
    assert best[pmx.TUNE_SPREAD_STREAM] == 1
    assert best[pmx.TUNE_PAIR_BLOCKS_PER_CU] == 0
    assert sy.shape[pmx.TUNE_PAIR_SCHEDULE] == 1 and sy.shape[pmx.TUNE_SPREAD_STREAM] == 1
    # each candidate's time is measured with the other knobs at their defaults
    assert abs(times[pmx.TUNE_PAIR_SCHEDULE][1] - 0.8) < 1e-9
    assert abs(times[pmx.TUNE_PAIR_BLOCKS_PER_CU][-2] - 1.3) < 1e-9
