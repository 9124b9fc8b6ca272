"""Autotuning of per-system launch parameters.

Two searches, both on the device-resident MD path and both timing the same
short trajectory from the same state for every candidate:

* tune_skin: the pair-list skin.  A larger skin rebuilds the cell sort and
  cluster-pair list less often but makes the pair kernel test more candidate
  pairs.
* tune_launch: the launch shapes of one system (pmx.TUNE_* knobs), one knob
  at a time with the others at their defaults: the pair-kernel schedule
  (the list's runs in contiguous shares per wave, or the first run static
  and the rest from a counter), the list prune after each build, the stream
  of the exclusion correction and bonded terms, the charge spread fused into
  the kick-drift launch or apart, the pair-kernel blocks per
  CU (fewer leave room for the PME chain on the second stream), the FFT
  lines per block and whole-plane FFT launches (the PME grid blocking) and
  the stream of the PME spread.

No knob changes a mixed-form result: the list's runs depend on the list
only, every FP32 partial sum is over a tile, and the accumulation is in
fixed point.  The FP64 form's atomics may add in another order.
"""
from __future__ import annotations

import time

from . import pmx

LAUNCH_CANDIDATES = (
    (pmx.TUNE_PAIR_SCHEDULE, (0, 1)),
    (pmx.TUNE_LIST_PRUNE, (1, 0)),
    (pmx.TUNE_FIXUP_STREAM, (1, 0)),
    (pmx.TUNE_FUSED_SPREAD, (1, 0)),
    (pmx.TUNE_PAIR_BLOCKS_PER_CU, (0, -1, -2)),
    (pmx.TUNE_FFT_LINES, (16, 8, 4)),
    (pmx.TUNE_FFT_PLANES, (1, 0)),
    (pmx.TUNE_SPREAD_STREAM, (0, 1)),
)


def _time_run(sy, x, v, steps: int, warmup: int = 20) -> float:
    sy.md_set_state(x, v)
    sy.md_run(warmup)
    t0 = time.perf_counter()
    sy.md_run(steps)
    return (time.perf_counter() - t0) * 1e3 / steps


def tune_skin(ctx, s, spec: pmx.ForceSpec, precision: int, x, v, timestep_fs: float = 1.0,
              candidates=(0.5, 0.7, 0.9, 1.2, 1.5), steps: int = 400, rounds: int = 2) -> tuple[float, dict]:
    """Returns (best skin, {skin: ms per step}); each time is the best of
    `rounds` runs (the rebuild cadence makes a single short run noisy)."""
    times = {}
    for skin in candidates:
        sy = pmx.System(ctx, s, spec, precision)
        try:
            sy.md_init(s, timestep_fs, skin=skin)
            times[skin] = min(_time_run(sy, x, v, steps) for _ in range(rounds))
        finally:
            sy.close()
    best = min(times, key=times.get)
    return best, times


def tune_launch(sy, x, v, steps: int = 400, candidates=LAUNCH_CANDIDATES, rounds: int = 2,
                margin: float = 0.01) -> tuple[dict, dict]:
    """Tunes the launch shapes of a System whose MD state is initialised
    (md_init done).  Leaves the best shapes set on it and returns
    ({knob: value}, {knob: {value: ms per step}}).  The first candidate of
    each knob is its default (below 100,000 atoms for the fixup stream);
    each time is the best of `rounds` runs, and a
    knob leaves its default only for a candidate faster by more than
    `margin` (run-to-run noise of a short timed run is about that)."""
    best = {k: vals[0] for k, vals in candidates}
    for k, val in best.items():
        sy.set_tuning(k, val)
    times: dict = {}
    for _ in range(rounds):
        for knob, vals in candidates:
            t = times.setdefault(knob, {})
            for val in vals:
                sy.set_tuning(knob, val)
                ms = _time_run(sy, x, v, steps)
                t[val] = min(ms, t.get(val, ms))
            sy.set_tuning(knob, best[knob])
    for knob, vals in candidates:
        t = times[knob]
        fastest = min(t, key=t.get)
        best[knob] = fastest if t[fastest] < (1.0 - margin) * t[vals[0]] else vals[0]
        sy.set_tuning(knob, best[knob])
    return best, times
