"""Autotuning of per-system launch parameters.

Currently tunes the pair-list skin of the device-resident MD path: a larger
skin rebuilds the cell sort and cluster-pair list less often but makes the
pair kernel test more candidate pairs.  Each candidate runs the same short
trajectory from the same state; the fastest is kept.
"""
from __future__ import annotations

import time

from . import pmx


def tune_skin(ctx, s, spec: pmx.ForceSpec, precision: int, x, v, timestep_fs: float = 1.0,
              candidates=(0.5, 0.7, 0.9, 1.2), steps: int = 200) -> tuple[float, dict]:
    """Returns (best skin, {skin: ms per step})."""
    times = {}
    for skin in candidates:
        sy = pmx.System(ctx, s, spec, precision)
        try:
            sy.md_init(s, timestep_fs, skin=skin)
            sy.md_set_state(x, v)
            sy.md_run(20)
            t0 = time.perf_counter()
            sy.md_run(steps)
            times[skin] = (time.perf_counter() - t0) * 1e3 / steps
        finally:
            sy.close()
    best = min(times, key=times.get)
    return best, times
