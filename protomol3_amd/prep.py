"""Preparation of generated starting structures.

The seeded lattice generators place randomly oriented molecules close
together, which leaves a few intermolecular contacts far inside the LJ
repulsive wall.  relax() removes them with capped steepest descent on the
full force field (nonbonded + bonds + angles) evaluated by the engine, so
that dynamics started from the result is well behaved (a stated
minimisation protocol before NVE runs)."""
from __future__ import annotations

import numpy as np

from . import pmx


def relax(ctx, s, spec: pmx.ForceSpec, iters: int = 300, gamma: float = 2e-4, cap: float = 0.05,
          precision: int = pmx.FP64) -> np.ndarray:
    """Positions after `iters` steps of x += clip(gamma * F, cap) per atom
    (gamma in A^2 mol/kcal, cap in A).  Returns the new positions."""
    sy = pmx.System(ctx, s, spec, precision)
    try:
        sy.md_init(s, 1.0, skin=0.0)  # enables the bonded terms in evaluate()
        x = np.array(s.pos, dtype=float)
        for _ in range(iters):
            f, _, _ = sy.evaluate(x)
            step = gamma * f
            norm = np.linalg.norm(step, axis=1, keepdims=True)
            step *= np.minimum(1.0, cap / np.maximum(norm, 1e-300))
            x += step
    finally:
        sy.close()
    return x
