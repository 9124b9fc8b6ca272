"""Batched replicas: pmx_md_run_many runs independent systems of uneven size
concurrently (systems spread over several contexts, chunks interleaved per
context); each system's trajectory must be the one pmx_md_run gives it alone.
The mixed form accumulates forces in int64 fixed point, so the trajectories
are reproducible to the last bit; the FP64 form uses double atomics, whose
order may differ."""
import numpy as np
import pytest

from protomol3_amd import pmx, serve, systems

pytestmark = pytest.mark.gpu

SIZES = [600, 1500, 2500, 4000, 900, 3100, 1200]
STEPS = 60


def _systems():
    out = []
    for i, n in enumerate(SIZES):
        s = systems.replica_box(n, 300 + i).with_velocities(300.0, seed=400 + i)
        out.append((s, serve.request_spec(s)))
    return out


def _make(ctx, s, spec, precision, seed):
    sy = pmx.System(ctx, s, spec, precision)
    sy.md_init(s, 1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=seed)
    sy.md_set_state(s.pos, s.vel)
    return sy


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_run_many_matches_each_system_alone(precision):
    reqs = _systems()
    alone = []
    ctx = pmx.Context(0)
    try:
        for i, (s, spec) in enumerate(reqs):
            sy = _make(ctx, s, spec, precision, 50 + i)
            try:
                sy.md_run(STEPS)
                alone.append(sy.md_get_state())
            finally:
                sy.close()
    finally:
        ctx.close()
    ctxs = [pmx.Context(0) for _ in range(3)]
    sys_ = []
    try:
        for i, (s, spec) in enumerate(reqs):
            sys_.append(_make(ctxs[i % 3], s, spec, precision, 50 + i))
        pmx.run_many(sys_, STEPS)
        many = [sy.md_get_state() for sy in sys_]
        stats = [sy.md_stats() for sy in sys_]
    finally:
        for sy in sys_:
            sy.close()
        for c in ctxs:
            c.close()
    tol = 0.0 if precision == pmx.MIXED else 1e-9
    for i, ((xa, va, ea), (xm, vm, em), st) in enumerate(zip(alone, many, stats)):
        dx, dv = np.abs(xa - xm).max(), np.abs(va - vm).max()
        print(f"system {i} n={SIZES[i]} prec={precision} dx={dx:.3e} dv={dv:.3e} stats={st}", flush=True)
        assert st["steps"] == STEPS
        assert dx <= tol * max(1.0, np.abs(xa).max())
        assert dv <= tol * max(1.0, np.abs(va).max())
        for k in ea:
            assert abs(ea[k] - em[k]) <= 1e-9 * max(1.0, abs(ea[k])), k


def test_run_many_rejects_a_repeated_system():
    s, spec = _systems()[0]
    ctx = pmx.Context(0)
    try:
        sy = _make(ctx, s, spec, pmx.MIXED, 1)
        try:
            with pytest.raises(pmx.PmxError):
                pmx.run_many([sy, sy], 5)
        finally:
            sy.close()
    finally:
        ctx.close()
