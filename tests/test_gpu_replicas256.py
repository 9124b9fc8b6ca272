"""The batched-replica workload at its stated size: 256 independent systems
with atom counts drawn log-uniform in [500, 20,000] (seed 11, water + 5-10%
bead-polymer solute at 0.0334 atoms/A^3, LJ C2 7-9 A, PME ~1 A grid, mixed
form, Langevin 300 K, 1/ps).

* Every system's step-0 nonbonded forces and energies against the oracle
  (RMS relative force error <= 1e-5, energies 3e-5 relative).
* Every system's 5-step Langevin trajectory run batched (pmx_md_run_many over
  four contexts, all 256 resident at once) against the same system run alone:
  bit-identical, since the mixed form accumulates in int64 fixed point and the
  Langevin noise is a counter-based stream keyed by (seed, atom, step).
* All 256 in one pmx_md_run_batch call against each in a batch of one:
  bit-identical."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from protomol3_amd import pmx, serve, systems

pytestmark = pytest.mark.gpu

NSYS, STEPS, LANES = 256, 5, 4


@pytest.fixture(scope="module")
def draw():
    rng = np.random.default_rng(11)
    sizes = np.exp(rng.uniform(np.log(500), np.log(20000), NSYS)).astype(int)
    boxes = [systems.replica_box(int(n), 2000 + i).with_velocities(300.0, seed=3000 + i)
             for i, n in enumerate(sizes)]
    print(f"replicas: {NSYS} systems, {sum(b.n for b in boxes)} atoms", flush=True)
    return boxes


def _oracle_spec(spec):
    return O.make_spec(spec.cellsize, lj=(O.SW_C2, spec.lj[1], spec.lj[2]), pme=spec.pme)


def test_every_replica_forces_match_the_oracle(draw):
    threads = min(os.cpu_count() or 1, 16)
    ctx = pmx.Context(0)
    worst = 0.0
    try:
        for i, b in enumerate(draw):
            spec = serve.request_spec(b)
            sy = pmx.System(ctx, b, spec, pmx.MIXED)
            try:
                f, ec, elj = sy.evaluate(b.pos)
            finally:
                sy.close()
            fr, ecr, eljr = O.evaluate(b, _oracle_spec(spec), nthreads=threads)
            rms = np.sqrt(np.mean(np.sum((f - fr) ** 2, 1)) / np.mean(np.sum(fr ** 2, 1)))
            worst = max(worst, rms)
            if i % 32 == 0 or rms > 1e-5:
                print(f"replica {i} n={b.n} box={b.box[0]:.2f} rms={rms:.2e} Ecoul {ec:.6f} vs {ecr:.6f} "
                      f"Elj {elj:.6f} vs {eljr:.6f}", flush=True)
            assert rms <= 1e-5, i
            assert abs(ec - ecr) <= 3e-5 * abs(ecr), i
            assert abs(elj - eljr) <= 3e-5 * max(abs(eljr), 1.0), i
    finally:
        ctx.close()
    print(f"replicas: worst RMS relative force error {worst:.2e}", flush=True)


def _make(ctx, b, i):
    sy = pmx.System(ctx, b, serve.request_spec(b), pmx.MIXED)
    sy.md_init(b, 1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=i)
    sy.md_set_state(b.pos, b.vel)
    return sy


def test_every_replica_batched_equals_alone(draw):
    alone = []
    ctx = pmx.Context(0)
    try:
        for i, b in enumerate(draw):
            sy = _make(ctx, b, i)
            try:
                sy.md_run(STEPS)
                alone.append(sy.md_get_state())
            finally:
                sy.close()
    finally:
        ctx.close()
    ctxs = [pmx.Context(0) for _ in range(LANES)]
    sys_ = []
    try:
        for i, b in enumerate(draw):
            sys_.append(_make(ctxs[i % LANES], b, i))
        pmx.run_many(sys_, STEPS)
        many = [sy.md_get_state() for sy in sys_]
    finally:
        for sy in sys_:
            sy.close()
        for c in ctxs:
            c.close()
    for i, ((xa, va, ea), (xm, vm, em)) in enumerate(zip(alone, many)):
        assert np.array_equal(xa, xm) and np.array_equal(va, vm), i
        for k in ea:
            assert ea[k] == em[k], (i, k)
    print(f"replicas: {NSYS} batched trajectories bit-identical to the runs alone", flush=True)


def test_every_replica_in_batched_launches_equals_its_batch_of_one(draw):
    """All 256 systems in one pmx_md_run_batch call (each kernel of a step
    launched once for all of them) against each system in a batch of its
    own: bit-identical, since a block reads only its own system's table entry
    and each system's list rebuilds are decided from its own displacements."""
    ctx = pmx.Context(0)
    try:
        ones = []
        for i, b in enumerate(draw):
            sy = _make(ctx, b, i)
            try:
                pmx.run_batch([sy], STEPS)
                ones.append(sy.md_get_state())
            finally:
                sy.close()
        sys_ = [_make(ctx, b, i) for i, b in enumerate(draw)]
        try:
            pmx.run_batch(sys_, STEPS)
            batch = [sy.md_get_state() for sy in sys_]
        finally:
            for sy in sys_:
                sy.close()
    finally:
        ctx.close()
    for i, ((xa, va, ea), (xb, vb, eb)) in enumerate(zip(ones, batch)):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), i
        for k in ea:
            assert ea[k] == eb[k], (i, k)
    print(f"replicas: {NSYS} trajectories in batched launches bit-identical to batches of one", flush=True)
