"""Batched launches: pmx_md_run_batch runs every kernel of an MD step once for
many systems of uneven size.

* A system's trajectory does not depend on the batch it runs in: each system
  run in a batch of one gives, in the mixed form, the same positions,
  velocities and energies to the last bit as in the batch of all.
* The trajectory is pmx_md_run's: in the FP64 form to 1e-9 (the two differ only
  in the steps at which the list is rebuilt, which FP64 forces do not see,
  and in the order of double atomics); in the mixed form to 1e-6 A after 60
  steps (its FP32 tile sums follow the list).
* The energies the batch leaves are those of a fresh evaluation at the final
  positions (the device-side list rebuilds kept every list valid).
* pmx_sys_compute_batch gives every system the forces and energies that
  pmx_sys_evaluate gives it.
* Batches that mix precisions or force kinds, repeat a system or hold a
  non-PME system are refused."""
import numpy as np
import pytest

from protomol3_amd import pmx, serve, systems

pytestmark = pytest.mark.gpu

# the largest box's PME planes (80 x 80) exceed the one-launch plane pass,
# the others take it: one batch holds both kinds of FFT launch
SIZES = [600, 1500, 2500, 4000, 900, 3100, 1200, 16000]
STEPS = 60


def _systems():
    out = []
    for i, n in enumerate(SIZES):
        s = systems.replica_box(n, 300 + i).with_velocities(300.0, seed=400 + i)
        out.append((s, serve.request_spec(s)))
    return out


def _make(ctx, s, spec, precision, seed, skin=1.0):
    sy = pmx.System(ctx, s, spec, precision)
    sy.md_init(s, 1.0, skin=skin, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=seed)
    sy.md_set_state(s.pos, s.vel)
    return sy


def _run(ctx, reqs, precision, how, idx=None):
    idx = range(len(reqs)) if idx is None else idx
    sys_ = [_make(ctx, *reqs[i], precision, 50 + i) for i in idx]
    try:
        if how == "batch":
            pmx.run_batch(sys_, STEPS)
        elif how == "alone_batch":
            for sy in sys_:
                pmx.run_batch([sy], STEPS)
        else:
            for sy in sys_:
                sy.md_run(STEPS)
        return [sy.md_get_state() for sy in sys_], [sy.md_stats() for sy in sys_]
    finally:
        for sy in sys_:
            sy.close()


@pytest.fixture(scope="module")
def ctx():
    c = pmx.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_batch_matches_batch_of_one_and_md_run(ctx, precision):
    reqs = _systems()
    batch, bstats = _run(ctx, reqs, precision, "batch")
    ones, _ = _run(ctx, reqs, precision, "alone_batch")
    chunked, cstats = _run(ctx, reqs, precision, "md_run")
    tol_one = 0.0 if precision == pmx.MIXED else 1e-9
    tol_run = 1e-6 if precision == pmx.MIXED else 1e-9
    for i, ((xb, vb, eb), (xo, vo, eo), (xr, vr, er)) in enumerate(zip(batch, ones, chunked)):
        dxo, dxr = np.abs(xb - xo).max(), np.abs(xb - xr).max()
        dvr = np.abs(vb - vr).max()
        print(f"system {i} n={SIZES[i]} prec={precision} vs one {dxo:.3e} vs md_run {dxr:.3e} dv {dvr:.3e} "
              f"stats {bstats[i]} md_run {cstats[i]}\n  energies batch {eb}\n  md_run {er}", flush=True)
        assert bstats[i]["steps"] == STEPS and bstats[i]["rebuilds"] > 0
        assert dxo <= tol_one * max(1.0, np.abs(xb).max())
        assert np.abs(vb - vo).max() <= tol_one * max(1.0, np.abs(vb).max())
        if precision == pmx.MIXED:
            for k in eb:
                assert eb[k] == eo[k], k
        assert dxr <= tol_run * max(1.0, np.abs(xr).max())
        assert dvr <= 10 * tol_run * max(1.0, np.abs(vr).max())
        # terms of opposite sign can leave a small total (the Coulomb energy
        # of these boxes): the bound is relative to the sum of the magnitudes
        scale = sum(abs(e) for e in er.values())
        for k in eb:
            assert abs(eb[k] - er[k]) <= 1e-6 * scale, k


def test_batch_energies_are_those_of_the_final_positions(ctx):
    reqs = _systems()[:4]
    sys_ = [_make(ctx, *reqs[i], pmx.FP64, 70 + i, skin=0.6) for i in range(len(reqs))]
    try:
        pmx.run_batch(sys_, 45)
        states = [sy.md_get_state() for sy in sys_]
    finally:
        for sy in sys_:
            sy.close()
    for (s, spec), (x, v, e) in zip(reqs, states):
        sy = pmx.System(ctx, s, spec, pmx.FP64)
        try:
            _, ec, elj = sy.evaluate(x)
        finally:
            sy.close()
        print(f"n={s.n} batch Ec {e['coulomb']:.10g} fresh {ec:.10g} Elj {e['lj']:.10g} fresh {elj:.10g}", flush=True)
        assert abs(e["coulomb"] - ec) <= 1e-9 * abs(ec)
        assert abs(e["lj"] - elj) <= 1e-9 * max(1.0, abs(elj))


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_force_batch_matches_each_system_alone(ctx, precision):
    """pmx_sys_compute_batch: every system's forces and energies at its own
    positions equal the system's in a batch of its own to the last bit
    (mixed form) and pmx_sys_evaluate's to FP32 rounding (mixed: batched
    lists are cut into coarser runs, so the FP32 tile sums differ) or to the
    order of double atomics (FP64), twice in a row on the same systems at new
    positions (lists rebuilt, accumulators left clear)."""
    reqs = _systems()
    rng = np.random.default_rng(5)
    sys_ = [pmx.System(ctx, s, spec, precision) for s, spec in reqs]
    try:
        for rnd in range(2):
            xs = [s.pos + rng.normal(0.0, 0.02, s.pos.shape) * rnd for s, _ in reqs]
            got = pmx.evaluate_batch(sys_, xs)
            for i, ((s, spec), x, (f, ec, elj)) in enumerate(zip(reqs, xs, got)):
                one = pmx.System(ctx, s, spec, precision)
                try:
                    f1, ec1, elj1 = one.evaluate(x)
                    fb, ecb, eljb = pmx.evaluate_batch([one], [x])[0]
                finally:
                    one.close()
                df = np.abs(f - f1).max() / np.abs(f1).max()
                rms = np.sqrt(np.mean(np.sum((f - f1) ** 2, 1)) / np.mean(np.sum(f1 ** 2, 1)))
                print(f"round {rnd} system {i} n={s.n} prec={precision} df {df:.2e} rms {rms:.2e} "
                      f"dEc {ec - ec1:.2e} dElj {elj - elj1:.2e}", flush=True)
                if precision == pmx.MIXED:
                    assert np.array_equal(f, fb) and ec == ecb and elj == eljb
                    assert rms <= 1e-6
                    # FP32 pair energies summed in other groups: rounding of
                    # terms of up to ~1e3 kcal/mol, whose total can be small
                    assert abs(ec - ec1) <= 1e-6 * abs(ec1) + 1e-3 and abs(elj - elj1) <= 1e-6 * abs(elj1) + 1e-3
                else:
                    assert df <= 1e-12
                    assert abs(ec - ec1) <= 1e-11 * abs(ec1) + 1e-9 and abs(elj - elj1) <= 1e-11 * abs(elj1) + 1e-9
    finally:
        for sy in sys_:
            sy.close()


def test_batch_refuses_mixed_batches(ctx):
    reqs = _systems()[:2]
    a = _make(ctx, *reqs[0], pmx.MIXED, 1)
    b = _make(ctx, *reqs[1], pmx.FP64, 2)
    s, _ = reqs[0]
    c = _make(ctx, s, pmx.ForceSpec(5.0, lj=(pmx.SW_C2, 7.0, 9.0), coul=(pmx.SW_C2, 7.0, 9.0)), pmx.MIXED, 3)
    pos = {id(a): reqs[0][0].pos, id(b): reqs[1][0].pos, id(c): s.pos}
    try:
        for bad in ([a, b], [a, a], [a, c]):
            with pytest.raises(pmx.PmxError):
                pmx.run_batch(bad, 5)
            with pytest.raises(pmx.PmxError):
                pmx.evaluate_batch(bad, [pos[id(sy)] for sy in bad])
        pmx.run_batch([a], 0)
    finally:
        for sy in (a, b, c):
            sy.close()


def test_batch_order6_and_skin0_follow_md_run(ctx):
    """PME stencil order 6 and a zero skin (a list build every step) in
    batched launches: the FP64 trajectories equal pmx_md_run's to 1e-9."""
    reqs = [(s, serve.request_spec(s, order=6)) for s, _ in _systems()[:4]]
    out = {}
    for how in ("batch", "md_run"):
        sys_ = [_make(ctx, *reqs[i], pmx.FP64, 90 + i, skin=0.0) for i in range(len(reqs))]
        try:
            if how == "batch":
                pmx.run_batch(sys_, 25)
            else:
                for sy in sys_:
                    sy.md_run(25)
            out[how] = [(sy.md_get_state(), sy.md_stats()) for sy in sys_]
        finally:
            for sy in sys_:
                sy.close()
    for i, (((xb, vb, _), sb), ((xr, vr, _), sr)) in enumerate(zip(out["batch"], out["md_run"])):
        print(f"order 6 skin 0 system {i}: dx {np.abs(xb - xr).max():.2e} stats {sb} md_run {sr}", flush=True)
        assert sb["rebuilds"] >= 25
        assert np.abs(xb - xr).max() <= 1e-9 * max(1.0, np.abs(xr).max())
        assert np.abs(vb - vr).max() <= 1e-9 * max(1.0, np.abs(vr).max())
