"""GPU device-resident MD (leapfrog + bonded + nonbonded, all on the device)
against the oracle integrator and the reference's golden trajectory."""
import numpy as np
import pytest

from oracle import md
from oracle import oracle as O
from protomol3_amd import pmx, systems
from test_golden_trajectory import SPEC, check_outputs, load_golden, water423

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pmx.Context(0)
    yield c
    c.close()


def run_gpu_md(ctx, s, spec, v0, dt, nsteps, precision, skin, chunks=1, initial=None):
    """nsteps of leapfrog on the device; initial (a list) receives the total
    forces at the start configuration."""
    sy = pmx.System(ctx, s, pmx.ForceSpec(spec["cellsize"], lj=spec.get("lj"), coul=spec.get("coul"),
                                          pme=spec.get("pme")), precision)
    try:
        sy.md_init(s, dt, skin=skin)
        sy.md_set_state(s.pos, v0)
        if initial is not None:
            sy.md_get_state()
            initial.append(sy.get_forces()[0])
        for _ in range(chunks):
            sy.md_run(nsteps // chunks)
        x, v, e = sy.md_get_state()
        f, _, _ = sy.get_forces()
        stats = sy.md_stats()
    finally:
        sy.close()
    return x, v, f, e, stats


@pytest.mark.parametrize("skin", [0.0, 1.0])
def test_gpu_md_reproduces_reference_golden_trajectory(ctx, skin, tmp_path):
    """water_CHARMM_PERIODIC_L: 200 steps of 0.5 fs on the GPU (FP64 form),
    compared with the reference's expected outputs at its harness epsilon;
    the force trajectory (XYZForceFile) and energies line written from the
    GPU's results match the reference's expected files line by line."""
    s, v0 = water423()
    pos, vel, f0, f200, en = load_golden()
    f_init = []
    x, v, f, e, stats = run_gpu_md(ctx, s, SPEC, v0, 0.5, 200, pmx.FP64, skin, initial=f_init)
    assert np.abs(f_init[0] - f0).max() < 1e-7
    check_outputs(s, f_init[0], f, e, tmp_path)
    print(f"skin={skin} stats={stats} dpos={np.abs(x - pos).max():.2e} dvel={np.abs(v - vel).max():.2e} "
          f"dforce={np.abs(f - f200).max():.2e} energies={e}", flush=True)
    assert stats["steps"] == 200
    assert np.abs(x - pos).max() < 1e-8
    assert np.abs(v - vel).max() < 1e-8
    assert np.abs(f - f200).max() < 1e-7
    for k, ref in en.items():
        assert abs(e[k] - ref) <= 6e-8 * abs(ref) + 1e-5, k


def test_gpu_md_pme_matches_oracle_integrator(ctx):
    s = systems.water_box(seed=21).with_velocities(300.0, seed=5)
    spec = dict(cellsize=5.0, lj=(O.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
    ospec = O.make_spec(5.0, lj=spec["lj"], pme=spec["pme"])
    xr, vr, fr, er = md.leapfrog_run(s, ospec, s.pos, s.vel, 1.0, 10)
    x, v, f, e, stats = run_gpu_md(ctx, s, spec, s.vel, 1.0, 10, pmx.FP64, 1.0)
    print(f"fp64 dx={np.abs(x - xr).max():.2e} dv={np.abs(v - vr).max():.2e} stats={stats}", flush=True)
    assert np.abs(x - xr).max() < 1e-9
    assert np.abs(v - vr).max() < 1e-8
    for k in er:
        assert abs(e[k] - er[k]) <= 1e-9 * max(1.0, abs(er[k])), k
    # mixed form: same trajectory within FP32 pair-arithmetic error
    xm, vm, fm, em, _ = run_gpu_md(ctx, s, spec, s.vel, 1.0, 10, pmx.MIXED, 1.0)
    print(f"mixed dx={np.abs(xm - xr).max():.2e} dv={np.abs(vm - vr).max():.2e}", flush=True)
    assert np.abs(xm - xr).max() < 1e-5
    assert np.abs(vm - vr).max() < 1e-4


def test_gpu_md_skin_reuse_is_exact_and_rolls_back(ctx):
    """A tiny skin forces rollbacks; the trajectory must equal the one that
    rebuilds the list every step (FP64 form: up to atomics ordering)."""
    s = systems.water_box(seed=22).with_velocities(600.0, seed=6)
    spec = dict(cellsize=5.0, lj=(O.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
    x0, v0, _, e0, s0 = run_gpu_md(ctx, s, spec, s.vel, 1.0, 60, pmx.FP64, 0.0)
    x1, v1, _, e1, s1 = run_gpu_md(ctx, s, spec, s.vel, 1.0, 60, pmx.FP64, 0.05, chunks=2)
    x2, v2, _, e2, s2 = run_gpu_md(ctx, s, spec, s.vel, 1.0, 60, pmx.FP64, 1.0)
    print(f"skin0 {s0} skin0.05 {s1} skin1 {s2} dx={np.abs(x1 - x0).max():.2e} {np.abs(x2 - x0).max():.2e}",
          flush=True)
    assert s1["rollbacks"] > 0
    assert s2["rebuilds"] < s0["rebuilds"] // 2
    for x, v in ((x1, v1), (x2, v2)):
        assert np.abs(x - x0).max() < 1e-10
        assert np.abs(v - v0).max() < 1e-9


def test_gpu_md_state_energies_before_any_step(ctx):
    """Energies read right after md_set_state are those of the given
    positions (forces computed on demand), not of an earlier evaluation."""
    s = systems.water_box(seed=23).with_velocities(300.0, seed=7)
    spec = dict(cellsize=5.0, lj=(O.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
    ospec = O.make_spec(5.0, lj=spec["lj"], pme=spec["pme"])
    _, eref = md.total_forces(s, ospec, s.pos)
    sy = pmx.System(ctx, s, pmx.ForceSpec(5.0, lj=spec["lj"], pme=spec["pme"]), pmx.FP64)
    try:
        sy.md_init(s, 1.0, skin=0.5)
        sy.md_set_state(s.pos, s.vel)
        _, _, e = sy.md_get_state()
        x2 = s.pos + 0.01
        sy.md_set_state(x2, s.vel)
        _, _, e2 = sy.md_get_state()
    finally:
        sy.close()
    _, eref2 = md.total_forces(s, ospec, x2)
    print(f"energies {e} oracle {eref}", flush=True)
    for k in eref:
        assert abs(e[k] - eref[k]) <= 1e-9 * max(1.0, abs(eref[k])), k
        assert abs(e2[k] - eref2[k]) <= 1e-9 * max(1.0, abs(eref2[k])), k
    assert e["kinetic"] > 0


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_gpu_md_graph_replay_matches_kernel_launches(ctx, precision):
    """Chunks replayed from captured HIP graphs give the trajectory of
    kernel-by-kernel launches: several runs with host work between them, each
    of several chunks (the same graph replayed with other step indices, which
    the Langevin noise reads from device memory), and a small skin that
    forces rolled-back chunks (replayed without graphs)."""
    s = systems.water_box(seed=31).with_velocities(300.0, seed=8)
    spec = pmx.ForceSpec(5.0, lj=(pmx.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
    out = {}
    for graphs in (False, True):
        sy = pmx.System(ctx, s, spec, precision)
        try:
            sy.md_init(s, 1.0, skin=0.3, integrator=pmx.LANGEVIN, temperature=300.0, gamma=5.0, seed=77,
                       graphs=graphs)
            sy.md_set_state(s.pos, s.vel)
            states = []
            for _ in range(4):
                sy.md_run(45)
                states.append(sy.md_get_state())
            out[graphs] = (states, sy.md_stats())
        finally:
            sy.close()
    (eager, st0), (graph, st1) = out[False], out[True]
    print(f"prec={precision} stats eager {st0} graphs {st1}", flush=True)
    if precision == pmx.MIXED:  # FP64 atomics may round differently and move a rollback
        assert st0 == st1
    assert st1["rollbacks"] > 0 and st1["steps"] == 180
    tol = 0.0 if precision == pmx.MIXED else 1e-10
    for (xa, va, ea), (xb, vb, eb) in zip(eager, graph):
        assert np.abs(xa - xb).max() <= tol * max(1.0, np.abs(xa).max())
        assert np.abs(va - vb).max() <= tol * max(1.0, np.abs(va).max())
        for k in ea:
            assert abs(ea[k] - eb[k]) <= 1e-9 * max(1.0, abs(ea[k])), k


@pytest.mark.parametrize("precision", [pmx.MIXED, pmx.FP64])
def test_gpu_md_launch_shapes_keep_the_trajectory(ctx, precision):
    """The autotuner's launch shapes change how the work is cut, not what is
    computed.  A run whose shapes change between its parts follows the
    unchanged run (the mixed form to the last bit: the list's runs depend on
    the list only and every FP32 partial sum is over a tile before it goes to
    fixed point), so the mixed form's forces and energies at one
    configuration are bit-identical under every schedule, block count, list
    prune, fixup stream and FFT blocking (lines or whole planes)."""
    s = systems.water_box(seed=32).with_velocities(300.0, seed=9)
    spec = pmx.ForceSpec(5.0, lj=(pmx.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 30, 24), order=4, alpha=0.32))
    shapes = [{}, {pmx.TUNE_FFT_LINES: 1, pmx.TUNE_SPREAD_STREAM: 1},
              {pmx.TUNE_PAIR_BLOCKS_PER_CU: 1, pmx.TUNE_FFT_LINES: 64, pmx.TUNE_FUSED_SPREAD: 0,
               pmx.TUNE_FFT_PLANES: 0},
              {pmx.TUNE_PAIR_BLOCKS_PER_CU: 3, pmx.TUNE_FFT_LINES: 5, pmx.TUNE_SPREAD_STREAM: 0,
               pmx.TUNE_FUSED_SPREAD: 1, pmx.TUNE_FFT_PLANES: 1}]
    out = {}
    for tuned in (False, True):
        sy = pmx.System(ctx, s, spec, precision)
        try:
            sy.set_tuning(pmx.TUNE_PAIR_SCHEDULE, 1)
            sy.md_init(s, 1.0, skin=0.5, graphs=tuned)
            sy.md_set_state(s.pos, s.vel)
            states = []
            for sh in shapes:
                for k, v in (sh.items() if tuned else ()):
                    sy.set_tuning(k, v)
                sy.md_run(30)
                states.append(sy.md_get_state())
            out[tuned] = (states, sy.md_stats())
        finally:
            sy.close()
    (s0, st0), (s1, st1) = out[False], out[True]
    print(f"prec={precision} stats default {st0} tuned {st1}", flush=True)
    tol = 0.0 if precision == pmx.MIXED else 1e-10
    for (xa, va, ea), (xb, vb, eb) in zip(s0, s1):
        assert np.abs(xa - xb).max() <= tol * max(1.0, np.abs(xa).max())
        assert np.abs(va - vb).max() <= tol * max(1.0, np.abs(va).max())
        for k in ea:
            assert abs(ea[k] - eb[k]) <= 1e-9 * max(1.0, abs(ea[k])), k

    x = s0[-1][0]
    sy = pmx.System(ctx, s, spec, precision)
    try:
        f0, ec0, elj0 = sy.evaluate(x)
        ftol = 0.0 if precision == pmx.MIXED else 1e-12
        cases = [((pmx.TUNE_PAIR_BLOCKS_PER_CU, 1),), ((pmx.TUNE_PAIR_SCHEDULE, 1),),
                 ((pmx.TUNE_PAIR_BLOCKS_PER_CU, -1),), ((pmx.TUNE_LIST_PRUNE, 0),), ((pmx.TUNE_FIXUP_STREAM, 1),),
                 ((pmx.TUNE_FFT_PLANES, 0),), ((pmx.TUNE_FFT_PLANES, 1), (pmx.TUNE_FFT_LINES, 3)),
                 ((pmx.TUNE_PAIR_BLOCKS_PER_CU, 0), (pmx.TUNE_PAIR_SCHEDULE, 0), (pmx.TUNE_LIST_PRUNE, 1))]
        for case in cases:
            for k, v in case:
                sy.set_tuning(k, v)
            f, ec, elj = sy.evaluate(x)
            rel = np.sqrt(np.mean(np.sum((f - f0) ** 2, 1)) / np.mean(np.sum(f0 ** 2, 1)))
            print(f"prec={precision} {case} rms_rel {rel:.3e}", flush=True)
            assert rel <= ftol
            if precision == pmx.MIXED:
                assert np.array_equal(f, f0) and ec == ec0 and elj == elj0
            assert abs(ec - ec0) <= 1e-9 * abs(ec0) and abs(elj - elj0) <= 1e-9 * max(1.0, abs(elj0))
        for knob, bad in ((pmx.TUNE_FFT_LINES, 0), (pmx.TUNE_PAIR_BLOCKS_PER_CU, -8), (pmx.TUNE_PAIR_SCHEDULE, 2),
                          (pmx.TUNE_SPREAD_STREAM, 2), (pmx.TUNE_LIST_PRUNE, 2), (pmx.TUNE_FIXUP_STREAM, 2),
                          (pmx.TUNE_FUSED_SPREAD, 2), (pmx.TUNE_FFT_PLANES, 2), (1, 16), (5, 1), (99, 1)):
            with pytest.raises(pmx.PmxError):
                sy.set_tuning(knob, bad)
    finally:
        sy.close()
