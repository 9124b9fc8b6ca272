"""Request loop: library exports (CPU); results of requests run together
equal the same requests run alone (GPU)."""
import ctypes as C
import re

import numpy as np
import pytest

from protomol3_amd import pmx, serve, systems


def test_serve_library_exports_header_symbols():
    lib = serve.load()
    names = re.findall(r"^\s*(?:int\s+|const char\s*\*\s*)(pmx_server_\w+)\s*\(",
                       open(serve.LIB_PATH.replace("protomol3_amd/lib/libpmx_serve.so",
                                                   "include/pmx_serve.h")).read(), re.M)
    assert len(names) >= 5
    for nm in names:
        assert hasattr(lib, nm), nm


def test_estimate_covers_the_descriptor_sizes():
    s = systems.replica_box(3000, 11)
    spec = serve.request_spec(s)
    b0 = pmx.estimate_bytes(s, spec, pmx.MIXED)
    b1 = pmx.estimate_bytes(s, spec, pmx.MIXED, trajectory=True)
    assert b1 > b0 > 3000 * 100


def _requests():
    reqs = []
    for i, (n, kind) in enumerate([(1200, 0), (2600, 1), (1800, 0), (3400, 1), (1000, 1), (2200, 0)]):
        s = systems.replica_box(n, 100 + i).with_velocities(300.0, seed=200 + i)
        reqs.append((s, serve.request_spec(s), kind))
    return reqs


@pytest.mark.gpu
def test_gpu_requests_alone_and_in_the_mix_agree():
    reqs = _requests()
    alone = []
    ctx = pmx.Context(0)
    try:
        for s, spec, kind in reqs:
            sy = pmx.System(ctx, s, spec, pmx.MIXED)
            try:
                if kind == 0:
                    f, ec, elj = sy.evaluate(s.pos)
                    alone.append((f, np.array([ec, elj])))
                else:
                    sy.md_init(s, 1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=9)
                    sy.md_set_state(s.pos, s.vel)
                    pmx.run_batch([sy], 20)  # one-shot trajectories run in batched launches
                    x, v, e = sy.md_get_state()
                    alone.append((x, v))
            finally:
                sy.close()
    finally:
        ctx.close()
    srv = serve.Server(0, slots=4)
    try:
        tickets = []
        for s, spec, kind in reqs:
            if kind == 0:
                tickets.append(srv.submit_forces(s, spec, pmx.MIXED))
            else:
                tickets.append(srv.submit_trajectory(s, spec, pmx.MIXED, 20, skin=1.0, gamma=1.0, seed=9))
        for t in tickets:
            srv.wait(t)
    finally:
        srv.close()
    for (s, spec, kind), t, ref in zip(reqs, tickets, alone):
        if kind == 0:
            d = np.abs(t.a - ref[0]).max() / np.abs(ref[0]).max()
            de = np.abs(t.e - ref[1]).max()
        else:
            d = max(np.abs(t.a - ref[0]).max(), np.abs(t.v - ref[1]).max())
            de = 0.0
        print(f"n={s.n} kind={kind} max diff {d:.2e} energy diff {de:.2e} start {t.start_ms:.1f} ms "
              f"done {t.done_ms:.1f} ms", flush=True)
        assert d <= 1e-12 and de <= 1e-9


MD = dict(timestep_fs=1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=5)


@pytest.mark.gpu
def test_gpu_session_keeps_the_system_resident_and_resumes():
    """A session's trajectory requests continue from the resident state: two
    requests of 25 and 35 steps give, to the last bit, what one system run
    for 25 and then 35 steps gives; force requests on a force session match
    one-shot requests at the same positions."""
    s = systems.replica_box(5000, 31).with_velocities(300.0, seed=32)
    spec = serve.request_spec(s)
    ctx = pmx.Context(0)
    try:
        sy = pmx.System(ctx, s, spec, pmx.MIXED)
        try:
            sy.md_init(s, **MD)
            sy.md_set_state(s.pos, s.vel)
            sy.md_run(25)
            x25, v25, _ = sy.md_get_state()
            sy.md_run(35)
            x60, v60, e60 = sy.md_get_state()
        finally:
            sy.close()
    finally:
        ctx.close()
    srv = serve.Server(0, slots=2)
    try:
        traj = srv.open(s, spec, pmx.MIXED, MD)
        frc = srv.open(s, spec, pmx.MIXED)
        t1, t2 = traj.run(25), traj.run(35)
        f1, f2 = frc.forces(x25), frc.forces(x60)
        one = srv.submit_forces(s, spec, pmx.MIXED, x=x60)
        traj.close()
        frc.close()
        for t in (t1, t2, f1, f2, one):
            srv.wait(t)
        peak, _ = srv.stats()
    finally:
        srv.close()
    print(f"session: dx25={np.abs(t1.a - x25).max():.2e} dx60={np.abs(t2.a - x60).max():.2e} "
          f"df={np.abs(f2.a - one.a).max():.2e} peak={peak}", flush=True)
    assert np.array_equal(t1.a, x25) and np.array_equal(t1.v, v25)
    assert np.array_equal(t2.a, x60) and np.array_equal(t2.v, v60)
    assert np.allclose(t2.e, [e60[k] for k in ("coulomb", "lj", "bond", "angle", "kinetic")], rtol=1e-12, atol=0)
    assert np.array_equal(f2.a, one.a) and np.array_equal(f2.e, one.e)
    assert not np.array_equal(f1.a, f2.a)


@pytest.mark.gpu
def test_gpu_admission_keeps_resident_footprints_within_the_budget():
    """Serving-size requests (1,000-100,000 atoms) under a 320 MB budget: the
    footprints resident at any moment (running one-shot requests plus open
    sessions) never exceed it, so some requests wait for others to finish;
    every result equals the same request run alone."""
    sizes = [100000, 90000, 60000, 40000, 20000, 10000, 5000, 3000, 2000, 1000, 70000, 30000]
    reqs = []
    for i, n in enumerate(sizes):
        s = systems.replica_box(n, 700 + i).with_velocities(300.0, seed=800 + i)
        reqs.append((s, serve.request_spec(s), i % 2))
    budget = 320 << 20
    srv = serve.Server(0, slots=4, hbm_budget=budget)
    try:
        ts = []
        for s, spec, kind in reqs:
            if kind == 0:
                ts.append(srv.submit_forces(s, spec, pmx.MIXED))
            else:
                ts.append(srv.submit_trajectory(s, spec, pmx.MIXED, 20, skin=1.0, gamma=1.0, seed=9))
        for t in ts:
            srv.wait(t)
        peak, peak_running = srv.stats()
    finally:
        srv.close()
    foot = [pmx.estimate_bytes(s, spec, pmx.MIXED, trajectory=bool(kind)) for s, spec, kind in reqs]
    print(f"admission: budget {budget} peak {peak} peak_running {peak_running} footprints {foot}", flush=True)
    assert max(foot) <= budget
    assert peak <= budget
    # the resident sum at every start, from the requests' own time lines
    waited = 0
    for i, t in enumerate(ts):
        resident = sum(f for f, u in zip(foot, ts) if u.start_at <= t.start_at < u.done_at)
        assert resident <= budget, (i, resident)
        waited += any(u.done_at <= t.start_at for u in ts if u is not t and u.submit_at <= t.submit_at)
    assert waited > 0  # the budget made requests queue
    ctx = pmx.Context(0)
    try:
        for (s, spec, kind), t in zip(reqs, ts):
            sy = pmx.System(ctx, s, spec, pmx.MIXED)
            try:
                if kind == 0:
                    f, ec, elj = sy.evaluate(s.pos)
                    assert np.array_equal(t.a, f) and t.e[0] == ec and t.e[1] == elj
                else:
                    sy.md_init(s, 1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=9)
                    sy.md_set_state(s.pos, s.vel)
                    pmx.run_batch([sy], 20)
                    x, v, _ = sy.md_get_state()
                    assert np.array_equal(t.a, x) and np.array_equal(t.v, v)
            finally:
                sy.close()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_gpu_queued_trajectories_run_as_one_batch():
    """One slot, busy with a force request while six trajectory requests
    queue behind it: the slot then starts all six together (one batch of
    launches), and each result equals the request run alone in a batch of
    one; a request with another step count waits for its own batch."""
    big = systems.replica_box(60000, 900)
    trajs = []
    for i, n in enumerate([1500, 800, 2400, 1100, 3000, 1900]):
        s = systems.replica_box(n, 910 + i).with_velocities(300.0, seed=920 + i)
        trajs.append((s, serve.request_spec(s)))
    srv = serve.Server(0, slots=1)
    try:
        first = srv.submit_forces(big, serve.request_spec(big), pmx.MIXED)
        ts = [srv.submit_trajectory(s, spec, pmx.MIXED, 30, skin=1.0, gamma=1.0, seed=9) for s, spec in trajs]
        other = srv.submit_trajectory(*trajs[0], pmx.MIXED, 10, skin=1.0, gamma=1.0, seed=9)
        for t in [first, *ts, other]:
            srv.wait(t)
    finally:
        srv.close()
    starts = {t.start_at for t in ts}
    print(f"batch: starts {sorted(starts)} first done {first.done_at} other start {other.start_at}", flush=True)
    assert len(starts) == 1 and min(starts) >= first.done_at
    assert other.start_at >= max(t.done_at for t in ts)
    ctx = pmx.Context(0)
    try:
        for (s, spec), t in zip(trajs, ts):
            sy = pmx.System(ctx, s, spec, pmx.MIXED)
            try:
                sy.md_init(s, 1.0, skin=1.0, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=9)
                sy.md_set_state(s.pos, s.vel)
                pmx.run_batch([sy], 30)
                x, v, _ = sy.md_get_state()
            finally:
                sy.close()
            assert np.array_equal(t.a, x) and np.array_equal(t.v, v)
    finally:
        ctx.close()
