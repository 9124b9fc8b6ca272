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
                    sy.md_run(20)
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
