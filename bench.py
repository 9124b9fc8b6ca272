#!/usr/bin/env python3
"""Benchmark: ProtoMol nonbonded hot path on one MI355X.

Workload: the 26,000-atom solvated bead polymer in a 64 A periodic box
(500-bead chain + 8,500 flexible 3-site waters), LJ C2 switched 10-12 A,
PME real space Cutoff 12 A, PME 64^3 grid order 4, alpha from accuracy 1e-6
(= 0.257951837636369), leapfrog NVE at 1 fs, Maxwell-Boltzmann 300 K start.

A step is one leapfrog MD step run entirely on the device: kick + drift,
the full nonbonded force pass (cluster-pair LJ + PME real space with exact
cutoff membership, B-spline spread, 3-D FFT convolution, force gather, FP64
exclusion correction), harmonic bonds and angles.  The cell sort and
cluster-pair list are rebuilt whenever an atom has moved more than half the
pair-list skin (chunks that cross it are rolled back and replayed), so no
step uses a stale list.  Positions and velocities stay in HBM.

value = ns/day of the mixed form (FP32 pair/spline math, int64 fixed-point
accumulation, FP64 corrections/bonded/integrator); the FP64 form is
measured in the same run and reported beside it, with the RMS relative force
error of the mixed form against it and both forms' NVE energy drift.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 69          # algorithmic FLOPs per in-cutoff pair (LJ C2 + PME real, counted once)
# HBM bytes per k_pair launch (mixed form, skin 1.5 A): rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0,
# scales from a plain 256 MiB copy (profiles/k_pair_traffic_polymer26000.md)
PAIR_TRAFFIC_BYTES = 41.65e6
PAIR_TRAFFIC_SKIN = 1.5
LAUNCH_KNOBS = {2: "fft_lines", 3: "pair_blocks_per_cu", 4: "pair_schedule", 6: "spread_on_second_stream",
                7: "list_prune", 8: "fixups_on_main_stream", 9: "spread_fused_with_kick_drift", 10: "fft_planes"}
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
TIMESTEP_FS = 1.0


def workload(ctx):
    """The generated structure, relaxed by 300 capped steepest-descent steps
    on the full force field (removes lattice overlaps), then 300 K
    Maxwell-Boltzmann velocities."""
    from protomol3_amd import pmx, prep, systems
    s = systems.polymer_water(n_total=26000, box=64.0, n_beads=500, seed=1)
    spec = pmx.ForceSpec(6.0, lj=(pmx.SW_C2, 10.0, 12.0),
                         pme=dict(rc=12.0, grid=(64, 64, 64), order=4, alpha=-1.0, accuracy=1e-6))
    s.pos = prep.relax(ctx, s, spec)
    s.with_velocities(300.0, seed=2)
    return s, spec


def ns_day(ms_per_step: float) -> float:
    return 86400.0 / (ms_per_step * 1e-3) * TIMESTEP_FS * 1e-6


def _round(x, nd):
    return None if x is None else round(x, nd)


def run_form(ctx, s, spec, precision, steps, warmup, skin, drift_steps, tune_launch=False):
    from protomol3_amd import pmx, tune
    sy = pmx.System(ctx, s, spec, precision)
    try:
        sy.md_init(s, TIMESTEP_FS, skin=skin)
        launch = tune.tune_launch(sy, s.pos, s.vel) if tune_launch else None
        sy.md_set_state(s.pos, s.vel)
        sy.md_run(warmup)
        st0 = sy.md_stats()
        t0 = time.perf_counter()
        sy.md_run(steps)
        t1 = time.perf_counter()
        st1 = sy.md_stats()
        # same number of steps again with per-stage HIP events recorded
        sy.timing_begin()
        sy.md_run(steps)
        stages, ncalls = sy.timing_end()
        x, v, e = sy.md_get_state()
        # in-cutoff pair count at this configuration (for the FLOP figure)
        sy.evaluate(x, flags=pmx.EVAL_PAIRSTATS)
        npairs, _ = sy.pair_stats()
        # NVE drift over drift_steps from here
        slope, etot = None, [None]
        if drift_steps > 0:
            sy.md_set_state(x, v)
            sy.md_run(1)
            _, _, e0 = sy.md_get_state()
            etot = [sum(e0.values())]
            chunk = max(1, drift_steps // 10)
            for _ in range(10):
                sy.md_run(chunk)
                _, _, ek = sy.md_get_state()
                etot.append(sum(ek.values()))
            tps = np.arange(len(etot)) * chunk * TIMESTEP_FS * 1e-6  # ns
            slope = float(np.polyfit(tps, etot, 1)[0])
    finally:
        sy.close()
    return dict(ms_per_step=(t1 - t0) * 1e3 / steps, stages=stages, ncalls=ncalls, npairs=npairs,
                rebuilds=st1["rebuilds"] - st0["rebuilds"], rollbacks=st1["rollbacks"] - st0["rollbacks"],
                x=x, drift_kcal_mol_ns=slope, etot_first=etot[0], etot_last=etot[-1], launch=launch)


def large(ctx, steps: int, skin: float, log=print):
    """The large water box: 333,000 3-site waters (999,000 atoms) in a 216 A
    box, LJ C2 8-10 A, PME real Cutoff 10 A, 256^3 grid, order 6, alpha 0.34;
    relaxed by 100 capped steepest-descent steps, 300 K velocities; `steps`
    leapfrog steps of each form, timed per stage; mixed vs FP64 force error
    at the end configuration."""
    from protomol3_amd import pmx, prep, systems
    t0 = time.perf_counter()
    s = systems.big_water()
    spec = pmx.ForceSpec(5.0, lj=(pmx.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(256, 256, 256), order=6, alpha=0.34))
    s.pos = prep.relax(ctx, s, spec, iters=100, precision=pmx.MIXED)
    s.with_velocities(300.0, seed=3)
    log(f"large: {s.n} atoms built and relaxed in {time.perf_counter() - t0:.1f} s")
    out = {"workload": "water999000: 216 A box, LJ C2 8-10 A, PME 256^3 order 6 rc 10 A alpha 0.34, leapfrog 1 fs",
           "atoms": s.n, "steps": steps, "skin_A": skin}
    xs = None
    for prec, name in ((pmx.MIXED, "mixed"), (pmx.FP64, "fp64")):
        r = run_form(ctx, s, spec, prec, steps, 20, skin, 0)
        log(f"large {name}: {r['ms_per_step']:.3f} ms/step rebuilds {r['rebuilds']} stages {r['stages']}")
        out[name] = {"ns_per_day": round(ns_day(r["ms_per_step"]), 3), "ms_per_step": round(r["ms_per_step"], 4),
                     "stage_ms": {k: round(v, 4) for k, v in r["stages"].items()}, "rebuilds": r["rebuilds"],
                     "pairs_per_step": int(r["npairs"])}
        if xs is None:
            xs = r["x"]
    f = {}
    for prec in (pmx.FP64, pmx.MIXED):
        sy = pmx.System(ctx, s, spec, prec)
        f[prec] = sy.evaluate(xs)[0]
        sy.close()
    d = f[pmx.MIXED] - f[pmx.FP64]
    out["rms_rel_force_err_mixed_vs_fp64"] = float(
        f"{np.sqrt(np.mean(np.sum(d * d, 1))) / np.sqrt(np.mean(np.sum(f[pmx.FP64] ** 2, 1))):.3e}")
    return out


def water1536_leg(ctx, steps: int = 1000, cpu_steps: int = 20, log=print):
    """The 1,536-atom water box (512 flexible 3-site waters in a 24.8 A box,
    LJ C2 8-10 A on O, PME 32^3 order 4 alpha 0.32 rc 10 A), leapfrog NVE
    1 fs from 300 K: `steps` steps of each form; mixed vs FP64 force error
    at the end configuration; the oracle's FP64 MD step on one thread for
    cpu_steps steps beside it."""
    from protomol3_amd import pmx, systems
    s = systems.water_box().with_velocities(300.0, seed=4)
    spec = pmx.ForceSpec(5.0, lj=(pmx.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
    out = {"workload": "water1536: 24.8 A box, LJ C2 8-10 A, PME 32^3 order 4 rc 10 A alpha 0.32, leapfrog NVE 1 fs",
           "atoms": s.n, "steps": steps}
    xs = None
    for prec, name in ((pmx.MIXED, "mixed"), (pmx.FP64, "fp64")):
        r = run_form(ctx, s, spec, prec, steps, 50, 1.0, steps)
        log(f"water1536 {name}: {r['ms_per_step']:.4f} ms/step stages {r['stages']}")
        out[name] = {"ns_per_day": round(ns_day(r["ms_per_step"]), 3), "ms_per_step": round(r["ms_per_step"], 5),
                     "nve_drift_kcal_mol_per_ns": _round(r["drift_kcal_mol_ns"], 3), "pairs_per_step": int(r["npairs"])}
        xs = r["x"] if xs is None else xs
    out["mixed_over_fp64"] = round(out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"], 3)
    f = {}
    for prec in (pmx.FP64, pmx.MIXED):
        sy = pmx.System(ctx, s, spec, prec)
        f[prec] = sy.evaluate(xs)
        sy.close()
    d = f[pmx.MIXED][0] - f[pmx.FP64][0]
    out["rms_rel_force_err_mixed_vs_fp64"] = float(
        f"{np.sqrt(np.mean(np.sum(d * d, 1))) / np.sqrt(np.mean(np.sum(f[pmx.FP64][0] ** 2, 1))):.3e}")
    e64 = f[pmx.FP64][1] + f[pmx.FP64][2]
    out["rel_energy_err_mixed_vs_fp64"] = float(f"{abs(f[pmx.MIXED][1] + f[pmx.MIXED][2] - e64) / abs(e64):.3e}")
    if cpu_steps > 0:
        from oracle import md
        from oracle import oracle as O
        ospec = O.make_spec(5.0, lj=(O.SW_C2, 8.0, 10.0), pme=dict(rc=10.0, grid=(32, 32, 32), order=4, alpha=0.32))
        x, v, fo, _ = md.leapfrog_run(s, ospec, s.pos, s.vel, TIMESTEP_FS, 1, nthreads=1)
        t0 = time.perf_counter()
        md.leapfrog_run(s, ospec, x, v, TIMESTEP_FS, cpu_steps, f=fo, nthreads=1)
        dt = (time.perf_counter() - t0) / cpu_steps
        out["cpu_baseline"] = {"value": round(ns_day(dt * 1e3), 5), "unit": "ns/day", "cores": 1, "kind": "port",
                               "sample": f"{cpu_steps} leapfrog MD steps of the oracle (FP64), {dt:.4f} s per step"}
    return out


def copies(s, spec, ncopies: int, steps: int, lanes: int, skin: float, log=print):
    """ncopies independent copies of the 26,000-atom workload (same start
    positions, velocities from per-copy seeds), leapfrog NVE, `steps` steps
    each: in the mixed form by one pmx_md_run_batch call (every kernel of a
    step launched once for all copies) and by one pmx_md_run_many call over
    `lanes` contexts (each copy's kernels on its context's streams), and in
    the FP64 form batched; aggregate ns/day = copies x simulated time / wall
    time."""
    from protomol3_amd import pmx
    v0 = s.vel
    vels = [s.with_velocities(300.0, seed=100 + i).vel.copy() for i in range(ncopies)]
    s.vel = v0

    def leg(precision, how):
        ctxs = [pmx.Context(0) for _ in range(lanes if how == "many" else 1)]
        sys_ = []
        try:
            for i, vel in enumerate(vels):
                sy = pmx.System(ctxs[i % len(ctxs)], s, spec, precision)
                sys_.append(sy)
                sy.md_init(s, TIMESTEP_FS, skin=skin)
                sy.md_set_state(s.pos, vel)
            run = pmx.run_batch if how == "batch" else pmx.run_many
            run(sys_, 20)
            t0 = time.perf_counter()
            run(sys_, steps)
            wall = time.perf_counter() - t0
        finally:
            for sy in sys_:
                sy.close()
            for c in ctxs:
                c.close()
        return wall, ncopies * steps * TIMESTEP_FS * 1e-6 / wall * 86400.0

    wall, agg = leg(pmx.MIXED, "batch")
    _, agg_many = leg(pmx.MIXED, "many")
    _, agg64 = leg(pmx.FP64, "batch")
    log(f"copies: {ncopies} x {s.n} atoms batched: mixed {agg:.1f} ns/day aggregate ({wall:.2f} s), FP64 "
        f"{agg64:.1f}; over {lanes} contexts {agg_many:.1f}")
    return {"copies": ncopies, "steps": steps, "skin_A": skin, "wall_s": round(wall, 3),
            "aggregate_ns_per_day": round(agg, 3), "fp64_aggregate_ns_per_day": round(agg64, 3),
            "mixed_over_fp64_batched": round(agg / agg64, 3),
            "per_context_streams": {"lanes": lanes, "aggregate_ns_per_day": round(agg_many, 3)}}


def drift_study(ctx, s, spec, ps: float = 10.0, log=print):
    """NVE total-energy drift of the 26,000-atom workload against the time
    step: for each form and h in (1, 0.5, 0.25) fs, ps picoseconds of
    leapfrog from the same start, total energy sampled every 0.1 ps, slope of
    a least-squares line (kcal/mol/ns); and at 1 fs the same run with the
    pair list rebuilt every step (skin 0) beside the tuned skin of 1.5 A.
    Leapfrog's energy error is O(h^2), so a drift caused by the time step
    falls about 4x per halving of h, and one caused by a force that is not the
    gradient of the energy does not."""
    from protomol3_amd import pmx
    out = []
    cases = [(pmx.FP64, h, 1.5) for h in (1.0, 0.5, 0.25)] + [(pmx.FP64, 1.0, 0.0)] + \
            [(pmx.MIXED, h, 1.5) for h in (1.0, 0.5, 0.25)]
    for prec, h, skin in cases:
        sy = pmx.System(ctx, s, spec, prec)
        try:
            sy.md_init(s, h, skin=skin)
            sy.md_set_state(s.pos, s.vel)
            per = int(round(100.0 / h))  # steps per 0.1 ps
            nsamp = int(round(ps * 10))
            t0 = time.perf_counter()
            sy.md_run(1)
            _, _, e = sy.md_get_state()
            etot, kin = [sum(e.values())], [e.get("kinetic", float("nan"))]
            for _ in range(nsamp):
                sy.md_run(per)
                _, _, e = sy.md_get_state()
                etot.append(sum(e.values()))
                kin.append(e.get("kinetic", float("nan")))
            wall = time.perf_counter() - t0
        finally:
            sy.close()
        tns = np.arange(len(etot)) * 0.1e-3
        slope = float(np.polyfit(tns, etot, 1)[0])
        r = {"form": "mixed" if prec == pmx.MIXED else "fp64", "timestep_fs": h, "skin_A": skin, "ps": ps,
             "drift_kcal_mol_per_ns": round(slope, 3), "etot_start": round(etot[0], 4),
             "etot_end": round(etot[-1], 4), "etot_std": round(float(np.std(etot)), 4),
             "kinetic_start": round(kin[0], 3), "kinetic_end": round(kin[-1], 3), "wall_s": round(wall, 2)}
        log(f"drift: {r}")
        out.append(r)
    return {"workload": "polymer_water26000 NVE leapfrog from the relaxed 300 K start", "runs": out}


def time_pair(ctx, s, spec, x, skin, launch=None, reps: int = 200) -> float:
    """Average device ms of one k_pair launch (mixed form) on the list an MD
    step with this skin builds at positions x: the kernel alone, launched
    back to back between two HIP events on its stream, with the launch
    shapes the MD run used (launch = (chosen knobs, timings) or None)."""
    from protomol3_amd import pmx
    sy = pmx.System(ctx, s, spec, pmx.MIXED)
    try:
        sy.md_init(s, TIMESTEP_FS, skin=skin)
        for k, v in (launch[0].items() if launch else ()):
            sy.set_tuning(k, v)
        sy.md_set_state(x, s.vel)
        sy.time_pair(10)
        return sy.time_pair(reps)
    finally:
        sy.close()


def serving(nreq: int, slots: int, seed: int = 7, pool: int = 256, rate: float = 0.0, log=print):
    """Serving stream: requests of log-uniform size in [1,000, 100,000] atoms
    (water + bead-polymer solute at 0.0334 atoms/A^3, per-system seeds),
    half single force evaluations and half 100-step Langevin trajectories
    (300 K, 1/ps, per-request noise seeds), mixed precision, LJ C2 7-9 A +
    PME ~1 A grid.  Up to `pool` distinct systems are built; request i uses
    system i mod pool.  First the first min(nreq, pool) requests at once
    (capacity), then all nreq requests as Poisson arrivals at `rate`
    requests/s (0: 75% of that capacity); latency = submit to result, per
    request."""
    fixed_rate = rate
    from protomol3_amd import pmx, serve, systems
    rng = np.random.default_rng(seed)
    npool = min(nreq, pool)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(100000), npool)).astype(int)
    kinds = rng.integers(0, 2, nreq)
    built = []
    for i, n in enumerate(sizes):
        s = systems.replica_box(int(n), 1000 + i).with_velocities(300.0, seed=5000 + i)
        built.append((s, serve.request_spec(s)))
    reqs = [(built[i % npool][0], built[i % npool][1], int(kinds[i]), i) for i in range(nreq)]
    natoms = int(sum(s.n for s, _, _, _ in reqs))
    log(f"serving: {nreq} requests over {npool} systems, {natoms} atoms in total, {slots} slots")

    def submit(srv, s, spec, kd, i):
        if kd == 0:
            return srv.submit_forces(s, spec, pmx.MIXED)
        return srv.submit_trajectory(s, spec, pmx.MIXED, 100, skin=1.0, gamma=1.0, seed=17 + i)

    srv = serve.Server(0, slots=slots)
    try:
        t0 = time.perf_counter()
        ts = [submit(srv, *r) for r in reqs[:npool]]
        for t in ts:
            srv.wait(t)
        makespan = time.perf_counter() - t0
        capacity = npool / makespan
        rate = rate if rate > 0 else 0.75 * capacity
        gaps = rng.exponential(1.0 / rate, nreq)
        t0 = time.perf_counter()
        ts, due = [], t0
        for k, (r, g) in enumerate(zip(reqs, gaps)):
            due += g
            now = time.perf_counter()
            if due > now:
                time.sleep(due - now)
            ts.append(submit(srv, *r))
            if k and k % 1000 == 0:
                log(f"serving: {k} requests submitted")
        lat = np.array([srv.wait(t).done_ms for t in ts])
        wall = time.perf_counter() - t0
        peak_bytes, peak_running = srv.stats()
    finally:
        srv.close()
    traj_steps = 100 * sum(1 for _, _, kd, _ in reqs if kd == 1)
    return {"requests": nreq, "distinct_systems": npool, "slots": slots, "atoms_total": natoms,
            "capacity_req_per_s": round(capacity, 3), "arrival_rate_req_per_s": round(rate, 3),
            "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
            "aggregate_ns_per_day": round(traj_steps * 1e-6 / wall * 86400.0, 3),
            "wall_s": round(wall, 2), "peak_device_bytes": int(peak_bytes), "peak_running": int(peak_running),
            "workload": "log-uniform 1k-100k atoms, 50% force evaluations / 50% 100-step Langevin 1 fs, mixed, "
                        + ("Poisson arrivals at 75% of measured capacity" if fixed_rate <= 0 else
                           f"Poisson arrivals at {fixed_rate} requests/s")}


def replicas(nsys: int, steps: int, lanes: int, seed: int = 11, skin: float = 1.0, graphs: bool = False,
             how: str = "batch", precision: int = 1, log=print):
    """Batched replicas: nsys independent systems with atom counts drawn
    log-uniform in [500, 20,000] (water + 5-10% bead-polymer solute at 0.0334
    atoms/A^3, per-system seeds), LJ C2 7-9 A + PME ~1 A grid, mixed
    precision, `steps` Langevin steps (300 K, 1/ps) each, all run by one
    pmx_md_run_batch call (how="batch": every kernel of a step launched once
    for all systems) or one pmx_md_run_many call over `lanes` contexts
    (how="many"); precision 1 is the mixed form, 0 FP64."""
    from protomol3_amd import pmx, serve, systems
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(500), np.log(20000), nsys)).astype(int)
    t0 = time.perf_counter()
    boxes = [systems.replica_box(int(n), 2000 + i).with_velocities(300.0, seed=3000 + i) for i, n in enumerate(sizes)]
    natoms = int(sum(b.n for b in boxes))
    log(f"replicas: {nsys} systems, {natoms} atoms in total, built in {time.perf_counter() - t0:.1f} s")
    ctxs = [pmx.Context(0) for _ in range(lanes if how == "many" else 1)]
    run = pmx.run_batch if how == "batch" else pmx.run_many
    sys_ = []
    try:
        for i, b in enumerate(boxes):
            sy = pmx.System(ctxs[i % len(ctxs)], b, serve.request_spec(b), precision)
            sys_.append(sy)
            sy.md_init(b, TIMESTEP_FS, skin=skin, integrator=pmx.LANGEVIN, temperature=300.0, gamma=1.0, seed=i,
                       graphs=graphs)
            sy.md_set_state(b.pos, b.vel)
        run(sys_, 5)  # warm-up (first list builds, kernel loads)
        st0 = [sy.md_stats() for sy in sys_]
        t0 = time.perf_counter()
        run(sys_, steps)
        wall = time.perf_counter() - t0
        st1 = [sy.md_stats() for sy in sys_]
        rebuilds = sum(b["rebuilds"] - a["rebuilds"] for a, b in zip(st0, st1))
        rollbacks = sum(b["rollbacks"] - a["rollbacks"] for a, b in zip(st0, st1))
    finally:
        for sy in sys_:
            sy.close()
        for c in ctxs:
            c.close()
    return {"systems": nsys, "atoms_total": natoms, "steps": steps, "launches": how, "lanes": len(ctxs),
            "wall_s": round(wall, 3), "skin_A": skin, "rebuilds": rebuilds, "rollbacks": rollbacks,
            "aggregate_ns_per_day": round(nsys * steps * TIMESTEP_FS * 1e-6 / wall * 86400.0, 3),
            "atom_steps_per_s": float(f"{natoms * steps / wall:.4e}"),
            "workload": "log-uniform 500-20,000 atoms, water + bead polymer at 0.0334 atoms/A^3, rc 9 A, "
                        f"PME ~1 A grid, Langevin 300 K 1/ps, {'mixed' if precision else 'FP64'}, one "
                        f"pmx_md_run_{how} call"}


# Reference build / oracle time per force evaluation of the 26,000-atom workload
# on one core of the same host (profiles/cpu_calibration.md, oracle/calibrate.py)
REF_OVER_ORACLE_1T = 1.015


def cpu_baseline(s, budget_s: float = 20.0):
    """The oracle (FP64 C++ restatement of the reference path + leapfrog) on
    the host: whole MD steps of the same workload, first on one thread, then
    on every core the box gives this job (at most 16), about budget_s in all.
    The reference itself cannot run on the GPU host; its speed there is
    estimated from the serial figure with the reference/oracle time ratio
    measured on one core of another host."""
    from oracle import md
    from oracle import oracle as O
    nproc = os.cpu_count() or 1
    cores = min(nproc, 16)
    ospec = O.make_spec(6.0, lj=(O.SW_C2, 10.0, 12.0),
                        pme=dict(rc=12.0, grid=(64, 64, 64), order=4, alpha=-1.0, accuracy=1e-6))
    x, v, f, _ = md.leapfrog_run(s, ospec, s.pos, s.vel, TIMESTEP_FS, 1, nthreads=cores)  # relaxed start

    def steps_for(nthreads, budget):
        nonlocal x, v, f
        n, t0 = 0, time.perf_counter()
        while True:
            x, v, f, _ = md.leapfrog_run(s, ospec, x, v, TIMESTEP_FS, 1, f=f, nthreads=nthreads)
            n += 1
            if time.perf_counter() - t0 > budget or n >= 50:
                return n, (time.perf_counter() - t0) / n

    n1, dt1 = steps_for(1, budget_s / 3.0)
    nc, dtc = steps_for(cores, 2.0 * budget_s / 3.0)
    return dict(value=round(ns_day(dtc * 1e3), 5), unit="ns/day", cores=cores, kind="port", nproc=nproc,
                sample=f"{nc} leapfrog MD steps of the 26,000-atom workload (oracle FP64 nonbonded on {cores} "
                       f"threads + bonded + integrator), {dtc:.3f} s per step; host has {nproc} logical cores",
                serial={"value": round(ns_day(dt1 * 1e3), 5), "cores": 1, "s_per_step": round(dt1, 4),
                        "steps": n1},
                reference_serial_estimate={
                    "value": round(ns_day(dt1 * REF_OVER_ORACLE_1T * 1e3), 5), "cores": 1,
                    "ref_over_oracle_1thread": REF_OVER_ORACLE_1T,
                    "note": "serial oracle time x the reference/oracle ratio measured on identical input "
                            "(profiles/cpu_calibration.md); no MPI toolchain on the host"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--skin", type=float, default=-1.0, help="pair-list skin in A (<0: autotune)")
    ap.add_argument("--drift-steps", type=int, default=10000)
    ap.add_argument("--no-tune-launch", action="store_true",
                    help="keep the default launch shapes (no tile/blocks/FFT-lines search)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--serve-requests", type=int, default=64, help="requests in the serving stream (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-rate", type=float, default=0.0,
                    help="Poisson arrival rate in requests/s (0: 75%% of the measured capacity)")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving stream (--serve-requests requests) and print it")
    ap.add_argument("--replicas", type=int, default=256,
                    help="batched replicas (log-uniform 500-20,000 atoms) run after the main metric (0: skip)")
    ap.add_argument("--copies", type=int, default=8,
                    help="copies of the 26,000-atom workload run side by side for an aggregate rate (0: skip)")
    ap.add_argument("--no-water1536", action="store_true", help="skip the 1,536-atom water box leg")
    ap.add_argument("--large", action="store_true",
                    help="only the 999,000-atom water box, both forms timed per stage (--steps steps, default 1000)")
    ap.add_argument("--pair-only", action="store_true",
                    help="only time the pair kernel alone (at --skin, default 0.7 A) and print it")
    ap.add_argument("--drift-study", action="store_true",
                    help="only the NVE drift against the time step (1, 0.5, 0.25 fs; skin 0 and 1.5 A)")
    ap.add_argument("--replica-steps", type=int, default=500)
    ap.add_argument("--replica-lanes", type=int, default=4)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {a.gpus}); "
              "sharding the force path over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    from protomol3_amd import pmx, tune
    if a.serve_only:
        log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
        print(json.dumps({"serving": serving(a.serve_requests, a.serve_slots, rate=a.serve_rate, log=log)}),
              flush=True)
        return
    ctx = pmx.Context(0)
    if a.large:
        log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
        out = large(ctx, a.steps if a.steps != 400 else 1000, a.skin if a.skin >= 0 else 1.2, log=log)
        print(json.dumps(out), flush=True)
        ctx.close()
        return
    s, spec = workload(ctx)
    log = lambda m: print(f"bench: {m}", file=sys.stderr, flush=True)
    log(f"{s.n} atoms, steps={a.steps} warmup={a.warmup}")
    if a.drift_study:
        print(json.dumps(drift_study(ctx, s, spec, log=log)), flush=True)
        ctx.close()
        return
    if a.pair_only:
        skin = a.skin if a.skin >= 0 else 0.7
        ms = time_pair(ctx, s, spec, s.pos, skin)
        print(json.dumps({"kernel": "k_pair<mixed,PME,C2,Cutoff>", "skin_A": skin, "avg_launch_ms": round(ms, 5)}),
              flush=True)
        ctx.close()
        return
    skins = {}
    if a.skin < 0:
        skin_m, times_m = tune.tune_skin(ctx, s, spec, pmx.MIXED, s.pos, s.vel)
        skin_d, times_d = tune.tune_skin(ctx, s, spec, pmx.FP64, s.pos, s.vel)
        skins = {"mixed": times_m, "fp64": times_d}
        log(f"skin tuning (ms/step) mixed {times_m} fp64 {times_d}")
    else:
        skin_m = skin_d = a.skin
    tl = not a.no_tune_launch
    mixed = run_form(ctx, s, spec, pmx.MIXED, a.steps, a.warmup, skin_m, a.drift_steps, tl)
    log(f"mixed {mixed['ms_per_step']:.4f} ms/step rebuilds {mixed['rebuilds']} rollbacks {mixed['rollbacks']} "
        f"stages {mixed['stages']}")
    fp64 = run_form(ctx, s, spec, pmx.FP64, a.steps, a.warmup, skin_d, a.drift_steps, tl)
    log(f"fp64 {fp64['ms_per_step']:.4f} ms/step rebuilds {fp64['rebuilds']} stages {fp64['stages']}")
    # force accuracy of the mixed form: nonbonded forces of both forms at the
    # same configuration (the end of the mixed run)
    forms = {}
    for prec in (pmx.FP64, pmx.MIXED):
        sy = pmx.System(ctx, s, spec, prec)
        forms[prec] = sy.evaluate(mixed["x"])
        sy.close()
    pair_launch_ms = time_pair(ctx, s, spec, mixed["x"], skin_m, mixed["launch"])
    small_out = None if a.no_water1536 else water1536_leg(ctx, cpu_steps=0 if a.no_cpu else 20, log=log)
    ctx.close()
    f64, fm = forms[pmx.FP64][0], forms[pmx.MIXED][0]
    e64, em = forms[pmx.FP64], forms[pmx.MIXED]
    rel_e = {"coulomb": float(f"{abs(em[1] - e64[1]) / abs(e64[1]):.3e}"),
             "lj": float(f"{abs(em[2] - e64[2]) / abs(e64[2]):.3e}")}
    err = float(np.sqrt(np.mean(np.sum((fm - f64) ** 2, 1))) / np.sqrt(np.mean(np.sum(f64 ** 2, 1))))
    from protomol3_amd import outputs
    cmp = outputs.compare_forces(f64, e64[1] + e64[2], fm, em[1] + em[2], s.mass)
    pair_ms = pair_launch_ms
    achieved = FLOP_PER_PAIR * mixed["npairs"] / (pair_ms * 1e-3) / 1e12
    cpu = None if a.no_cpu else cpu_baseline(s, a.cpu_budget)
    serve_out = serving(a.serve_requests, a.serve_slots, log=log) if a.serve_requests > 0 else None
    rep_out = None
    if a.replicas > 0:
        rep_out = replicas(a.replicas, a.replica_steps, a.replica_lanes, log=log)
        r64 = replicas(a.replicas, a.replica_steps, a.replica_lanes, precision=0, log=log)
        many = replicas(a.replicas, a.replica_steps, a.replica_lanes, how="many", log=log)
        rep_out["fp64_aggregate_ns_per_day"] = r64["aggregate_ns_per_day"]
        rep_out["mixed_over_fp64_batched"] = round(rep_out["aggregate_ns_per_day"] / r64["aggregate_ns_per_day"], 3)
        rep_out["per_context_streams"] = {k: many[k] for k in ("lanes", "wall_s", "aggregate_ns_per_day")}
        log(f"replicas: batched {rep_out['aggregate_ns_per_day']:.1f} ns/day (FP64 {r64['aggregate_ns_per_day']:.1f}), "
            f"per-context streams {many['aggregate_ns_per_day']:.1f}")
    copies_out = copies(s, spec, a.copies, a.steps, a.replica_lanes, skin_m, log=log) if a.copies > 0 else None
    if copies_out is not None:
        copies_out["over_fp64_single"] = round(
            copies_out["aggregate_ns_per_day"] / ns_day(fp64["ms_per_step"]), 3)
    out = {
        "metric": "ns/day (26k-atom PME) mixed vs FP64 GPU at <=1e-5 force err; p50/p99 latency",
        "value": round(ns_day(mixed["ms_per_step"]), 3),
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "dtype_detail": "mixed form: f32 pair/spline arithmetic, i64 fixed-point force/grid accumulation, f64 "
                        "corrections and integrator; the f64 form is measured alongside",
        "data": "synthetic (seeded 26,000-atom bead polymer + water, random orientations, relaxed by 300 "
                "capped steepest-descent steps, 300 K velocities)",
        "config": {"workload": "polymer_water26000: 64 A box, LJ C2 10-12 A, PME 64^3 order 4 rc 12 A "
                               "alpha 0.257951837636369, leapfrog NVE 1 fs; step = full device MD step",
                   "global_batch": 1, "seq_len": s.n, "parallelism": "single GPU"},
        "fp64_ns_per_day": round(ns_day(fp64["ms_per_step"]), 3),
        "fp64_ms_per_step": round(fp64["ms_per_step"], 5),
        "mixed_over_fp64": round(fp64["ms_per_step"] / mixed["ms_per_step"], 3),
        "rms_rel_force_err_mixed_vs_fp64": float(f"{err:.3e}"),
        "rel_energy_err_mixed_vs_fp64": rel_e,
        "compare_mixed_vs_fp64": {k: float(f"{v:.4e}") for k, v in cmp.items()},
        "nve_drift_kcal_mol_per_ns": {"mixed": _round(mixed["drift_kcal_mol_ns"], 3),
                                      "fp64": _round(fp64["drift_kcal_mol_ns"], 3),
                                      "steps": a.drift_steps},
        "launch_tuning": {form: None if r["launch"] is None else
                          {"chosen": {LAUNCH_KNOBS[k]: v for k, v in r["launch"][0].items()},
                           "ms_per_step": {LAUNCH_KNOBS[k]: {str(val): round(ms, 5) for val, ms in t.items()}
                                           for k, t in r["launch"][1].items()}}
                          for form, r in (("mixed", mixed), ("fp64", fp64))},
        "pair_list": {"skin_A": {"mixed": skin_m, "fp64": skin_d}, "tuning_ms_per_step": skins,
                      "rebuilds_in_timed_steps": {"mixed": mixed["rebuilds"], "fp64": fp64["rebuilds"]},
                      "rollbacks_in_timed_steps": {"mixed": mixed["rollbacks"], "fp64": fp64["rollbacks"]}},
        "pairs_per_step": int(mixed["npairs"]),
        "pair_interactions_per_s": float(f"{mixed['npairs'] / (mixed['ms_per_step'] * 1e-3):.4e}"),
        "stage_ms_mixed": {k: round(v, 5) for k, v in mixed["stages"].items()},
        "stage_ms_fp64": {k: round(v, 5) for k, v in fp64["stages"].items()},
        "roofline": {"bound": "valu", "kernel": "k_pair<mixed,PME,C2,Cutoff>",
                     "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": PAIR_TRAFFIC_BYTES,
                     "traffic_unit": "bytes per launch (rocprofv3 PMC, skin %.1f A)" % PAIR_TRAFFIC_SKIN,
                     "flop_per_pair": FLOP_PER_PAIR, "pairs_per_launch": int(mixed["npairs"]),
                     "avg_launch_ms": round(pair_ms, 5),
                     "avg_launch_ms_in_md_step": round(mixed["stages"]["pair"], 5)},
        "cpu_baseline": cpu,
        "serving": serve_out,
    }
    out["water1536"] = small_out
    out["replicas"] = rep_out
    out["copies_polymer26000"] = copies_out
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
