#!/usr/bin/env python3
"""Benchmark: ProtoMol nonbonded hot path on one MI355X.

Workload: the 26,000-atom solvated bead polymer in a 64 A periodic box
(500-bead chain + 8,500 flexible 3-site waters), LJ C2 switched 10-12 A,
PME real space Cutoff 12 A, PME 64^3 grid order 4, alpha from accuracy 1e-6
(= 0.257951837636369), 1 fs per step.

A step is one full nonbonded force pass over the resident system: cell sort,
cluster-pair list build, LJ + PME real-space pair kernel, B-spline spread,
3-D FFT convolution, force gather, FP64 exclusion correction, and the
accumulator-to-FP64 force conversion.  Positions stay in HBM; nothing is
cached between steps (the sort, list and grids are rebuilt every step, as the
reference rebuilds its cell lists every evaluation).

Prints one JSON line.  value = ns/day of the mixed form (FP32 pair/spline
math, int64 fixed-point accumulation); the FP64 form is measured in the same
run and reported beside it.
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
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector
HBM_PEAK_GBS = 8000.0


def workload():
    from protomol3_amd import systems
    s = systems.polymer_water(n_total=26000, box=64.0, n_beads=500, seed=1)
    spec = dict(cellsize=6.0, lj=(3, 10.0, 12.0), pme=dict(rc=12.0, grid=(64, 64, 64), order=4, alpha=-1.0,
                                                            accuracy=1e-6))
    return s, spec


def run_form(ctx, s, spec, precision, steps, warmup):
    from protomol3_amd import pmx
    gs = pmx.ForceSpec(spec["cellsize"], lj=spec["lj"], pme=spec["pme"])
    sysh = pmx.System(ctx, s, gs, precision)
    sysh.set_positions(s.pos)
    # pair count for the algorithmic FLOP figure (outside the timed region)
    f, ec, elj = sysh.evaluate(s.pos, flags=pmx.EVAL_PAIRSTATS)
    npairs, _ = sysh.pair_stats()
    sysh.set_positions(s.pos)
    for _ in range(warmup):
        sysh.compute()
    sysh.sync()
    # timed region: K back-to-back steps, events per stage on the system's stream
    sysh.timing_begin()
    t0 = time.perf_counter()
    for _ in range(steps):
        sysh.compute(pmx.EVAL_TIMING)
    sysh.sync()
    t1 = time.perf_counter()
    stages, ncalls = sysh.timing_end()
    f2, ec2, elj2 = sysh.get_forces()
    sysh.close()
    return dict(ms_per_step=(t1 - t0) * 1e3 / steps, stages=stages, ncalls=ncalls, npairs=npairs,
                forces=f2, ecoul=ec2, elj=elj2)


def cpu_baseline(s, spec, budget_s: float = 20.0):
    """The oracle (FP64 C++ restatement of the reference path) on the host's
    cores: whole force evaluations of the same workload until ~budget_s."""
    from oracle import oracle as O
    cores = min(os.cpu_count() or 1, 16)
    p = spec["pme"]
    ospec = O.make_spec(spec["cellsize"], lj=spec["lj"], pme=dict(p))
    O.evaluate(s, ospec, nthreads=cores)  # warm
    n = 0
    t0 = time.perf_counter()
    while True:
        O.evaluate(s, ospec, nthreads=cores)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 50:
            break
    dt = (time.perf_counter() - t0) / n
    return dict(value=86400.0 / dt * 1e-6, unit="ns/day", cores=cores, kind="port",
                sample=f"{n} full force evaluations (LJ C2 + PME real/recip/correction) of the 26,000-atom "
                       f"workload on {cores} threads, {dt:.3f} s each")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU sharding design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    from protomol3_amd import pmx
    s, spec = workload()
    ctx = pmx.Context(0)
    print(f"bench: {s.n} atoms, steps={a.steps} warmup={a.warmup}", file=sys.stderr, flush=True)
    mixed = run_form(ctx, s, spec, pmx.MIXED, a.steps, a.warmup)
    print(f"bench: mixed {mixed['ms_per_step']:.4f} ms/step stages {mixed['stages']}", file=sys.stderr, flush=True)
    fp64 = run_form(ctx, s, spec, pmx.FP64, max(1, a.steps // 2), max(1, a.warmup // 2))
    print(f"bench: fp64 {fp64['ms_per_step']:.4f} ms/step stages {fp64['stages']}", file=sys.stderr, flush=True)
    ctx.close()

    ref = fp64["forces"]
    err = np.sqrt(np.mean(np.sum((mixed["forces"] - ref) ** 2, 1))) / np.sqrt(np.mean(np.sum(ref ** 2, 1)))
    ns_day = lambda ms: 86400.0 / (ms * 1e-3) * 1e-6  # 1 fs per step
    pair_ms = mixed["stages"]["pair"]
    achieved = FLOP_PER_PAIR * mixed["npairs"] / (pair_ms * 1e-3) / 1e12
    cpu = None if a.no_cpu else cpu_baseline(s, spec, a.cpu_budget)
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
        "dtype": "f32+i64 (mixed); f64 reference form measured alongside",
        "data": "synthetic (seeded 26,000-atom polymer + water, random orientations)",
        "config": {"workload": "polymer_water26000: 64 A box, LJ C2 10-12 A, PME 64^3 order 4 rc 12 A "
                               "alpha 0.257951837636369, 1 fs/step; step = full nonbonded force pass",
                   "global_batch": 1, "seq_len": s.n, "parallelism": "single GPU"},
        "fp64_ns_per_day": round(ns_day(fp64["ms_per_step"]), 3),
        "fp64_ms_per_step": round(fp64["ms_per_step"], 5),
        "mixed_over_fp64": round(fp64["ms_per_step"] / mixed["ms_per_step"], 3),
        "rms_rel_force_err_mixed_vs_fp64": float(f"{err:.3e}"),
        "pairs_per_step": int(mixed["npairs"]),
        "pair_interactions_per_s": float(f"{mixed['npairs'] / (mixed['ms_per_step'] * 1e-3):.4e}"),
        "stage_ms_mixed": {k: round(v, 5) for k, v in mixed["stages"].items()},
        "stage_ms_fp64": {k: round(v, 5) for k, v in fp64["stages"].items()},
        "roofline": {"bound": "valu_fp32", "kernel": "k_pair<mixed,PME>", "achieved": round(achieved, 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                     "traffic": None, "flop_per_pair": FLOP_PER_PAIR},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
