#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/*.npz from the reference
ProtoMol path, compiled from the reference sources by oracle/ref/Makefile
(oracle/_ref/pm_ref).  Run from the repository root in the container that
holds /root/reference:

    make -f oracle/ref/Makefile -j8 && python tests/golden/make_fixtures.py

Each fixture holds the inputs (positions; the generator and seed that made
the rest of the system) and the reference's outputs: forces, Coulomb and
Lennard-Jones energies, the PME alpha the reference chose, the cell of every
atom, and the count and set hash of the in-cutoff non-excluded pairs visited
by the reference's cell-pair enumerator.  Large systems keep a strided
subsample of the forces.
"""
from __future__ import annotations

import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from protomol3_amd import charmm, systems, sysfile  # noqa: E402

PM_REF = os.path.join(ROOT, "oracle", "_ref", "pm_ref")
OUT = os.path.join(ROOT, "tests", "golden")
W423 = os.path.join(OUT, "water_423")


def make_system(gen: str, args: dict):
    if gen == "charmm_water423":
        return charmm.build_system(os.path.join(W423, "water_423.psf"), os.path.join(W423, "water_423.pos.pdb"),
                                   os.path.join(W423, "par_all27_prot_lipid.inp"), box=(28.0, 28.0, 28.0))
    return getattr(systems, gen)(**args)


# name: (generator, generator args, cellsize, pm_ref force spec, membership cutoff, force stride)
CASES = {
    "water423_ljc2_coulc1": ("charmm_water423", {}, 6.5, ["coul_c1", "6.5", "lj_c2", "0.1", "6.5"], 6.5, 1),
    "water648_ljc2_coulc1": ("water_box", dict(n_side=6, box=18.6, seed=3), 4.5,
                             ["lj_c2", "0.1", "9", "coul_c1", "9"], 9.0, 1),
    "water1536_ljc2_pme": ("water_box", {}, 5.0,
                           ["lj_c2", "8", "10", "pme", "32", "32", "32", "4", "10", "1e-6", "0.32", "111"], 10.0, 1),
    "water1536_pme_real": ("water_box", {}, 5.0,
                           ["pme", "32", "32", "32", "4", "10", "1e-6", "0.32", "100"], 10.0, 1),
    "water1536_pme_recip": ("water_box", {}, 5.0,
                            ["pme", "32", "32", "32", "4", "10", "1e-6", "0.32", "010"], 10.0, 1),
    "water1536_pme_corr": ("water_box", {}, 5.0,
                           ["pme", "32", "32", "32", "4", "10", "1e-6", "0.32", "001"], 10.0, 1),
    "water1536_ljcut_coulc2": ("water_box", dict(seed=5), 4.0,
                               ["lj_cutoff", "9.5", "coul_c2", "7", "9.5"], 9.5, 1),
    "water1536_pme_order6_grid30": ("water_box", dict(seed=6), 5.0,
                                    ["lj_c2", "8", "10", "pme", "30", "30", "30", "6", "10", "1e-6", "0.34", "111"],
                                    10.0, 1),
    "water1536_pme_c1": ("water_box", dict(seed=7), 5.0,
                         ["pme_c1", "32", "32", "32", "4", "10", "1e-6", "0.32", "111"], 10.0, 1),
    "polymer3200_pme_auto": ("polymer_water", dict(n_total=3200, box=32.0, n_beads=50, seed=4), 6.0,
                             ["lj_c2", "10", "12", "pme", "32", "32", "32", "4", "12", "1e-6", "-1", "111"], 12.0, 1),
    "polymer26000_pme_auto": ("polymer_water", dict(n_total=26000, box=64.0, n_beads=500, seed=1), 6.0,
                              ["lj_c2", "10", "12", "pme", "64", "64", "64", "4", "12", "1e-6", "-1", "111"], 12.0, 13),
    # scaled 1-4 chain pairs (MODIFIED class, coulombScalingFactor 0.5, separate A14/B14)
    "polymer3200_mod14_pme": ("polymer_water", dict(n_total=3200, box=32.0, n_beads=50, seed=4, scaled14=0.5),
                              6.0, ["lj_c2", "10", "12", "pme", "32", "32", "32", "4", "12", "1e-6", "-1", "111"],
                              12.0, 1),
    "polymer3200_mod14_pme_corr": ("polymer_water",
                                   dict(n_total=3200, box=32.0, n_beads=50, seed=4, scaled14=0.5), 6.0,
                                   ["pme", "32", "32", "32", "4", "12", "1e-6", "-1", "001"], 12.0, 1),
    # scaled 1-4 with cutoff Coulomb, LJ and Coulomb cutoffs apart (9 and 11.5 A)
    "polymer3200_mod14_ljc2_coulc2": ("polymer_water",
                                      dict(n_total=3200, box=32.0, n_beads=50, seed=4, scaled14=0.5), 5.0,
                                      ["lj_c2", "7", "9", "coul_c2", "9", "11.5"], 11.5, 1),
    # net charge +0.5 e (odd bead count): the charged-system energy term;
    # LJ cutoff 10 A beside a PME real-space cutoff of 12 A
    "polymer3201_charged_pme": ("polymer_water", dict(n_total=3201, box=32.0, n_beads=51, seed=4), 6.0,
                                ["lj_c2", "8", "10", "pme", "32", "32", "32", "4", "12", "1e-6", "-1", "111"],
                                12.0, 1),
}


def run_case(name: str, gen: str, args: dict, cellsize: float, spec: list, rc: float, stride: int) -> dict:
    s = make_system(gen, args)
    with tempfile.TemporaryDirectory() as td:
        sf, of, pf = (os.path.join(td, x) for x in ("sys", "out", "pairs"))
        sysfile.write_sysfile(sf, s)
        env = dict(os.environ, PM_REF_PAIRS=pf, PM_REF_PAIRS_RC=repr(rc), PM_REF_PAIRS_NOLIST="1")
        r = subprocess.run([PM_REF, sf, of, "cellsize", repr(cellsize), *spec], capture_output=True, text=True,
                           env=env, check=True)
        ec, elj, frc = sysfile.read_forces(of)
        m = re.search(r"alpha=([0-9.eE+-]+)", r.stdout + r.stderr)
        alpha = float(m.group(1)) if m else 0.0
        with open(pf) as f:
            dims = tuple(int(v) for v in f.readline().split()[1:])
            cells = np.zeros(s.n, np.int32)
            for _ in range(s.n):
                i, cx, cy, cz = (int(v) for v in f.readline().split())
                cells[i] = (cx * dims[1] + cy) * dims[2] + cz
            head = f.readline().split()
            npairs, phash = int(head[1]), int(head[2])
    idx = np.arange(0, s.n, stride)
    # large systems: positions are regenerated from the seeded generator and
    # checked against this digest instead of being stored
    pos = s.pos if stride == 1 else np.zeros((0, 3))
    return dict(
        generator=np.array(gen), generator_args=np.array(json.dumps(args)), cellsize=np.float64(cellsize),
        spec=np.array(" ".join(spec)), rc=np.float64(rc), pos=pos, n=np.int64(s.n),
        pos_sha256=np.array(hashlib.sha256(np.ascontiguousarray(s.pos).tobytes()).hexdigest()),
        force_index=idx.astype(np.int32),
        forces=frc[idx], e_coulomb=np.float64(ec), e_lj=np.float64(elj), alpha=np.float64(alpha),
        cells=cells, cell_dims=np.array(dims, np.int32), npairs=np.int64(npairs),
        pair_hash=np.uint64(phash))


def main():
    if not os.path.exists(PM_REF):
        sys.exit(f"{PM_REF} missing: build it with make -f oracle/ref/Makefile")
    names = sys.argv[1:] or list(CASES)
    for name in names:
        d = run_case(name, *CASES[name])
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **d)
        print(f"{name}: N={int(d['n'])} Ecoul={float(d['e_coulomb']):.10g} Elj={float(d['e_lj']):.10g} "
              f"alpha={float(d['alpha'])} pairs={int(d['npairs'])}")


if __name__ == "__main__":
    main()
