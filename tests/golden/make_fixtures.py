#!/usr/bin/env python3
"""Writes the golden fixtures in tests/golden/*.npz from the reference
ProtoMol path, compiled from the reference sources by oracle/ref/Makefile
(oracle/_ref/pm_ref).  Run from the repository root in the container that
holds /root/reference:

    make -f oracle/ref/Makefile -j8 && python tests/golden/make_fixtures.py

The topology fixtures (topo_*.npz) hold what the reference's topology builder
(compiled into oracle/_ref/pm_ref_topo) makes of the reference's own test
molecules in tests/golden/charmm/ for each `exclude` keyword: LJ type, scaled
charge and molecule of every atom, the LJ table, and the exclusion table with
the class check() returns for each entry.

Each force fixture holds the inputs (positions; the generator and seed that made
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
    if gen == "charmm_file":
        return charmm_file_system(**args)
    return getattr(systems, gen)(**args)


CHARMM = os.path.join(OUT, "charmm")
PAR27 = os.path.join(W423, "par_all27_prot_lipid.inp")


def charmm_file_system(stem: str, box: float, exclude: str, csf: float):
    """A reference test molecule (tests/golden/charmm/<stem>.psf/.pdb with the
    CHARMM27 parameters) in a cubic periodic box whose origin puts the
    molecule's centre at the box centre."""
    psf, pdb = os.path.join(CHARMM, stem + ".psf"), os.path.join(CHARMM, stem + ".pdb")
    pos = charmm.read_pdb(pdb)
    centre = 0.5 * (pos.min(0) + pos.max(0))
    return charmm.build_system(psf, pdb, PAR27, box=(box, box, box), origin=tuple(centre - 0.5 * box),
                               exclude=exclude, csf=csf, name=stem)


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
    # reference test molecules with dihedrals and impropers (CHARMM27):
    # 1-4 pairs from the topology builder, rings where a pair is both 1-3 and
    # 1-4, a protein with net charge
    "wwdomain_scaled14_pme": ("charmm_file", dict(stem="ww_domain", box=48.0, exclude="scaled1-4", csf=1.0), 5.0,
                              ["lj_c2", "8", "10", "pme", "48", "48", "48", "4", "10", "1e-6", "-1", "111"], 10.0, 1),
    "droplet_alan_scaled14_coulc2": ("charmm_file",
                                     dict(stem="large_water_droplet_alan", box=40.0, exclude="scaled1-4",
                                          csf=0.8333), 5.0,
                                     ["lj_c2", "8", "10", "coul_c2", "8", "10"], 10.0, 1),
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


PM_REF_TOPO = os.path.join(ROOT, "oracle", "_ref", "pm_ref_topo")

# topology fixtures: (molecule stem, exclude keyword) -> the reference
# topology builder's output for that PSF with the CHARMM27 parameters
TOPO_CASES = {
    "ww_domain_scaled14": ("ww_domain", "scaled1-4"),
    "ww_domain_one4": ("ww_domain", "1-4"),
    "ww_domain_one3": ("ww_domain", "1-3"),
    "droplet_alan_scaled14": ("large_water_droplet_alan", "scaled1-4"),
    "droplet_alan_one2": ("large_water_droplet_alan", "1-2"),
}


def run_topo_case(stem: str, exclude: str) -> dict:
    with tempfile.TemporaryDirectory() as td:
        of = os.path.join(td, "topo")
        subprocess.run([PM_REF_TOPO, os.path.join(CHARMM, stem + ".psf"), PAR27, exclude, of], check=True,
                       capture_output=True)
        with open(of) as f:
            n, t, ne = (int(v) for v in f.readline().split())
            atoms = np.array([f.readline().split() for _ in range(n)], dtype=object)
            lj = np.array([[float(v) for v in f.readline().split()] for _ in range(t * t)]).reshape(t, t, 4)
            table = np.array([[int(v) for v in f.readline().split()] for _ in range(ne)],
                             dtype=np.int32).reshape(-1, 4)
            counts = np.array([int(v) for v in f.readline().split()], dtype=np.int32)
    return dict(stem=np.array(stem), exclude=np.array(exclude), type=atoms[:, 0].astype(np.int32),
                scaled_charge=atoms[:, 1].astype(np.float64), molecule=atoms[:, 2].astype(np.int32), lj=lj,
                table=table, counts=counts)


# compare fixtures: the reference's `compare` report (CompareForce.cpp:82-130)
# of force spec B against force spec A on one system with O/H masses
COMPARE_CASES = {
    "compare_water1536_ljc2": ("water_box", dict(seed=3), 5.0, ["lj_c2", "8", "10"], ["lj_c2", "7", "10"]),
    "compare_water1536_pme": ("water_box", dict(seed=3), 5.0,
                              ["pme", "32", "32", "32", "4", "10", "1e-6", "0.32", "111"],
                              ["pme", "24", "24", "24", "4", "10", "1e-6", "0.32", "111"]),
}


def run_compare_case(gen: str, args: dict, cellsize: float, spec_a: list, spec_b: list) -> dict:
    s = make_system(gen, args)
    mass = np.where(np.arange(s.n) % 3 == 0, 15.9994, 1.008)
    out = {}
    with tempfile.TemporaryDirectory() as td:
        sf, of, mf = (os.path.join(td, x) for x in ("sys", "out", "mass"))
        sysfile.write_sysfile(sf, s)
        np.savetxt(mf, mass, fmt="%.17g")
        for tag, spec in (("a", spec_a), ("b", spec_b)):
            subprocess.run([PM_REF, sf, of, "cellsize", repr(cellsize), *spec], capture_output=True, text=True,
                           check=True)
            ec, elj, frc = sysfile.read_forces(of)
            out[f"forces_{tag}"] = frc
            out[f"pe_{tag}"] = np.float64(ec + elj)
        r = subprocess.run([PM_REF, sf, of, "cellsize", repr(cellsize), *spec_a, *spec_b], capture_output=True,
                           text=True, check=True, env=dict(os.environ, PM_REF_COMPARE="1", PM_REF_MASSES=mf))
    lines = [ln for ln in (r.stdout + r.stderr).splitlines() if ln.startswith("Comparing")]
    out.update(generator=np.array(gen), generator_args=np.array(json.dumps(args)), mass=mass,
               pos=s.pos, report_lines=np.array(lines))
    return out


def main():
    if not os.path.exists(PM_REF):
        sys.exit(f"{PM_REF} missing: build it with make -f oracle/ref/Makefile")
    names = sys.argv[1:] or list(CASES) + [f"topo_{k}" for k in TOPO_CASES] + list(COMPARE_CASES)
    for name in names:
        if name in COMPARE_CASES:
            d = run_compare_case(*COMPARE_CASES[name])
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **d)
            print(f"{name}: " + " | ".join(str(x) for x in d["report_lines"]))
            continue
        if name.startswith("topo_"):
            d = run_topo_case(*TOPO_CASES[name[5:]])
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **d)
            print(f"{name}: N={len(d['type'])} types={d['lj'].shape[0]} exclusions={len(d['table'])} "
                  f"bonded={d['counts'].tolist()}")
            continue
        d = run_case(name, *CASES[name])
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **d)
        print(f"{name}: N={int(d['n'])} Ecoul={float(d['e_coulomb']):.10g} Elj={float(d['e_lj']):.10g} "
              f"alpha={float(d['alpha'])} pairs={int(d['npairs'])}")


if __name__ == "__main__":
    main()
