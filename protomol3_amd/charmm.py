"""Minimal CHARMM input readers (PSF, PDB, PAR) that build a MolSystem the way
ProtoMol's topology builder does for the nonbonded path.

Follows protomol/src/protomol/io/PSFReader.cpp (atoms, bonds, angles,
dihedrals, impropers), io/PDBReader.cpp (ATOM/HETATM coordinates),
io/PARReader.cpp:416-450 (NONBONDED: atom, ignored, epsilon, Rmin/2
[, ignored, epsilon14, Rmin/2_14]), topology/BuildTopology.cpp:92 (scaled
charges), :663-673 (CHARMM combining), :900-1095 (molecules: connected
components over bonds, angles, dihedrals and impropers, numbered by lowest
atom) and :1102-1158 (exclusions by the .conf `exclude` keyword: 1-2 pairs
from bonds, 1-3 from angles, 1-4 from dihedrals, FULL except 1-4 pairs under
scaled1-4, which are MODIFIED).  Dihedral and improper energies are not part
of this engine (the device MD step evaluates bonds and angles only); their
atom lists are kept for exclusions and molecules.
"""
from __future__ import annotations

import math

import numpy as np

from .systems import FULL, MODIFIED, SQRTCOULOMBCONSTANT, MolSystem


def read_psf(path: str) -> dict:
    with open(path) as f:
        lines = f.read().splitlines()
    out = {"atoms": [], "bonds": [], "angles": [], "dihedrals": np.zeros((0, 4), np.int32),
           "impropers": np.zeros((0, 4), np.int32)}
    i = 0
    while i < len(lines):
        ln = lines[i]
        if "!NATOM" in ln:
            n = int(ln.split()[0])
            for k in range(n):
                t = lines[i + 1 + k].split()
                out["atoms"].append(dict(name=t[4], type=t[5], charge=float(t[6]), mass=float(t[7])))
            i += n
        elif any(k in ln for k in ("!NBOND", "!NTHETA", "!NPHI", "!NIMPHI")):
            n = int(ln.split()[0])
            key, per = next((k, p) for t, k, p in (("!NBOND", "bonds", 2), ("!NTHETA", "angles", 3),
                                                   ("!NPHI", "dihedrals", 4), ("!NIMPHI", "impropers", 4))
                             if t in ln)
            vals = []
            j = i + 1
            while len(vals) < n * per:
                vals += [int(v) - 1 for v in lines[j].split()]
                j += 1
            out[key] = np.array(vals, dtype=np.int32).reshape(n, per)
            i = j - 1
        i += 1
    return out


def read_pdb(path: str) -> np.ndarray:
    xyz = []
    with open(path) as f:
        for ln in f:
            if ln.startswith(("ATOM", "HETATM")):
                xyz.append((float(ln[30:38]), float(ln[38:46]), float(ln[46:54])))
    return np.array(xyz, dtype=np.float64)


def read_par(path: str) -> dict:
    sec = None
    par = {"bonds": {}, "angles": {}, "nonbonded": {}}
    with open(path) as f:
        for raw in f:
            ln = raw.split("!")[0].strip()
            if not ln:
                continue
            w = ln.split()
            head = w[0].upper()
            if head.startswith("BOND"):
                sec = "bonds"; continue
            if head.startswith("ANGL") or head.startswith("THET"):
                sec = "angles"; continue
            if head.startswith("DIHE") or head.startswith("PHI") or head.startswith("IMPR") \
                    or head.startswith("IMPH") or head.startswith("CMAP") or head.startswith("NBFIX") \
                    or head.startswith("HBOND") or head.startswith("END"):
                sec = None; continue
            if head.startswith("NONB") or head.startswith("NBON"):
                sec = "nonbonded"; continue
            if sec == "nonbonded" and head.startswith(("CUTNB", "WMIN")):
                continue
            try:
                if sec == "bonds" and len(w) >= 4:
                    k, r0 = float(w[2]), float(w[3])
                    par["bonds"][(w[0], w[1])] = (k, r0)
                    par["bonds"][(w[1], w[0])] = (k, r0)
                elif sec == "angles" and len(w) >= 5:
                    k, th = float(w[3]), float(w[4])
                    par["angles"][(w[0], w[1], w[2])] = (k, th)
                    par["angles"][(w[2], w[1], w[0])] = (k, th)
                elif sec == "nonbonded" and len(w) >= 4:
                    eps, rm = float(w[2]), float(w[3])
                    e14, r14 = (float(w[5]), float(w[6])) if len(w) >= 7 else (eps, rm)
                    par["nonbonded"][w[0]] = (eps, rm, e14, r14)
            except ValueError:
                continue
    return par


EXCLUDE = ("none", "1-2", "1-3", "1-4", "scaled1-4")


def build_system(psf: str, pdb: str, par: str, box, origin=(0.0, 0.0, 0.0), exclude: str = "scaled1-4",
                 csf: float = 1.0, name: str = "charmm") -> MolSystem:
    """The MolSystem ProtoMol builds from a PSF/PDB/PAR set for the .conf
    keywords `exclude` and `coulombScalingFactor` (csf)."""
    if exclude not in EXCLUDE:
        raise ValueError(f"exclude must be one of {EXCLUDE}")
    p = read_psf(psf)
    pos = read_pdb(pdb)
    prm = read_par(par)
    atoms = p["atoms"]
    n = len(atoms)
    if pos.shape[0] != n:
        raise ValueError("PDB and PSF atom counts differ")
    # atom types in order of first appearance (BuildTopology.cpp:80-84)
    tnames: list[str] = []
    tid = {}
    for a in atoms:
        if a["type"] not in tid:
            tid[a["type"]] = len(tnames)
            tnames.append(a["type"])
    T = len(tnames)
    lj = np.zeros((T, T, 4))
    for i, ti in enumerate(tnames):
        ei, ri, e14i, r14i = prm["nonbonded"][ti]
        for j, tj in enumerate(tnames):
            ej, rj, e14j, r14j = prm["nonbonded"][tj]
            r, e = ri + rj, math.sqrt(ei * ej)
            r14, e14 = r14i + r14j, math.sqrt(e14i * e14j)
            lj[i, j] = (r ** 12 * e, 2 * r ** 6 * e, r14 ** 12 * e14, 2 * r14 ** 6 * e14)
    bonds = np.asarray(p["bonds"], dtype=np.int32).reshape(-1, 2)
    angles = np.asarray(p["angles"], dtype=np.int32).reshape(-1, 3)
    bpar = np.array([prm["bonds"][(atoms[a]["type"], atoms[b]["type"])] for a, b in bonds]).reshape(-1, 2)
    apar = np.array([(k, math.radians(t)) for k, t in
                     (prm["angles"][(atoms[a]["type"], atoms[b]["type"], atoms[c]["type"])]
                      for a, b, c in angles)]).reshape(-1, 2)
    dihedrals = np.asarray(p["dihedrals"], dtype=np.int32).reshape(-1, 4)
    impropers = np.asarray(p["impropers"], dtype=np.int32).reshape(-1, 4)
    # molecules: connected components over bonds, angles, dihedrals and
    # impropers (BuildTopology.cpp:900-1060), numbered by lowest atom
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in bonds:
        parent[find(a)] = find(b)
    for a, b, c in angles:
        parent[find(a)] = find(b)
        parent[find(c)] = find(b)
    for quad in (*dihedrals, *impropers):
        for x in quad[1:]:
            parent[find(x)] = find(quad[0])
    roots = {}
    mol = np.array([roots.setdefault(find(i), len(roots)) for i in range(n)], dtype=np.int32)
    # exclusion entries in the reference's insertion order: bonds, angles,
    # dihedrals (BuildTopology.cpp:1118-1156).  A pair may come twice with
    # different classes (a 1-3 pair that is also 1-4 around a five-membered
    # ring); the engine keeps both rules of ExclusionTable::add for it: the
    # last class for membership, the first for the PME correction table.
    table = []
    if exclude != "none":
        table += [(min(a, b), max(a, b), FULL) for a, b in bonds]
        if exclude != "1-2":
            table += [(min(a, c), max(a, c), FULL) for a, _, c in angles]
            if exclude != "1-3":
                cls = FULL if exclude == "1-4" else MODIFIED
                table += [(min(a, d), max(a, d), cls) for a, _, _, d in dihedrals]
    ex = np.array(table, dtype=np.int32).reshape(-1, 3)
    return MolSystem(
        name=name, pos=pos, charge=np.array([a["charge"] for a in atoms]) * SQRTCOULOMBCONSTANT,
        atype=np.array([tid[a["type"]] for a in atoms], dtype=np.int32), molecule=mol,
        mass=np.array([a["mass"] for a in atoms]), lj=lj, box=np.asarray(box, dtype=np.float64),
        origin=np.asarray(origin, dtype=np.float64), bonds=bonds, bond_par=bpar, angles=angles,
        angle_par=apar, exclusions=ex, csf=csf, dihedrals=dihedrals, impropers=impropers)
