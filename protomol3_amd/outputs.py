"""ProtoMol output files written from engine results, for line-by-line
comparison with the reference's own expected outputs.

* XYZ trajectory files (io/XYZTrajectoryWriter.cpp:64-160), as written for
  ``finXYZForceFile`` / ``XYZForceFile`` (output/OutputXYZTrajectoryForce.cpp:
  39-44) and position/velocity trajectories: a first line holding the frame
  count left-aligned in 19 characters, then per frame the atom count and one
  line per atom, ``name<TAB>`` and three fields of width 24 at 15 significant
  digits.
* The all-energies file (output/OutputEnergies.cpp:73-106): time in fs with
  two decimals in width 14, then E_potential, E_kinetic, E_total,
  temperature, E_bond, E_angle, E_dihedral, E_improper, E_VdW, E_coulomb,
  E_other and the volume at 8 significant digits in width 14, and E_shadow
  at 16 digits in width 20.  The reference sets both ``fixed`` and
  ``scientific`` there, which newer C++ libraries print as hexadecimal
  floats; the general format written here is the one its expected files hold.
"""
from __future__ import annotations

import os

import numpy as np

BOLTZMANN = 0.001987191  # protomol/base/PMConstants.cpp:70


def _g(v: float, prec: int) -> str:
    return f"{v:.{prec}g}"


def write_xyz_frame(path: str, coords: np.ndarray, names) -> None:
    """Appends one frame to an XYZ trajectory file, creating it or
    rewriting its frame-count line as XYZTrajectoryWriter::reopen does."""
    coords = np.asarray(coords, float).reshape(-1, 3)
    frames = 0
    if os.path.exists(path):
        with open(path) as f:
            first = f.readline()
        try:
            frames = int(first.split()[0])
        except (ValueError, IndexError):
            frames = 0
    head = (str(frames + 1) + " " * 23)[:19] + "\n"
    if frames > 0:
        with open(path, "r+") as f:
            f.seek(0)
            f.write(head)
    else:
        with open(path, "w") as f:
            f.write(head)
    lines = [f"{len(coords)}\n"]
    for nm, (x, y, z) in zip(names, coords):
        lines.append(f"{nm}\t{_g(x, 15):>24}{_g(y, 15):>24}{_g(z, 15):>24}\n")
    with open(path, "a") as f:
        f.writelines(lines)


def temperature(kinetic: float, dof: int) -> float:
    """ProtoMol::temperature (topology/TopologyUtilities.cpp): 2 E_kin /
    (dof k_B)."""
    return 2.0 * kinetic / (dof * BOLTZMANN)


def energies_line(time_fs: float, e: dict, dof: int, volume: float, shadow: float = 0.0) -> str:
    """One line of the all-energies file for energies e (keys coulomb, lj,
    bond, angle, kinetic; dihedral, improper, other default to 0)."""
    pot = sum(e.get(k, 0.0) for k in ("coulomb", "lj", "bond", "angle", "dihedral", "improper", "other"))
    kin = e.get("kinetic", 0.0)
    vals = [pot, kin, pot + kin, temperature(kin, dof), e.get("bond", 0.0), e.get("angle", 0.0),
            e.get("dihedral", 0.0), e.get("improper", 0.0), e.get("lj", 0.0), e.get("coulomb", 0.0),
            e.get("other", 0.0), volume]
    return (f"{time_fs:14.2f} " + " ".join(f"{_g(v, 8):>14}" for v in vals) + " " + f"{_g(shadow, 16):>20}")


# Potential-energy slots of ScalarStructure::potentialEnergy
# (protomol/type/ScalarStructure.h:140-147).
POTENTIAL_KEYS = ("coulomb", "lj", "bond", "angle", "dihedral", "improper", "other")


def compare_forces(f_ref: np.ndarray, pe_ref: float, f: np.ndarray, pe: float, mass: np.ndarray) -> dict:
    """The A/B error report of ProtoMol's `compare` force keyword for one
    evaluation (CompareForce::postprocess, protomol/force/CompareForce.cpp:
    89-130): `f`, `pe` are the forces and potential energy of the force under
    test, `f_ref`, `pe_ref` those of the force it is compared with, `mass`
    the atoms' masses.  With e_i = |f_i - f_ref,i| / sqrt(m_i) and
    g_i = |f_i| / sqrt(m_i):
      absF2 = max_i |f_i - f_ref,i|,
      rFavg = sum_i e_i / sum_i g_i,
      rFmax = N max_i e_i / sum_i g_i,
      rPE   = |(pe_ref - pe) / pe|."""
    f = np.asarray(f, float).reshape(-1, 3)
    f_ref = np.asarray(f_ref, float).reshape(-1, 3)
    rm = 1.0 / np.asarray(mass, float)
    diff = np.sum((f - f_ref) ** 2, axis=1)
    err = np.sqrt(diff * rm)
    total = float(np.sum(np.sqrt(np.sum(f * f, axis=1) * rm)))
    n = len(f)
    return {"absF2": float(np.sqrt(diff.max())) if n else 0.0,
            "rFavg": float(err.sum() / total) if total != 0.0 else 0.0,
            "rFmax": float(n * err.max() / total) if total != 0.0 else 0.0,
            "rPE": abs((pe_ref - pe) / pe) if pe != 0.0 else 0.0}


def compare_summary(reports: list) -> dict:
    """The summary CompareForce prints when it is destroyed (CompareForce.cpp:
    33-71): the mean of every figure over the evaluations and its sample
    standard deviation (0 for a single evaluation), and their count."""
    n = len(reports)
    out = {"n": n}
    for k in ("absF2", "rFavg", "rFmax", "rPE"):
        v = np.array([r[k] for r in reports], float)
        mean = float(np.sum(v / n)) if n else 0.0
        sd = float(np.sqrt(np.sum((v - mean) ** 2) / (n - 1.0))) if n > 1 else 0.0
        out[k] = (mean, sd)
    return out


def compare_line(idx: int, r: dict) -> str:
    """One "Comparing" line as the reference prints it (15 significant
    digits, StringUtilities.cpp:114-119)."""
    return (f"Comparing {idx} : absF2 {r['absF2']:.15g}, rFavg {r['rFavg']:.15g}, rFmax {r['rFmax']:.15g}, "
            f"rPE {r['rPE']:.15g}")
