"""CPU oracle for the device-resident MD step -- TEST INFRASTRUCTURE ONLY.

FP64 restatement of the reference's leapfrog integrator and the harmonic
bonded terms of the systems the engine runs:
  * LeapfrogIntegrator::run (protomol/src/protomol/integrator/leapfrog/
    LeapfrogIntegrator.cpp:89-104): half kick + drift, forces,
    (kick + drift, forces) x (n-1), half kick; h = dt * INV_TIMEFACTOR;
  * BondSystemForce::calcBond (force/bonded/BondSystemForce.h:90-121),
    E = k (r - r0)^2;
  * AngleSystemForce::calcAngle (force/bonded/AngleSystemForce.h:90-160)
    without Urey-Bradley, E = k (theta - theta0)^2.
Nonbonded forces come from the C oracle (oracle.evaluate).
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

INV_TIMEFACTOR = 0.02045482828087295      # protomol/base/PMConstants.cpp:54
PDBVELSCALINGFACTOR = 20.45482706          # protomol/base/PMConstants.cpp:71


def _min_image(d: np.ndarray, box: np.ndarray) -> np.ndarray:
    return d - box * np.rint(d / box)


def bonded(s, pos: np.ndarray):
    """(forces (N,3), E_bond, E_angle)."""
    f = np.zeros_like(pos)
    box = np.asarray(s.box, float)
    eb = ea = 0.0
    if len(s.bonds):
        a1, a2 = s.bonds[:, 0], s.bonds[:, 1]
        k, r0 = s.bond_par[:, 0], s.bond_par[:, 1]
        r12 = _min_image(pos[a1] - pos[a2], box)
        r = np.linalg.norm(r12, axis=1)
        dpot = 2.0 * k * (r - r0)
        f1 = r12 * (-dpot / r)[:, None]
        np.add.at(f, a1, f1)
        np.add.at(f, a2, -f1)
        eb = float(np.sum(k * (r - r0) ** 2))
    if len(s.angles):
        a1, a2, a3 = s.angles[:, 0], s.angles[:, 1], s.angles[:, 2]
        k, t0 = s.angle_par[:, 0], s.angle_par[:, 1]
        r12 = _min_image(pos[a1] - pos[a2], box)
        r32 = _min_image(pos[a3] - pos[a2], box)
        d12 = np.linalg.norm(r12, axis=1)
        d32 = np.linalg.norm(r32, axis=1)
        th = np.arctan2(np.linalg.norm(np.cross(r12, r32), axis=1), np.einsum("ij,ij->i", r12, r32))
        st, ct = np.sin(th), np.cos(th)
        dpot = 2.0 * k * (th - t0)
        dr12, dr32 = r12 / d12[:, None], r32 / d32[:, None]
        dt1 = (dr12 * ct[:, None] - dr32) / (st * d12)[:, None]
        dt3 = (dr32 * ct[:, None] - dr12) / (st * d32)[:, None]
        f1 = -dt1 * dpot[:, None]
        f3 = -dt3 * dpot[:, None]
        np.add.at(f, a1, f1)
        np.add.at(f, a2, -f1 - f3)
        np.add.at(f, a3, f3)
        ea = float(np.sum(k * (th - t0) ** 2))
    return f, eb, ea


def total_forces(s, spec, pos, nthreads=8):
    f_nb, ec, elj = O.evaluate(s, spec, pos=pos, nthreads=nthreads)
    f_b, eb, ea = bonded(s, pos)
    return f_nb + f_b, dict(coulomb=ec, lj=elj, bond=eb, angle=ea)


def leapfrog_run(s, spec, x, v, dt_fs: float, nsteps: int, f=None, nthreads=8):
    """LeapfrogIntegrator::run(nsteps) from state (x, v) with forces f at x
    (computed if None).  Returns (x, v, f, energies)."""
    h = dt_fs * INV_TIMEFACTOR
    minv = (1.0 / np.asarray(s.mass, float))[:, None]
    x = np.array(x, dtype=float)
    v = np.array(v, dtype=float)
    if f is None:
        f, e = total_forces(s, spec, x, nthreads)
    for step in range(nsteps):
        v += f * (0.5 if step == 0 else 1.0) * h * minv
        x += v * h
        f, e = total_forces(s, spec, x, nthreads)
    v += f * 0.5 * h * minv
    e = dict(e)
    e["kinetic"] = float(0.5 * np.sum(np.sum(v * v, axis=1) / minv[:, 0]))
    return x, v, f, e
