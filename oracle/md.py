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


# ---------------------------------------------------------------------------
# Langevin leapfrog (LangevinLeapfrogIntegrator.cpp:28, 44-71) with the noise
# stream of the engine: Philox4x32-10 (Salmon et al., SC 2011) at counter
# (atom, half-kick index lo, hi, 'LANG') and key (seed lo, hi), uniforms
# (x + 1/2) 2^-32, Box-Muller on (u0, u1) and (u2, u3).

BOLTZMANN = 0.001987191                    # protomol/base/PMConstants.cpp:70
_M0, _M1, _W0, _W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    """ctr: (4, m) uint64 arrays of 32-bit words; key: two ints.  Returns the
    four output words, each an (m,) uint64 array."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & _MASK for c in ctr)
    k0, k1 = int(key[0]) & _MASK, int(key[1]) & _MASK
    for _ in range(10):
        p0 = np.uint64(_M0) * c0
        p1 = np.uint64(_M1) * c2
        n0 = ((p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)) & np.uint64(_MASK)
        n2 = ((p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)) & np.uint64(_MASK)
        c1 = p1 & np.uint64(_MASK)
        c3 = p0 & np.uint64(_MASK)
        c0, c2 = n0, n2
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def langevin_gauss3(seed: int, half: int, n: int) -> np.ndarray:
    """(n, 3) standard normal numbers of half kick `half` for atoms 0..n-1."""
    atoms = np.arange(n, dtype=np.uint64)
    m = np.full(n, 1, dtype=np.uint64)
    x = philox4x32_10((atoms, m * np.uint64(half & _MASK), m * np.uint64((half >> 32) & _MASK),
                       m * np.uint64(0x4C414E47)), (seed & _MASK, (seed >> 32) & _MASK))
    s = 2.3283064365386963e-10
    u = [(xi.astype(np.float64) + 0.5) * s for xi in x]
    r0 = np.sqrt(-2.0 * np.log(u[0]))
    r1 = np.sqrt(-2.0 * np.log(u[2]))
    tp = 6.283185307179586
    return np.stack([r0 * np.cos(tp * u[1]), r0 * np.sin(tp * u[1]), r1 * np.cos(tp * u[3])], axis=1)


def langevin_constants(dt_fs: float, temperature: float, gamma_ps: float):
    """(h, vdt, fdt, ndt, fc) of LangevinLeapfrogIntegrator.cpp:28, 46-51
    (gamma = 0: the limits fdt = h/2, ndt = 0)."""
    h = dt_fs * INV_TIMEFACTOR
    g = gamma_ps / (1000.0 * INV_TIMEFACTOR)
    vdt = np.exp(-0.5 * g * h)
    fdt = (1.0 - np.exp(-0.5 * g * h)) / g if g > 0 else 0.5 * h
    ndt = np.sqrt((1.0 - np.exp(-g * h)) / (2.0 * g)) if g > 0 else 0.0
    fc = 2.0 * BOLTZMANN * temperature * g
    return h, vdt, fdt, ndt, fc


def langevin_run(s, spec, x, v, dt_fs, nsteps, temperature, gamma_ps, seed, step0=0, f=None, nthreads=8,
                 forces=None):
    """StandardIntegrator::run(nsteps) with the Langevin half kicks
    v = v vdt + F fdt / m + g sqrt(fc / m) ndt; half kick 2t opens step t and
    2t + 1 closes it (t counted from step0).  forces(x) -> (f, energies)
    replaces the oracle force field when given."""
    h, vdt, fdt, ndt, fc = langevin_constants(dt_fs, temperature, gamma_ps)
    minv = (1.0 / np.asarray(s.mass, float))[:, None]
    sfm = np.sqrt(fc * minv) * ndt
    calc = forces or (lambda xx: total_forces(s, spec, xx, nthreads))
    x = np.array(x, dtype=float)
    v = np.array(v, dtype=float)
    n = len(x)
    if f is None:
        f, e = calc(x)
    e = {}
    for t in range(step0, step0 + nsteps):
        v = v * vdt + f * fdt * minv + langevin_gauss3(seed, 2 * t, n) * sfm
        x += v * h
        f, e = calc(x)
        v = v * vdt + f * fdt * minv + langevin_gauss3(seed, 2 * t + 1, n) * sfm
    e = dict(e)
    e["kinetic"] = float(0.5 * np.sum(np.sum(v * v, axis=1) / minv[:, 0]))
    return x, v, f, e
