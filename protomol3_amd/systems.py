"""Seeded synthetic molecular systems in ProtoMol units (Angstrom, kcal/mol, amu).

Each generator returns a :class:`MolSystem` holding exactly what ProtoMol's
topology carries for the nonbonded path: scaled charges
(q * SQRTCOULOMBCONSTANT, protomol/topology/BuildTopology.cpp:92), per-atom LJ
type with the CHARMM-combined table {A, B, A14, B14}
(protomol/topology/BuildTopology.cpp:663-673), molecule ids (connected
components of the bond graph, BuildTopology.cpp:900-1095) and the exclusion
list built from bonds and angles (BuildTopology.cpp:1102-1158, FULL for 1-2 and
1-3 pairs).  Bonds and angles use ProtoMol's harmonic convention
E = k (x - x0)^2 (force/bonded/BondSystemForce.h:117-118,
AngleSystemForce.h:147).

The workloads:
  * water_box      -- 512 flexible 3-site waters in a 24.8 A cube (1,536 atoms)
  * polymer_water  -- 500-bead polymer + 8,500 waters, 64 A cube (26,000 atoms)
  * big_water      -- 333,000 waters in a 216 A cube (999,000 atoms)
  * replica_set    -- many water+polymer boxes of uneven size
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

SQRTCOULOMBCONSTANT = 18.2226123264      # protomol/base/PMConstants.cpp:67
BOLTZMANN = 0.001987191                  # protomol/base/PMConstants.cpp:70
INV_TIMEFACTOR = 0.02045482828087295     # protomol/base/PMConstants.cpp:54

FULL = 1
MODIFIED = 2

# Water model (flexible 3-site): O-H 0.9572 A, H-O-H 104.52 deg, k_b 450, k_theta 55
Q_O, Q_H = -0.834, 0.417
EPS_O, SIGMA_O = 0.1521, 3.1507
R_OH, THETA_HOH = 0.9572, math.radians(104.52)
KB_OH, KTH_HOH = 450.0, 55.0
M_O, M_H = 15.9994, 1.008
# Bead polymer: bond 3.8 A, LJ eps 0.2 / sigma 4.0, charges +-0.5 alternating
BEAD_BOND, BEAD_EPS, BEAD_SIGMA, BEAD_Q = 3.8, 0.2, 4.0, 0.5
BEAD_KB, BEAD_MASS = 100.0, 100.0
# 1-4 LJ of the bead when the chain carries scaled 1-4 pairs (CHARMM's
# separate eps14 / Rmin14/2 columns of a PAR file)
BEAD_EPS14, BEAD_SIGMA14 = 0.1, 3.5


def rmin_half(sigma: float) -> float:
    return sigma * 2.0 ** (1.0 / 6.0) / 2.0


def charmm_table(eps: list[float], rmin2: list[float], eps14: list[float] | None = None,
                 rmin2_14: list[float] | None = None) -> np.ndarray:
    """T x T x 4 table {A, B, A14, B14} with CHARMM combining
    (protomol/topology/BuildTopology.cpp:663-673).  The 1-4 columns combine
    eps14 / rmin2_14 the same way; without them they equal the normal ones."""
    t = len(eps)
    eps14 = eps if eps14 is None else eps14
    rmin2_14 = rmin2 if rmin2_14 is None else rmin2_14
    tab = np.zeros((t, t, 4))
    for i in range(t):
        for j in range(t):
            r = rmin2[i] + rmin2[j]
            e = math.sqrt(eps[i] * eps[j])
            r14 = rmin2_14[i] + rmin2_14[j]
            e14 = math.sqrt(eps14[i] * eps14[j])
            tab[i, j] = (r ** 12 * e, 2.0 * r ** 6 * e, r14 ** 12 * e14, 2.0 * r14 ** 6 * e14)
    return tab


@dataclasses.dataclass
class MolSystem:
    name: str
    pos: np.ndarray            # (N, 3) f64
    charge: np.ndarray         # (N,) scaled charge
    atype: np.ndarray          # (N,) int32 LJ type
    molecule: np.ndarray       # (N,) int32
    mass: np.ndarray           # (N,) amu
    lj: np.ndarray             # (T, T, 4)
    box: np.ndarray            # (3,) orthogonal edge lengths
    origin: np.ndarray         # (3,)
    bonds: np.ndarray          # (NB, 2) int32
    bond_par: np.ndarray       # (NB, 2) k, r0
    angles: np.ndarray         # (NA, 3) int32
    angle_par: np.ndarray      # (NA, 2) k, theta0 (rad)
    exclusions: np.ndarray     # (NE, 3) int32 a1 < a2, class
    csf: float = 1.0
    vel: np.ndarray | None = None
    # dihedral and improper atom lists (CHARMM inputs): they define 1-4
    # exclusions and molecules; their energies are not evaluated here
    dihedrals: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.int32))
    impropers: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 4), np.int32))

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def with_velocities(self, temperature: float, seed: int) -> "MolSystem":
        """Maxwell-Boltzmann velocities in ProtoMol internal units, sqrt(kT/m)
        per component (protomol/topology/TopologyUtilities.cpp:21-49)."""
        rng = np.random.default_rng(seed)
        s = np.sqrt(temperature * BOLTZMANN / self.mass)[:, None]
        self.vel = s * rng.standard_normal((self.n, 3))
        return self


def _water_exclusions(first: int) -> list[tuple[int, int, int]]:
    o, h1, h2 = first, first + 1, first + 2
    return [(o, h1, FULL), (o, h2, FULL), (h1, h2, FULL)]


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from normalised quaternions."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def _water_atoms(o_pos: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """(W, 3, 3) O, H1, H2 positions for oxygens at o_pos, random orientation."""
    half = THETA_HOH / 2.0
    local = np.array([[0.0, 0.0, 0.0],
                      [R_OH * math.sin(half), 0.0, R_OH * math.cos(half)],
                      [-R_OH * math.sin(half), 0.0, R_OH * math.cos(half)]])
    rot = _random_rotations(rng, o_pos.shape[0])
    return o_pos[:, None, :] + np.einsum("wij,aj->wai", rot, local)


def _assemble(name, box, parts, lj, csf=1.0) -> MolSystem:
    """parts: list of dicts with pos, charge, atype, mass, bonds, bond_par,
    angles, angle_par, exclusions (atom indices local to the part) and
    molecule (local molecule ids)."""
    pos, chg, typ, mol, mass = [], [], [], [], []
    bonds, bpar, angles, apar, excl = [], [], [], [], []
    off = 0
    moff = 0
    for p in parts:
        n = len(p["charge"])
        pos.append(p["pos"]); chg.append(p["charge"]); typ.append(p["atype"]); mass.append(p["mass"])
        mol.append(np.asarray(p["molecule"]) + moff)
        moff += int(np.max(p["molecule"])) + 1 if n else 0
        if len(p["bonds"]):
            bonds.append(np.asarray(p["bonds"]) + off); bpar.append(p["bond_par"])
        if len(p["angles"]):
            angles.append(np.asarray(p["angles"]) + off); apar.append(p["angle_par"])
        if len(p["exclusions"]):
            e = np.asarray(p["exclusions"]).copy()
            e[:, :2] += off
            excl.append(e)
        off += n
    cat = lambda xs, shape, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(shape, dt)
    box = np.asarray(box, dtype=np.float64)
    return MolSystem(
        name=name,
        pos=np.concatenate(pos).astype(np.float64),
        charge=np.concatenate(chg).astype(np.float64) * SQRTCOULOMBCONSTANT,
        atype=np.concatenate(typ).astype(np.int32),
        molecule=np.concatenate(mol).astype(np.int32),
        mass=np.concatenate(mass).astype(np.float64),
        lj=lj, box=box, origin=np.zeros(3),
        bonds=cat(bonds, (0, 2), np.int32), bond_par=cat(bpar, (0, 2), np.float64),
        angles=cat(angles, (0, 3), np.int32), angle_par=cat(apar, (0, 2), np.float64),
        exclusions=cat(excl, (0, 3), np.int32), csf=csf)


def _water_part(o_pos: np.ndarray, rng: np.random.Generator, o_type=0, h_type=1) -> dict:
    w = o_pos.shape[0]
    xyz = _water_atoms(o_pos, rng).reshape(-1, 3)
    idx = np.arange(w) * 3
    bonds = np.stack([np.stack([idx, idx + 1], 1), np.stack([idx, idx + 2], 1)], 1).reshape(-1, 2)
    angles = np.stack([idx + 1, idx, idx + 2], 1)
    excl = np.array([e for k in range(w) for e in _water_exclusions(3 * k)], dtype=np.int32).reshape(-1, 3)
    return dict(pos=xyz,
                charge=np.tile([Q_O, Q_H, Q_H], w),
                atype=np.tile([o_type, h_type, h_type], w),
                mass=np.tile([M_O, M_H, M_H], w),
                molecule=np.repeat(np.arange(w), 3),
                bonds=bonds, bond_par=np.tile([KB_OH, R_OH], (bonds.shape[0], 1)),
                angles=angles, angle_par=np.tile([KTH_HOH, THETA_HOH], (w, 1)),
                exclusions=excl)


def _jittered_lattice(n_side: int, spacing: float, jitter: float, rng) -> np.ndarray:
    g = (np.arange(n_side) + 0.5) * spacing
    o = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return o + rng.uniform(-jitter, jitter, o.shape)


def water_box(n_side: int = 8, box: float = 24.8, seed: int = 0, jitter: float = 0.3) -> MolSystem:
    """n_side^3 flexible 3-site waters: oxygens on a seeded jittered lattice
    (+-jitter A per coordinate), random orientations.  Default: the 1,536-atom
    24.8 A water box."""
    rng = np.random.default_rng(seed)
    o = _jittered_lattice(n_side, box / n_side, jitter, rng)
    lj = charmm_table([EPS_O, 0.0], [rmin_half(SIGMA_O), 0.0])
    return _assemble(f"water{3 * n_side ** 3}", [box] * 3, [_water_part(o, rng)], lj)


def _polymer_walk(n_beads: int, box: float, rng, min_sep: float = 4.0) -> np.ndarray:
    """Seeded self-avoiding random walk with fixed bond length, in the box:
    a step is rejected when it comes within min_sep of any earlier bead other
    than its predecessor (minimum image).  Earlier beads are kept in a cell
    grid of edge >= min_sep, so each test looks at 27 cells."""
    pts = [rng.uniform(0.25 * box, 0.75 * box, 3)]
    nc = int(box // min_sep)
    grid = {} if nc >= 3 else None
    cell = lambda p: tuple(int(c) % nc for c in np.floor((p - box * np.floor(p / box)) / (box / nc)))

    def near(cand, last):
        if grid is None:
            prev = np.asarray(pts[:-1])
            if not len(prev):
                return False
        else:
            cx, cy, cz = cell(cand)
            idx = [i for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
                   for i in grid.get(((cx + dx) % nc, (cy + dy) % nc, (cz + dz) % nc), ()) if i != last]
            if not idx:
                return False
            prev = np.asarray([pts[i] for i in idx])
        d = prev - cand
        d -= box * np.rint(d / box)
        return np.min(np.einsum("ij,ij->i", d, d)) < min_sep * min_sep

    if grid is not None:
        grid.setdefault(cell(pts[0]), []).append(0)
    tries = streak = 0
    while len(pts) < n_beads:
        v = rng.standard_normal(3)
        cand = pts[-1] + BEAD_BOND * v / np.linalg.norm(v)
        tries += 1
        if tries > 2000000:
            raise RuntimeError("polymer walk failed")
        if near(cand, len(pts) - 1):
            streak += 1
            if streak > 20000 and len(pts) > 8:
                # trapped: step back 8 beads and continue from there
                for _ in range(8):
                    p = pts.pop()
                    if grid is not None:
                        grid[cell(p)].remove(len(pts))
                streak = 0
            continue
        streak = 0
        pts.append(cand)
        if grid is not None:
            grid.setdefault(cell(cand), []).append(len(pts) - 1)
    return np.asarray(pts)


def _clear_of(wat: np.ndarray, beads: np.ndarray, box: float, overlap: float) -> np.ndarray:
    """Mask of water sites (S, 3 atoms, xyz) with no atom within `overlap` of
    any bead (minimum image); water atoms are binned in cells of edge >=
    overlap so each bead tests its 27 neighbouring cells."""
    S = wat.shape[0]
    keep = np.ones(S, bool)
    nc = int(box // overlap)
    if nc < 3:
        for b in beads:
            d = wat - b
            d -= box * np.rint(d / box)
            keep &= ~(np.einsum("sak,sak->sa", d, d).min(1) < overlap * overlap)
        return keep
    flat = wat.reshape(-1, 3)
    cid = np.floor((flat - box * np.floor(flat / box)) / (box / nc)).astype(np.int64) % nc
    key = (cid[:, 0] * nc + cid[:, 1]) * nc + cid[:, 2]
    order = np.argsort(key, kind="stable")
    start = np.searchsorted(key[order], np.arange(nc ** 3 + 1))
    for b in beads:
        c = np.floor((b - box * np.floor(b / box)) / (box / nc)).astype(np.int64) % nc
        cells = [(((c[0] + dx) % nc) * nc + (c[1] + dy) % nc) * nc + (c[2] + dz) % nc
                 for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
        atoms = np.concatenate([order[start[q]:start[q + 1]] for q in cells])
        if atoms.size == 0:
            continue
        d = flat[atoms] - b
        d -= box * np.rint(d / box)
        hit = np.einsum("ij,ij->i", d, d) < overlap * overlap
        keep[atoms[hit] // 3] = False
    return keep


def polymer_water(n_total: int = 26000, box: float = 64.0, n_beads: int = 500, seed: int = 1,
                  overlap: float = 2.4, jitter: float = 0.3, oversample: float = 1.25,
                  scaled14: float | None = None) -> MolSystem:
    """Solvated bead polymer: a seeded random-walk chain of n_beads (bond 3.8 A,
    LJ eps 0.2 / sigma 4.0, charges +-0.5 alternating) plus 3-site waters on a
    jittered lattice; waters with any atom within `overlap` A of a bead are
    removed and the remaining lattice sites are drawn in seeded order until the
    system holds exactly n_total atoms.  The lattice holds `oversample`
    times the waters needed.  An odd n_beads leaves a net charge of +0.5 e.

    Chain exclusions: bonded pairs (i, i+1) are FULL.  With scaled14 set, the
    chain's exclusions follow ProtoMol's `exclude scaled1-4` rule
    (BuildTopology.cpp:1110-1158): 1-2 and 1-3 pairs FULL, 1-4 pairs (i, i+3)
    MODIFIED, coulombScalingFactor = scaled14, and the bead's 1-4 LJ columns
    from eps 0.1 / sigma 3.5."""
    rng = np.random.default_rng(seed)
    n_water = (n_total - n_beads) // 3
    if n_beads + 3 * n_water != n_total:
        raise ValueError("n_total - n_beads must be a multiple of 3")
    beads = _polymer_walk(n_beads, box, rng)
    n_side = int(math.ceil((n_water * oversample) ** (1.0 / 3.0)))
    o = _jittered_lattice(n_side, box / n_side, jitter, rng)
    wat = _water_atoms(o, rng)                     # (S, 3, 3)
    keep = _clear_of(wat, beads, box, overlap)
    sites = np.flatnonzero(keep)
    if sites.size < n_water:
        raise RuntimeError("not enough lattice sites after overlap removal")
    sites = np.sort(rng.permutation(sites)[:n_water])
    wpart = _water_part(o[sites], np.random.default_rng(seed + 7), o_type=0, h_type=1)
    wpart["pos"] = wat[sites].reshape(-1, 3)
    bonds = np.stack([np.arange(n_beads - 1), np.arange(1, n_beads)], 1)
    excl = np.concatenate([bonds, np.full((n_beads - 1, 1), FULL)], 1)
    if scaled14 is not None:
        b = np.arange(n_beads)
        e13 = np.stack([b[:-2], b[2:], np.full(n_beads - 2, FULL)], 1)
        e14 = np.stack([b[:-3], b[3:], np.full(n_beads - 3, MODIFIED)], 1)
        excl = np.concatenate([excl, e13, e14])
    ppart = dict(pos=beads,
                 charge=np.where(np.arange(n_beads) % 2 == 0, BEAD_Q, -BEAD_Q),
                 atype=np.full(n_beads, 2), mass=np.full(n_beads, BEAD_MASS),
                 molecule=np.zeros(n_beads, int),
                 bonds=bonds, bond_par=np.tile([BEAD_KB, BEAD_BOND], (n_beads - 1, 1)),
                 angles=np.zeros((0, 3), int), angle_par=np.zeros((0, 2)),
                 exclusions=excl)
    eps, rmin2 = [EPS_O, 0.0, BEAD_EPS], [rmin_half(SIGMA_O), 0.0, rmin_half(BEAD_SIGMA)]
    if scaled14 is None:
        lj = charmm_table(eps, rmin2)
    else:
        lj = charmm_table(eps, rmin2, [EPS_O, 0.0, BEAD_EPS14],
                          [rmin_half(SIGMA_O), 0.0, rmin_half(BEAD_SIGMA14)])
    s = _assemble(f"polymer_water{n_total}", [box] * 3, [ppart, wpart], lj,
                  csf=1.0 if scaled14 is None else float(scaled14))
    # keep every coordinate inside [0, box): the lattice and walk already are,
    # except walk beads that wandered; wrap them (positions are periodic)
    s.pos = s.pos - box * np.floor(s.pos / box)
    return s


def big_water(n_water: int = 333000, box: float = 216.0, seed: int = 2, jitter: float = 0.3) -> MolSystem:
    """Large water box: n_water 3-site waters on the first n_water sites of a
    seeded jittered cubic lattice filling the box."""
    rng = np.random.default_rng(seed)
    n_side = int(math.ceil(n_water ** (1.0 / 3.0)))
    o = _jittered_lattice(n_side, box / n_side, jitter, rng)
    o = o[np.sort(rng.permutation(o.shape[0])[:n_water])]
    lj = charmm_table([EPS_O, 0.0], [rmin_half(SIGMA_O), 0.0])
    return _assemble(f"water{3 * n_water}", [box] * 3, [_water_part(o, rng)], lj)


def tiled(s: MolSystem, reps=(2, 2, 2)) -> MolSystem:
    """The periodic system s repeated reps[k] times along axis k: the box
    grows by reps, copy (i, j, k) is shifted by (i Lx, j Ly, k Lz) and gets its
    own atom and molecule numbers, so each copy's atoms see the neighbourhood
    they see in s (for cutoffs below half the small box).  Atom a of copy c is
    atom c * s.n + a, copies ordered x-major."""
    reps = tuple(int(r) for r in reps)
    shifts = np.array([(i, j, k) for i in range(reps[0]) for j in range(reps[1]) for k in range(reps[2])],
                      dtype=np.float64) * s.box
    nc, n = len(shifts), s.n
    nmol = int(s.molecule.max()) + 1 if n else 0
    off = (np.arange(nc) * n)[:, None, None]

    def rep_idx(a):
        return (a[None, :, :] + off).reshape(-1, a.shape[1]).astype(np.int32)

    excl = np.tile(s.exclusions, (nc, 1)).astype(np.int32)
    excl[:, :2] = rep_idx(s.exclusions[:, :2])
    return MolSystem(
        name=f"{s.name}_x{reps[0]}{reps[1]}{reps[2]}",
        pos=(s.pos[None, :, :] + shifts[:, None, :]).reshape(-1, 3),
        charge=np.tile(s.charge, nc), atype=np.tile(s.atype, nc),
        molecule=(s.molecule[None, :] + (np.arange(nc) * nmol)[:, None]).reshape(-1).astype(np.int32),
        mass=np.tile(s.mass, nc), lj=s.lj, box=s.box * np.array(reps, dtype=np.float64),
        origin=s.origin.copy(),
        bonds=rep_idx(s.bonds), bond_par=np.tile(s.bond_par, (nc, 1)),
        angles=rep_idx(s.angles), angle_par=np.tile(s.angle_par, (nc, 1)),
        exclusions=excl, csf=s.csf,
        vel=None if s.vel is None else np.tile(s.vel, (nc, 1)))


def pme_grid_size(length: float, spacing: float = 1.0) -> int:
    """Smallest 2^a 3^b 5^c >= length / spacing (even, as the PME order
    check requires grid >= order)."""
    target = max(8, int(math.ceil(length / spacing)))
    n = target
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and n % 2 == 0:
            return n
        n += 1


def replica_box(n_atoms_target: int, seed: int, density: float = 0.0334,
                solute_frac: tuple[float, float] = (0.05, 0.10)) -> MolSystem:
    """One water box with a random bead-polymer solute holding 5-10% of the
    atoms; box edge from `density` atoms/A^3 (literal reading)."""
    rng = np.random.default_rng(seed)
    frac = rng.uniform(*solute_frac)
    n_beads = max(2, int(round(n_atoms_target * frac)))
    n_water = max(1, (n_atoms_target - n_beads) // 3)
    n_total = n_beads + 3 * n_water
    box = (n_total / density) ** (1.0 / 3.0)
    # a dense solute can clear too many lattice sites: denser lattices then
    for over in (1.25, 1.6, 2.2, 3.0):
        try:
            return polymer_water(n_total=n_total, box=box, n_beads=n_beads, seed=seed, oversample=over)
        except RuntimeError as err:
            last = err
    raise last
