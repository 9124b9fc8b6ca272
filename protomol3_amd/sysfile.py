"""Plain-text system file (PMXSYS) shared by every checker and tool.

Layout (whitespace separated, '#' starts a comment line):
    PMXSYS 1
    N T NE
    box ex ey ez ox oy oz
    csf c
    T*T lines   A B A14 B14
    N lines     type scaledCharge molecule
    NE lines    a1 a2 cls            (1 FULL, 2 MODIFIED)
    N lines     x y z
"""
from __future__ import annotations

import io

import numpy as np

from .systems import MolSystem


def write_sysfile(path: str, s: MolSystem, pos: np.ndarray | None = None) -> None:
    pos = s.pos if pos is None else pos
    t = s.lj.shape[0]
    buf = io.StringIO()
    buf.write("PMXSYS 1\n")
    buf.write(f"{s.n} {t} {s.exclusions.shape[0]}\n")
    b, o = s.box, s.origin
    buf.write("box " + " ".join(repr(float(v)) for v in (*b, *o)) + "\n")
    buf.write(f"csf {float(s.csf)!r}\n")
    for i in range(t):
        for j in range(t):
            buf.write(" ".join(repr(float(v)) for v in s.lj[i, j]) + "\n")
    for i in range(s.n):
        buf.write(f"{int(s.atype[i])} {float(s.charge[i])!r} {int(s.molecule[i])}\n")
    for a1, a2, c in s.exclusions:
        buf.write(f"{int(a1)} {int(a2)} {int(c)}\n")
    np.savetxt(buf, pos, fmt="%.17g")
    with open(path, "w") as f:
        f.write(buf.getvalue())


def read_forces(path: str) -> tuple[float, float, np.ndarray]:
    """Output of the reference driver: 'E_coulomb E_lj alpha' then N force rows."""
    with open(path) as f:
        head = f.readline().split()
        frc = np.loadtxt(f, ndmin=2)
    return float(head[0]), float(head[1]), frc
