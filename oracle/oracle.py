"""ctypes wrapper for the CPU oracle (oracle/pm_oracle.cpp) -- TEST
INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by the product path."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "_build", "libpmoracle.so")

SW_NONE, SW_CUTOFF, SW_C1, SW_C2 = 0, 1, 2, 3
COUL_NONE, COUL_CUTOFF, COUL_PME = 0, 1, 2


class PmoSystem(C.Structure):
    _fields_ = [("n", C.c_int), ("ntypes", C.c_int), ("nexcl", C.c_int),
                ("pos", C.POINTER(C.c_double)), ("charge", C.POINTER(C.c_double)),
                ("type", C.POINTER(C.c_int)), ("molecule", C.POINTER(C.c_int)),
                ("lj", C.POINTER(C.c_double)), ("excl", C.POINTER(C.c_int)),
                ("box", C.c_double * 3), ("origin", C.c_double * 3), ("csf", C.c_double)]


class PmoSpec(C.Structure):
    _fields_ = [("cellsize", C.c_double), ("lj_switch", C.c_int), ("lj_on", C.c_double),
                ("lj_rc", C.c_double), ("coul_kind", C.c_int), ("coul_switch", C.c_int),
                ("coul_on", C.c_double), ("coul_rc", C.c_double), ("pme_grid", C.c_int * 3),
                ("pme_order", C.c_int), ("pme_accuracy", C.c_double), ("pme_alpha", C.c_double),
                ("pme_real", C.c_int), ("pme_recip", C.c_int), ("pme_corr", C.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-f", "oracle/Makefile"], cwd=ROOT, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.pmo_evaluate.argtypes = [C.POINTER(PmoSystem), C.POINTER(PmoSpec), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.c_int]
        L.pmo_pme_terms.argtypes = [C.POINTER(PmoSystem), C.POINTER(PmoSpec), C.POINTER(C.c_double)]
        L.pmo_pme_alpha.argtypes = [C.c_double, C.c_double]
        L.pmo_pme_alpha.restype = C.c_double
        L.pmo_cells.argtypes = [C.POINTER(PmoSystem), C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.pmo_pairs.argtypes = [C.POINTER(PmoSystem), C.c_double, C.c_double, C.POINTER(C.c_int), C.c_long,
                                C.POINTER(C.c_long)]
        L.pmo_pairs_bruteforce.argtypes = [C.POINTER(PmoSystem), C.c_double, C.POINTER(C.c_int), C.c_long,
                                           C.POINTER(C.c_long)]
        L.pmo_ewald_direct.argtypes = [C.POINTER(PmoSystem), C.c_double, C.c_double, C.c_int,
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pmo_bspline.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.pmo_poly5.argtypes = [C.c_double]
        L.pmo_poly5.restype = C.c_double
        L.pmo_fft3d.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OracleSystem:
    """Holds contiguous copies of a MolSystem's arrays for the C oracle."""

    def __init__(self, s, pos=None):
        self.pos = np.ascontiguousarray(s.pos if pos is None else pos, dtype=np.float64)
        self.charge = np.ascontiguousarray(s.charge, dtype=np.float64)
        self.type = np.ascontiguousarray(s.atype, dtype=np.int32)
        self.molecule = np.ascontiguousarray(s.molecule, dtype=np.int32)
        self.lj = np.ascontiguousarray(s.lj, dtype=np.float64)
        self.excl = np.ascontiguousarray(s.exclusions, dtype=np.int32).reshape(-1, 3)
        self.n = self.pos.shape[0]
        st = PmoSystem()
        st.n, st.ntypes, st.nexcl = self.n, self.lj.shape[0], self.excl.shape[0]
        st.pos, st.charge = _dp(self.pos), _dp(self.charge)
        st.type, st.molecule = _ip(self.type), _ip(self.molecule)
        st.lj, st.excl = _dp(self.lj), _ip(self.excl)
        for k in range(3):
            st.box[k] = float(s.box[k])
            st.origin[k] = float(s.origin[k])
        st.csf = float(s.csf)
        self.st = st


def make_spec(cellsize, lj=None, coul=None, pme=None) -> PmoSpec:
    """lj = (switch, on, rc); coul = (switch, on, rc) for cutoff Coulomb;
    pme = dict(rc, grid, order, accuracy, alpha, real, recip, corr, switch)."""
    sp = PmoSpec()
    sp.cellsize = cellsize
    if lj is not None:
        sp.lj_switch, sp.lj_on, sp.lj_rc = lj
    if coul is not None:
        sp.coul_kind = COUL_CUTOFF
        sp.coul_switch, sp.coul_on, sp.coul_rc = coul
    if pme is not None:
        sp.coul_kind = COUL_PME
        sp.coul_switch = pme.get("switch", SW_CUTOFF)
        sp.coul_rc = pme["rc"]
        for k in range(3):
            sp.pme_grid[k] = pme["grid"][k]
        sp.pme_order = pme.get("order", 4)
        sp.pme_accuracy = pme.get("accuracy", 1e-6)
        sp.pme_alpha = pme.get("alpha", -1.0)
        sp.pme_real = int(pme.get("real", True))
        sp.pme_recip = int(pme.get("recip", True))
        sp.pme_corr = int(pme.get("corr", True))
    return sp


def evaluate(s, spec: PmoSpec, pos=None, nthreads: int = 1):
    """Returns (forces (N,3), e_coulomb, e_lj)."""
    os_ = OracleSystem(s, pos)
    f = np.zeros((os_.n, 3))
    e = np.zeros(2)
    rc = lib().pmo_evaluate(C.byref(os_.st), C.byref(spec), _dp(f), _dp(e), nthreads)
    if rc != 0:
        raise ValueError(f"pmo_evaluate failed: {rc}")
    return f, float(e[0]), float(e[1])


def pme_terms(s, spec: PmoSpec, pos=None) -> dict:
    os_ = OracleSystem(s, pos)
    out = np.zeros(6)
    if lib().pmo_pme_terms(C.byref(os_.st), C.byref(spec), _dp(out)) != 0:
        raise ValueError("pmo_pme_terms failed")
    return dict(zip(["real", "recip", "intra", "self", "charged", "alpha"], out.tolist()))


def pme_alpha(rc: float, accuracy: float) -> float:
    return lib().pmo_pme_alpha(rc, accuracy)


def cells(s, cellsize: float, pos=None):
    os_ = OracleSystem(s, pos)
    out = np.zeros(os_.n, np.int32)
    dims = np.zeros(3, np.int32)
    if lib().pmo_cells(C.byref(os_.st), cellsize, _ip(out), _ip(dims)) != 0:
        raise ValueError("pmo_cells failed")
    return out, tuple(int(d) for d in dims)


def pairs(s, cellsize: float, rc: float, pos=None) -> np.ndarray:
    """(P, 2) in-cutoff, non-FULL pairs in the reference's visit order."""
    os_ = OracleSystem(s, pos)
    n = C.c_long(0)
    lib().pmo_pairs(C.byref(os_.st), cellsize, rc, None, 0, C.byref(n))
    out = np.zeros((max(n.value, 1), 2), np.int32)
    if lib().pmo_pairs(C.byref(os_.st), cellsize, rc, _ip(out), n.value, C.byref(n)) != 0:
        raise ValueError("pmo_pairs failed")
    return out[: n.value]


def pairs_bruteforce(s, rc: float, pos=None) -> np.ndarray:
    os_ = OracleSystem(s, pos)
    n = C.c_long(0)
    lib().pmo_pairs_bruteforce(C.byref(os_.st), rc, None, 0, C.byref(n))
    out = np.zeros((max(n.value, 1), 2), np.int32)
    if lib().pmo_pairs_bruteforce(C.byref(os_.st), rc, _ip(out), n.value, C.byref(n)) != 0:
        raise ValueError("pmo_pairs_bruteforce failed")
    return out[: n.value]


def ewald_direct(s, alpha: float, rc_real: float, kmax: int, pos=None):
    os_ = OracleSystem(s, pos)
    f = np.zeros((os_.n, 3))
    e = C.c_double(0)
    if lib().pmo_ewald_direct(C.byref(os_.st), alpha, rc_real, kmax, _dp(f), C.byref(e)) != 0:
        raise ValueError("pmo_ewald_direct failed")
    return f, e.value


def bspline(order: int, w: float):
    th = np.zeros(order)
    dth = np.zeros(order)
    lib().pmo_bspline(order, w, _dp(th), _dp(dth))
    return th, dth


def poly5(x: float) -> float:
    return lib().pmo_poly5(x)


def fft3d(a: np.ndarray, sign: int) -> np.ndarray:
    b = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nx, ny, nz = b.shape
    lib().pmo_fft3d(b.ctypes.data_as(C.POINTER(C.c_double)), nx, ny, nz, sign)
    return b
