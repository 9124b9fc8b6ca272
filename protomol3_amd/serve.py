"""ctypes binding of the request loop (include/pmx_serve.h,
protomol3_amd/lib/libpmx_serve.so) and the serving workload generator."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import pmx, systems

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libpmx_serve.so")
REQ_FORCES, REQ_TRAJECTORY = 0, 1


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("sys", C.POINTER(pmx.SysDesc)), ("md", C.POINTER(pmx.MdDesc)),
                ("x", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)), ("nsteps", C.c_int),
                ("out_a", C.POINTER(C.c_double)), ("out_v", C.POINTER(C.c_double)),
                ("out_energies", C.POINTER(C.c_double))]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    pmx.load()  # the engine library first (loud failure if missing)
    if not os.path.exists(LIB_PATH):
        raise pmx.PmxError(-5, f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.pmx_server_create.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.POINTER(vp)]
    L.pmx_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_longlong)]
    L.pmx_server_wait.argtypes = [vp, C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pmx_server_pending.argtypes = [vp, C.POINTER(C.c_int)]
    L.pmx_server_destroy.argtypes = [vp]
    L.pmx_server_last_error.argtypes = [vp]
    L.pmx_server_last_error.restype = C.c_char_p
    _lib = L
    return L


class Ticket:
    """A submitted request: its inputs are kept alive until the result is read."""

    def __init__(self, kind, n, keep):
        self.kind, self.n, self.keep = kind, n, keep
        self.a = np.zeros((n, 3))
        self.v = np.zeros((n, 3)) if kind == REQ_TRAJECTORY else None
        self.e = np.zeros(5 if kind == REQ_TRAJECTORY else 2)
        self.id = C.c_longlong()
        self.start_ms = self.done_ms = None


class Server:
    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 64 << 30):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.pmx_server_create(device, slots, hbm_budget, C.byref(self.h))
        if rc:
            raise pmx.PmxError(rc, "pmx_server_create")

    def submit_forces(self, s: systems.MolSystem, spec: pmx.ForceSpec, precision: int, x=None) -> Ticket:
        d, keep = pmx.make_desc(s, spec, precision)
        x = np.ascontiguousarray(s.pos if x is None else x, np.float64)
        t = Ticket(REQ_FORCES, s.n, [d, keep, x])
        r = Request(kind=REQ_FORCES, sys=C.pointer(d), x=pmx._dp(x), out_a=pmx._dp(t.a), out_energies=pmx._dp(t.e))
        t.keep.append(r)
        self._submit(r, t)
        return t

    def submit_trajectory(self, s: systems.MolSystem, spec: pmx.ForceSpec, precision: int, nsteps: int,
                          timestep_fs: float = 1.0, skin: float = 1.0, integrator: int = pmx.LANGEVIN,
                          temperature: float = 300.0, gamma: float = 1.0, seed: int = 1234) -> Ticket:
        d, keep = pmx.make_desc(s, spec, precision)
        md, mkeep = pmx.make_md_desc(s, timestep_fs, skin, integrator, temperature, gamma, seed)
        x = np.ascontiguousarray(s.pos, np.float64)
        v = np.ascontiguousarray(s.vel, np.float64)
        t = Ticket(REQ_TRAJECTORY, s.n, [d, keep, md, mkeep, x, v])
        r = Request(kind=REQ_TRAJECTORY, sys=C.pointer(d), md=C.pointer(md), x=pmx._dp(x), v=pmx._dp(v),
                    nsteps=nsteps, out_a=pmx._dp(t.a), out_v=pmx._dp(t.v), out_energies=pmx._dp(t.e))
        t.keep.append(r)
        self._submit(r, t)
        return t

    def _submit(self, r, t):
        rc = self.lib.pmx_server_submit(self.h, C.byref(r), C.byref(t.id))
        if rc:
            raise pmx.PmxError(rc, "pmx_server_submit")

    def wait(self, t: Ticket) -> Ticket:
        s0, s1 = C.c_double(), C.c_double()
        rc = self.lib.pmx_server_wait(self.h, t.id, C.byref(s0), C.byref(s1))
        if rc:
            raise pmx.PmxError(rc, "request failed: " + self.lib.pmx_server_last_error(self.h).decode())
        t.start_ms, t.done_ms = s0.value, s1.value
        return t

    def close(self):
        if self.h:
            self.lib.pmx_server_destroy(self.h)
            self.h = C.c_void_p()


def request_spec(s: systems.MolSystem, rc: float = 9.0, order: int = 4) -> pmx.ForceSpec:
    """Force lines of the batched and served systems: LJ C2 from rc-2 to rc,
    PME with a ~1 A grid (2^a 3^b 5^c), alpha from accuracy 1e-6."""
    grid = tuple(systems.pme_grid_size(L, 1.0) for L in s.box)
    return pmx.ForceSpec(max(rc / 2.0, 4.0), lj=(pmx.SW_C2, rc - 2.0, rc),
                         pme=dict(rc=rc, grid=grid, order=order, alpha=-1.0, accuracy=1e-6))
