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
                ("out_energies", C.POINTER(C.c_double)), ("session", C.c_longlong)]


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
    L.pmx_server_wait_times.argtypes = [vp, C.c_longlong, C.POINTER(C.c_double)]
    L.pmx_server_open.argtypes = [vp, C.POINTER(pmx.SysDesc), C.POINTER(pmx.MdDesc), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
    L.pmx_server_close.argtypes = [vp, C.c_longlong]
    L.pmx_server_stats.argtypes = [vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_int)]
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
        # ms since the request's submit (start, done) and since the server's
        # creation (submit_at, start_at, done_at)
        self.start_ms = self.done_ms = None
        self.submit_at = self.start_at = self.done_at = None


class Session:
    """A system kept resident on the device between requests
    (pmx_server_open / pmx_server_close)."""

    def __init__(self, srv, sid, n, keep, md: bool):
        self.srv, self.id, self.n, self.keep, self.md = srv, sid, n, keep, md

    def forces(self, x: np.ndarray) -> Ticket:
        """Forces and energies at positions x (a session opened without MD)."""
        x = np.ascontiguousarray(x, np.float64)
        pmx._check_shape(x, (self.n, 3), "positions")
        t = Ticket(REQ_FORCES, self.n, [x])
        r = Request(kind=REQ_FORCES, x=pmx._dp(x), out_a=pmx._dp(t.a), out_energies=pmx._dp(t.e), session=self.id)
        t.keep.append(r)
        self.srv._submit(r, t)
        return t

    def run(self, nsteps: int, x: np.ndarray | None = None, v: np.ndarray | None = None) -> Ticket:
        """nsteps more MD steps from the resident state (or from x, v)."""
        t = Ticket(REQ_TRAJECTORY, self.n, [])
        r = Request(kind=REQ_TRAJECTORY, nsteps=nsteps, out_a=pmx._dp(t.a), out_v=pmx._dp(t.v),
                    out_energies=pmx._dp(t.e), session=self.id)
        if x is not None:
            x = np.ascontiguousarray(x, np.float64)
            v = np.ascontiguousarray(v, np.float64)
            pmx._check_shape(x, (self.n, 3), "positions")
            pmx._check_shape(v, (self.n, 3), "velocities")
            t.keep += [x, v]
            r.x, r.v = pmx._dp(x), pmx._dp(v)
        t.keep.append(r)
        self.srv._submit(r, t)
        return t

    def close(self):
        if self.id:
            rc = self.srv.lib.pmx_server_close(self.srv.h, self.id)
            self.id = 0
            if rc:
                raise pmx.PmxError(rc, "pmx_server_close")


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

    def open(self, s: systems.MolSystem, spec: pmx.ForceSpec, precision: int, md: dict | None = None) -> Session:
        """Opens a session: the system stays resident on the device until
        close().  md: None for a force-only session, else the arguments of
        make_md_desc after the system (timestep_fs, skin, integrator,
        temperature, gamma, seed); the state starts at s.pos, s.vel."""
        d, keep = pmx.make_desc(s, spec, precision)
        keep = [d, keep]
        sid = C.c_longlong()
        if md is None:
            rc = self.lib.pmx_server_open(self.h, C.byref(d), None, None, None, C.byref(sid))
        else:
            m, mkeep = pmx.make_md_desc(s, **md)
            x = np.ascontiguousarray(s.pos, np.float64)
            v = np.ascontiguousarray(s.vel, np.float64)
            keep += [m, mkeep, x, v]
            rc = self.lib.pmx_server_open(self.h, C.byref(d), C.byref(m), pmx._dp(x), pmx._dp(v), C.byref(sid))
        if rc:
            raise pmx.PmxError(rc, "pmx_server_open")
        return Session(self, sid.value, s.n, keep, md is not None)

    def stats(self) -> tuple[int, int]:
        """(largest footprint sum resident at once, most items running at once)."""
        b, r = C.c_ulonglong(), C.c_int()
        self.lib.pmx_server_stats(self.h, C.byref(b), C.byref(r))
        return int(b.value), int(r.value)

    def _submit(self, r, t):
        rc = self.lib.pmx_server_submit(self.h, C.byref(r), C.byref(t.id))
        if rc:
            raise pmx.PmxError(rc, "pmx_server_submit")

    def wait(self, t: Ticket) -> Ticket:
        tm = (C.c_double * 3)()
        rc = self.lib.pmx_server_wait_times(self.h, t.id, tm)
        if rc:
            raise pmx.PmxError(rc, "request failed: " + self.lib.pmx_server_last_error(self.h).decode())
        t.submit_at, t.start_at, t.done_at = tm[0], tm[1], tm[2]
        t.start_ms, t.done_ms = tm[1] - tm[0], tm[2] - tm[0]
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
