"""ctypes binding of the pmx C ABI (include/pmx.h).

The library is the in-tree build protomol3_amd/lib/libpmx.so (gfx950 code
objects).  There is no fallback: if the library or a gfx950 device is
missing, every call raises PmxError.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from .systems import MolSystem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpmx.so")

FP64, MIXED = 0, 1
SW_NONE, SW_CUTOFF, SW_C1, SW_C2 = 0, 1, 2, 3
COUL_NONE, COUL_CUTOFF, COUL_PME = 0, 1, 2
PME_REAL, PME_RECIP, PME_CORR = 1, 2, 4
EVAL_PAIRSTATS, EVAL_TIMING = 1, 2
STAGES = ("sort", "list", "pair", "spread", "fft", "gather", "fixups", "finalize")

ERRORS = {-1: "PMX_ERR_ARG", -2: "PMX_ERR_OOM", -3: "PMX_ERR_HIP", -4: "PMX_ERR_UNSUPPORTED",
          -5: "PMX_ERR_NODEV", -6: "PMX_ERR_OVERFLOW"}


class PmxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class SysDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("ntypes", C.c_int), ("nexcl", C.c_int),
                ("charge", C.POINTER(C.c_double)), ("type", C.POINTER(C.c_int)),
                ("molecule", C.POINTER(C.c_int)), ("lj", C.POINTER(C.c_double)),
                ("excl", C.POINTER(C.c_int)), ("box", C.c_double * 3), ("origin", C.c_double * 3),
                ("csf", C.c_double), ("cellsize", C.c_double), ("lj_switch", C.c_int),
                ("lj_on", C.c_double), ("lj_rc", C.c_double), ("coul_kind", C.c_int),
                ("coul_switch", C.c_int), ("coul_on", C.c_double), ("coul_rc", C.c_double),
                ("pme_grid", C.c_int * 3), ("pme_order", C.c_int), ("pme_accuracy", C.c_double),
                ("pme_alpha", C.c_double), ("pme_terms", C.c_int), ("precision", C.c_int)]


class MdDesc(C.Structure):
    _fields_ = [("nbonds", C.c_int), ("nangles", C.c_int), ("bonds", C.POINTER(C.c_int)),
                ("bond_par", C.POINTER(C.c_double)), ("angles", C.POINTER(C.c_int)),
                ("angle_par", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("timestep_fs", C.c_double), ("skin", C.c_double), ("integrator", C.c_int),
                ("temperature", C.c_double), ("gamma", C.c_double), ("seed", C.c_ulonglong),
                ("graphs", C.c_int)]


LEAPFROG, LANGEVIN = 0, 1
TUNE_FFT_LINES, TUNE_PAIR_BLOCKS_PER_CU, TUNE_PAIR_SCHEDULE = 2, 3, 4
TUNE_SPREAD_STREAM, TUNE_LIST_PRUNE, TUNE_FIXUP_STREAM, TUNE_FUSED_SPREAD = 6, 7, 8, 9
TUNE_FFT_PLANES = 10


EXPORTED = ["pmx_ctx_create", "pmx_ctx_destroy", "pmx_last_error", "pmx_sys_create", "pmx_sys_destroy",
            "pmx_sys_uncache", "pmx_sys_evaluate", "pmx_sys_set_positions", "pmx_sys_compute",
            "pmx_sys_get_forces", "pmx_sys_sync", "pmx_sys_alpha", "pmx_sys_pair_stats", "pmx_pair_hash",
            "pmx_sys_cells", "pmx_sys_stage_times", "pmx_sys_timing_begin", "pmx_sys_timing_end",
            "pmx_md_init", "pmx_md_set_state", "pmx_md_run", "pmx_md_get_state", "pmx_md_stats",
            "pmx_langevin_noise", "pmx_sys_device_bytes", "pmx_estimate_bytes", "pmx_md_run_many",
            "pmx_sys_time_pair", "pmx_sys_set_tuning", "pmx_sys_compute_noalloc", "pmx_md_run_batch",
            "pmx_sys_compute_batch"]

_lib = None


def load() -> C.CDLL:
    """Loads libpmx.so; raises PmxError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PmxError(-5, f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    L.pmx_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.pmx_ctx_destroy.argtypes = [vp]
    L.pmx_last_error.argtypes = [vp]
    L.pmx_last_error.restype = C.c_char_p
    L.pmx_sys_create.argtypes = [vp, C.POINTER(SysDesc), C.POINTER(vp)]
    L.pmx_sys_destroy.argtypes = [vp]
    L.pmx_sys_uncache.argtypes = [vp]
    L.pmx_sys_evaluate.argtypes = [vp, dp, dp, dp, C.c_uint]
    L.pmx_sys_set_positions.argtypes = [vp, dp]
    L.pmx_sys_compute.argtypes = [vp, C.c_uint]
    L.pmx_sys_compute_noalloc.argtypes = [vp, C.c_uint]
    L.pmx_sys_get_forces.argtypes = [vp, dp, dp]
    L.pmx_sys_sync.argtypes = [vp]
    L.pmx_sys_alpha.argtypes = [vp, dp]
    L.pmx_sys_pair_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_ulonglong)]
    L.pmx_pair_hash.argtypes = [C.c_int, C.c_int]
    L.pmx_pair_hash.restype = C.c_ulonglong
    L.pmx_sys_cells.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pmx_sys_stage_times.argtypes = [vp, C.POINTER(C.c_float)]
    L.pmx_sys_timing_begin.argtypes = [vp]
    L.pmx_sys_timing_end.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    L.pmx_md_init.argtypes = [vp, C.POINTER(MdDesc)]
    L.pmx_md_set_state.argtypes = [vp, dp, dp]
    L.pmx_md_run.argtypes = [vp, C.c_int]
    L.pmx_md_run_many.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
    L.pmx_md_run_batch.argtypes = [C.POINTER(vp), C.c_int, C.c_int]
    L.pmx_sys_compute_batch.argtypes = [C.POINTER(vp), C.c_int, C.c_uint]
    L.pmx_sys_time_pair.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
    L.pmx_sys_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.pmx_md_get_state.argtypes = [vp, dp, dp, dp]
    L.pmx_md_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.pmx_langevin_noise.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_int, C.POINTER(C.c_double)]
    L.pmx_sys_device_bytes.argtypes = [vp, C.POINTER(C.c_ulonglong)]
    L.pmx_estimate_bytes.argtypes = [C.POINTER(SysDesc), C.c_int]
    L.pmx_estimate_bytes.restype = C.c_ulonglong
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def langevin_noise(seed: int, half: int, n: int) -> np.ndarray:
    """(n, 3) Langevin noise the device draws for half kick `half` (host)."""
    g = np.zeros((n, 3))
    rc = load().pmx_langevin_noise(seed, half, n, _dp(g))
    if rc:
        raise PmxError(rc, "pmx_langevin_noise")
    return g


def pair_hash_set(pairs: np.ndarray) -> int:
    """The library's order-independent pair-set hash, computed on the host
    (splitmix64 of (min << 32 | max), summed mod 2^64)."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    lo = np.minimum(p[:, 0], p[:, 1]).astype(np.uint64)
    hi = np.maximum(p[:, 0], p[:, 1]).astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (lo << np.uint64(32)) | hi
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return int(np.sum(z, dtype=np.uint64))


class Context:
    def __init__(self, device: int = 0):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.pmx_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise PmxError(rc, f"pmx_ctx_create(device={device}) failed")

    def check(self, rc: int, what: str):
        if rc != 0:
            raise PmxError(rc, f"{what}: {self.lib.pmx_last_error(self.h).decode()}")

    def close(self):
        if self.h:
            self.lib.pmx_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class ForceSpec:
    """The force lines of a ProtoMol .conf for this engine:
    lj = (switch, switchon, cutoff) for `LennardJones -algorithm NonbondedCutoff`;
    coul = (switch, switchon, cutoff) for `Coulomb -algorithm NonbondedCutoff`;
    pme = dict(rc, grid, order, accuracy, alpha, terms, switch) for
    `Coulomb -algorithm PMEwald`."""

    def __init__(self, cellsize: float, lj=None, coul=None, pme=None):
        self.cellsize, self.lj, self.coul, self.pme = cellsize, lj, coul, pme


def _check_shape(a: np.ndarray, shape: tuple, what: str) -> None:
    if a.shape != shape:
        raise ValueError(f"{what}: shape {a.shape}, expected {shape}")


def make_desc(s: MolSystem, spec: ForceSpec, precision: int = FP64):
    """(pmx_sys_desc, arrays it points into) for a system and force spec.
    Raises ValueError when an array's shape does not match the atom count
    (the library reads exactly n, ntypes^2 * 4 and 3 * nexcl values)."""
    keep = [np.ascontiguousarray(s.charge, np.float64), np.ascontiguousarray(s.atype, np.int32),
            np.ascontiguousarray(s.molecule, np.int32), np.ascontiguousarray(s.lj, np.float64),
            np.ascontiguousarray(s.exclusions, np.int32).reshape(-1, 3)]
    q, t, m, lj, ex = keep
    n = s.n
    _check_shape(q, (n,), "charge")
    _check_shape(t, (n,), "atype")
    _check_shape(m, (n,), "molecule")
    if lj.ndim != 3 or lj.shape[0] != lj.shape[1] or lj.shape[2] != 4:
        raise ValueError(f"lj: shape {lj.shape}, expected (ntypes, ntypes, 4)")
    if n and (t.min() < 0 or t.max() >= lj.shape[0]):
        raise ValueError("atype: LJ type outside the table")
    if ex.size and (ex[:, :2].min() < 0 or ex[:, :2].max() >= n):
        raise ValueError("exclusions: atom index outside the system")
    d = SysDesc()
    d.n, d.ntypes, d.nexcl = s.n, lj.shape[0], ex.shape[0]
    d.charge = _dp(q)
    d.type = t.ctypes.data_as(C.POINTER(C.c_int))
    d.molecule = m.ctypes.data_as(C.POINTER(C.c_int))
    d.lj = _dp(lj)
    d.excl = ex.ctypes.data_as(C.POINTER(C.c_int))
    for k in range(3):
        d.box[k] = float(s.box[k])
        d.origin[k] = float(s.origin[k])
    d.csf = float(s.csf)
    d.cellsize = float(spec.cellsize)
    if spec.lj is not None:
        d.lj_switch, d.lj_on, d.lj_rc = spec.lj
    if spec.coul is not None:
        d.coul_kind = COUL_CUTOFF
        d.coul_switch, d.coul_on, d.coul_rc = spec.coul
    if spec.pme is not None:
        p = spec.pme
        d.coul_kind = COUL_PME
        d.coul_switch = p.get("switch", SW_CUTOFF)
        d.coul_rc = p["rc"]
        for k in range(3):
            d.pme_grid[k] = p["grid"][k]
        d.pme_order = p.get("order", 4)
        d.pme_accuracy = p.get("accuracy", 1e-6)
        d.pme_alpha = p.get("alpha", -1.0)
        d.pme_terms = p.get("terms", PME_REAL | PME_RECIP | PME_CORR)
    d.precision = precision
    return d, keep


def make_md_desc(s: MolSystem, timestep_fs: float, skin: float = 1.0, integrator: int = LEAPFROG,
                 temperature: float = 300.0, gamma: float = 1.0, seed: int = 1234, graphs: bool = False):
    """(pmx_md_desc, arrays it points into)."""
    keep = [np.ascontiguousarray(s.bonds, np.int32).reshape(-1, 2),
            np.ascontiguousarray(s.bond_par, np.float64).reshape(-1, 2),
            np.ascontiguousarray(s.angles, np.int32).reshape(-1, 3),
            np.ascontiguousarray(s.angle_par, np.float64).reshape(-1, 2),
            np.ascontiguousarray(s.mass, np.float64)]
    b, bp, a, ap, m = keep
    _check_shape(m, (s.n,), "mass")
    _check_shape(bp, (b.shape[0], 2), "bond_par")
    _check_shape(ap, (a.shape[0], 2), "angle_par")
    if b.size and (b.min() < 0 or b.max() >= s.n):
        raise ValueError("bonds: atom index outside the system")
    if a.size and (a.min() < 0 or a.max() >= s.n):
        raise ValueError("angles: atom index outside the system")
    d = MdDesc()
    d.nbonds, d.nangles = b.shape[0], a.shape[0]
    d.bonds = b.ctypes.data_as(C.POINTER(C.c_int))
    d.bond_par = _dp(bp)
    d.angles = a.ctypes.data_as(C.POINTER(C.c_int))
    d.angle_par = _dp(ap)
    d.mass = _dp(m)
    d.timestep_fs = timestep_fs
    d.skin = skin
    d.integrator = integrator
    d.temperature = temperature
    d.gamma = gamma
    d.seed = seed
    d.graphs = 1 if graphs else 0
    return d, keep


def run_many(systems_: list, nsteps: int) -> None:
    """pmx_md_run_many: nsteps MD steps on every System (each after md_init
    and md_set_state); systems on different Contexts run concurrently."""
    if not systems_:
        return
    lib = systems_[0].ctx.lib
    arr = (C.c_void_p * len(systems_))(*[sy.h for sy in systems_])
    rc = lib.pmx_md_run_many(arr, len(systems_), nsteps)
    if rc:
        msgs = {sy.ctx.lib.pmx_last_error(sy.ctx.h).decode() for sy in systems_}
        raise PmxError(rc, "pmx_md_run_many: " + "; ".join(m for m in msgs if m))


def run_batch(systems_: list, nsteps: int) -> None:
    """pmx_md_run_batch: nsteps MD steps on every System (each after
    md_init and md_set_state; one Context, one precision, the same kinds of
    force, PME of one stencil order), every kernel of a step launched once
    for all of them."""
    if not systems_:
        return
    ctx = systems_[0].ctx
    arr = (C.c_void_p * len(systems_))(*[sy.h for sy in systems_])
    ctx.check(ctx.lib.pmx_md_run_batch(arr, len(systems_), nsteps), "pmx_md_run_batch")


def evaluate_batch(systems_: list, positions: list) -> list:
    """pmx_sys_compute_batch: the forces and energies of every System at its
    positions (an (n, 3) array each), every kernel of the force pass launched
    once for all of them; returns [(forces, ecoul, elj)] like
    System.evaluate."""
    if not systems_:
        return []
    ctx = systems_[0].ctx
    for sy, x in zip(systems_, positions):
        sy.set_positions(x)
    arr = (C.c_void_p * len(systems_))(*[sy.h for sy in systems_])
    ctx.check(ctx.lib.pmx_sys_compute_batch(arr, len(systems_), 0), "pmx_sys_compute_batch")
    return [sy.get_forces() for sy in systems_]


def estimate_bytes(s: MolSystem, spec: ForceSpec, precision: int = FP64, trajectory: bool = False) -> int:
    """Device bytes the engine would allocate for the system (host-side)."""
    d, keep = make_desc(s, spec, precision)
    return int(load().pmx_estimate_bytes(C.byref(d), 1 if trajectory else 0))


class System:
    """A resident device system (pmx_sys)."""

    def __init__(self, ctx: Context, s: MolSystem, spec: ForceSpec, precision: int = FP64):
        self.ctx, self.n = ctx, s.n
        d, self._keep = make_desc(s, spec, precision)
        self.desc = d
        self.h = C.c_void_p()
        ctx.check(ctx.lib.pmx_sys_create(ctx.h, C.byref(d), C.byref(self.h)), "pmx_sys_create")

    def evaluate(self, pos: np.ndarray, flags: int = 0):
        """Returns (forces (N,3), e_coulomb, e_lj) for positions (N,3)."""
        x = np.ascontiguousarray(pos, dtype=np.float64)
        _check_shape(x, (self.n, 3), "positions")
        f = np.zeros((self.n, 3))
        e = np.zeros(2)
        self.ctx.check(self.ctx.lib.pmx_sys_evaluate(self.h, _dp(x), _dp(f), _dp(e), flags), "pmx_sys_evaluate")
        return f, float(e[0]), float(e[1])

    def set_positions(self, pos: np.ndarray):
        x = np.ascontiguousarray(pos, dtype=np.float64)
        _check_shape(x, (self.n, 3), "positions")
        self.ctx.check(self.ctx.lib.pmx_sys_set_positions(self.h, _dp(x)), "pmx_sys_set_positions")

    def compute(self, flags: int = 0):
        self.ctx.check(self.ctx.lib.pmx_sys_compute(self.h, flags), "pmx_sys_compute")

    def compute_noalloc(self, flags: int = 0):
        """pmx_sys_compute without any device allocation (the pass's scratch
        is the library's static per-device array)."""
        self.ctx.check(self.ctx.lib.pmx_sys_compute_noalloc(self.h, flags), "pmx_sys_compute_noalloc")

    def sync(self):
        self.ctx.check(self.ctx.lib.pmx_sys_sync(self.h), "pmx_sys_sync")

    def get_forces(self):
        f = np.zeros((self.n, 3))
        e = np.zeros(2)
        self.ctx.check(self.ctx.lib.pmx_sys_get_forces(self.h, _dp(f), _dp(e)), "pmx_sys_get_forces")
        return f, float(e[0]), float(e[1])

    def alpha(self) -> float:
        a = C.c_double()
        self.ctx.check(self.ctx.lib.pmx_sys_alpha(self.h, C.byref(a)), "pmx_sys_alpha")
        return a.value

    def device_bytes(self) -> int:
        b = C.c_ulonglong()
        self.ctx.check(self.ctx.lib.pmx_sys_device_bytes(self.h, C.byref(b)), "pmx_sys_device_bytes")
        return b.value

    def pair_stats(self):
        n = C.c_longlong()
        h = C.c_ulonglong()
        self.ctx.check(self.ctx.lib.pmx_sys_pair_stats(self.h, C.byref(n), C.byref(h)), "pmx_sys_pair_stats")
        return n.value, h.value

    def cells(self):
        out = np.zeros(self.n, np.int32)
        dims = np.zeros(3, np.int32)
        self.ctx.check(self.ctx.lib.pmx_sys_cells(self.h, out.ctypes.data_as(C.POINTER(C.c_int)),
                                                  dims.ctypes.data_as(C.POINTER(C.c_int))), "pmx_sys_cells")
        return out, tuple(int(v) for v in dims)

    def stage_times(self) -> dict:
        ms = (C.c_float * len(STAGES))()
        self.ctx.check(self.ctx.lib.pmx_sys_stage_times(self.h, ms), "pmx_sys_stage_times")
        return dict(zip(STAGES, [float(v) for v in ms]))

    # -- device-resident MD -------------------------------------------------
    def md_init(self, s: MolSystem, timestep_fs: float, skin: float = 1.0, integrator: int = LEAPFROG,
                temperature: float = 300.0, gamma: float = 1.0, seed: int = 1234, graphs: bool = False):
        """Device MD set-up: leapfrog, or Langevin leapfrog at `temperature` (K)
        with friction `gamma` (1/ps) and noise stream `seed`; graphs: replay
        each chunk of steps as a captured HIP graph."""
        d, self._md_keep = make_md_desc(s, timestep_fs, skin, integrator, temperature, gamma, seed, graphs)
        self.ctx.check(self.ctx.lib.pmx_md_init(self.h, C.byref(d)), "pmx_md_init")

    def md_set_state(self, x: np.ndarray, v: np.ndarray):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        _check_shape(x, (self.n, 3), "positions")
        _check_shape(v, (self.n, 3), "velocities")
        self.ctx.check(self.ctx.lib.pmx_md_set_state(self.h, _dp(x), _dp(v)), "pmx_md_set_state")

    def md_run(self, nsteps: int):
        self.ctx.check(self.ctx.lib.pmx_md_run(self.h, nsteps), "pmx_md_run")

    def time_pair(self, reps: int = 100) -> float:
        """Average device ms of one pair-kernel launch at the last positions."""
        ms = C.c_float()
        self.ctx.check(self.ctx.lib.pmx_sys_time_pair(self.h, reps, C.byref(ms)), "pmx_sys_time_pair")
        return float(ms.value)

    def set_tuning(self, knob: int, value: int):
        """Sets a launch shape (TUNE_PAIR_SCHEDULE, TUNE_FFT_LINES,
        TUNE_PAIR_BLOCKS_PER_CU, TUNE_SPREAD_STREAM, TUNE_LIST_PRUNE,
        TUNE_FIXUP_STREAM or TUNE_FUSED_SPREAD);
        the resident state is kept."""
        self.ctx.check(self.ctx.lib.pmx_sys_set_tuning(self.h, knob, value), "pmx_sys_set_tuning")

    def md_get_state(self):
        """(x, v, energies dict) after the last run."""
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        e = np.zeros(5)
        self.ctx.check(self.ctx.lib.pmx_md_get_state(self.h, _dp(x), _dp(v), _dp(e)), "pmx_md_get_state")
        return x, v, dict(zip(("coulomb", "lj", "bond", "angle", "kinetic"), e.tolist()))

    def md_stats(self):
        st, rb, rl = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self.ctx.check(self.ctx.lib.pmx_md_stats(self.h, C.byref(st), C.byref(rb), C.byref(rl)), "pmx_md_stats")
        return dict(steps=st.value, rebuilds=rb.value, rollbacks=rl.value)

    def timing_begin(self):
        self.ctx.check(self.ctx.lib.pmx_sys_timing_begin(self.h), "pmx_sys_timing_begin")

    def timing_end(self):
        """(stage -> average ms, number of timed calls)."""
        ms = (C.c_float * len(STAGES))()
        n = C.c_int()
        self.ctx.check(self.ctx.lib.pmx_sys_timing_end(self.h, ms, C.byref(n)), "pmx_sys_timing_end")
        return dict(zip(STAGES, [float(v) for v in ms])), n.value

    def close(self):
        if self.h:
            self.ctx.lib.pmx_sys_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
