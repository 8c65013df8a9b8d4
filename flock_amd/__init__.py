"""flock_amd -- MI355X-native balanced transportation-problem solver.

Python binding of the C ABI in include/flock_amd.h (libflock_amd.so, built in-tree by
flock_amd/build.py).  The native library is required: importing works without it,
but every solve raises if the library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import instances  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libflock_amd.so")

OK, E_ITERATION_LIMIT, E_FOREST, E_UNBALANCED, E_INVALID, E_RANGE, E_TOO_LARGE, E_DEVICE, \
    E_NONINTEGRAL, E_BUDGET, E_BUSY = range(11)
STATUS_NAMES = ["ok", "iteration_limit", "forest", "unbalanced", "invalid", "range", "too_large",
                "device", "nonintegral", "budget", "busy"]

PRICING = {"dantzig": 0, "candidates": 1, "multi": 2, "auto": 3, "switch": 4}
INITIAL = {"vogel": 0, "nwc": 1}


class FlockError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else status}: {msg}")
        self.status = status


class _Options(ctypes.Structure):
    _fields_ = [("initial", ctypes.c_int32), ("pricing", ctypes.c_int32),
                ("max_iterations", ctypes.c_int64), ("threads", ctypes.c_int32),
                ("team", ctypes.c_int32), ("reserved", ctypes.c_int32 * 8)]


class _Instance(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int32), ("n", ctypes.c_int32),
                ("cost", ctypes.c_void_p), ("supply", ctypes.c_void_p), ("demand", ctypes.c_void_p)]


class _Flow(ctypes.Structure):
    _fields_ = [("source", ctypes.c_int32), ("destination", ctypes.c_int32), ("qty", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("status", ctypes.c_int32), ("n_flows", ctypes.c_int32), ("solve_ms", ctypes.c_double),
                ("flows", ctypes.c_void_p), ("basis_row", ctypes.c_void_p), ("basis_col", ctypes.c_void_p),
                ("basis_flow", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("sweeps", ctypes.c_int64), ("phase_us", ctypes.c_double * 6),
                ("start_us", ctypes.c_double), ("done_us", ctypes.c_double)]


class _Tuning(ctypes.Structure):
    _fields_ = [("wave_max_v", ctypes.c_int32), ("wave_pack", ctypes.c_int32), ("wg256_max_v", ctypes.c_int32),
                ("minor_limit", ctypes.c_int32), ("few_256_max_v", ctypes.c_int32),
                ("few_512_max_v", ctypes.c_int32), ("team_min_cells", ctypes.c_int32), ("graph_min_cells", ctypes.c_int32),
                ("prio_permille", ctypes.c_int32), ("sweep_vectors", ctypes.c_int32),
                ("large_threads", ctypes.c_int32), ("large_minor_limit", ctypes.c_int32)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_slots", ctypes.c_int32),
                ("step_pivots", ctypes.c_int32), ("slice_us", ctypes.c_int32), ("minor_limit", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 6)]


_lib = None

EXPORTS = ["flock_last_error", "flock_init", "flock_default_options", "flock_solve", "flock_solve_noalloc",
           "flock_batch_solve",
           "flock_batch_create", "flock_batch_run", "flock_batch_last_ms", "flock_batch_results",
           "flock_batch_free", "flock_server_create", "flock_server_submit", "flock_server_poll",
           "flock_server_pending", "flock_server_destroy", "flock_scale_float_costs", "flock_measure_sweep",
           "flock_autotune", "flock_autotune_large"]


def lib():
    """Load libflock_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlockError(E_DEVICE, f"native library missing: {LIB_PATH} (run flock_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.flock_last_error.restype = ctypes.c_char_p
    L.flock_init.argtypes = [ctypes.c_int]
    L.flock_default_options.argtypes = [ctypes.POINTER(_Options)]
    L.flock_default_options.restype = None
    L.flock_solve.argtypes = [ctypes.POINTER(_Instance), ctypes.POINTER(_Options), ctypes.POINTER(_Result)]
    L.flock_solve_noalloc.argtypes = [ctypes.POINTER(_Instance), ctypes.POINTER(_Options), ctypes.POINTER(_Result)]
    L.flock_batch_solve.argtypes = [i32, ctypes.POINTER(_Instance), ctypes.POINTER(_Options), ctypes.POINTER(_Result)]
    L.flock_batch_create.argtypes = [i32, ctypes.POINTER(_Instance), ctypes.POINTER(vp)]
    L.flock_batch_run.argtypes = [vp, ctypes.POINTER(_Options)]
    L.flock_batch_last_ms.argtypes = [vp]
    L.flock_batch_last_ms.restype = ctypes.c_double
    L.flock_batch_results.argtypes = [vp, ctypes.POINTER(_Result)]
    L.flock_batch_free.argtypes = [vp]
    L.flock_batch_free.restype = None
    if hasattr(L, "flock_server_create"):
        L.flock_server_create.argtypes = [ctypes.POINTER(_ServerConfig), ctypes.POINTER(_Options), ctypes.POINTER(vp)]
        L.flock_server_submit.argtypes = [vp, ctypes.POINTER(_Instance), ctypes.POINTER(i64)]
        L.flock_server_poll.argtypes = [vp, ctypes.c_double, i32, ctypes.POINTER(i64), ctypes.POINTER(_Result),
                                        ctypes.POINTER(i32)]
        L.flock_server_pending.argtypes = [vp]
        L.flock_server_pending.restype = i32
        L.flock_server_destroy.argtypes = [vp]
        L.flock_server_destroy.restype = None
    L.flock_scale_float_costs.argtypes = [i64, vp, vp, ctypes.POINTER(i32)]
    L.flock_measure_sweep.argtypes = [ctypes.POINTER(_Instance), i32, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
    L.flock_tuning_get.argtypes = [ctypes.POINTER(_Tuning)]
    L.flock_tuning_set.argtypes = [ctypes.POINTER(_Tuning)]
    L.flock_autotune.argtypes = [i32, ctypes.POINTER(_Instance), ctypes.POINTER(_Options), ctypes.POINTER(_Tuning),
                                 ctypes.POINTER(ctypes.c_double)]
    L.flock_autotune_large.argtypes = [ctypes.POINTER(_Instance), ctypes.POINTER(_Options), ctypes.POINTER(_Tuning),
                                       ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def get_tuning() -> dict:
    t = _Tuning()
    lib().flock_tuning_get(ctypes.byref(t))
    return {"wave_max_v": t.wave_max_v, "wave_pack": t.wave_pack, "wg256_max_v": t.wg256_max_v,
            "minor_limit": t.minor_limit, "few_256_max_v": t.few_256_max_v, "few_512_max_v": t.few_512_max_v,
            "team_min_cells": t.team_min_cells, "graph_min_cells": t.graph_min_cells,
            "prio_permille": t.prio_permille, "sweep_vectors": t.sweep_vectors, "large_threads": t.large_threads,
            "large_minor_limit": t.large_minor_limit}


def set_tuning(**kw) -> dict:
    cur = get_tuning()
    cur.update(kw)
    t = _Tuning(cur["wave_max_v"], cur["wave_pack"], cur["wg256_max_v"], cur["minor_limit"], cur["few_256_max_v"],
                cur["few_512_max_v"], cur["team_min_cells"], cur["graph_min_cells"], cur["prio_permille"],
                cur["sweep_vectors"], cur["large_threads"], cur["large_minor_limit"])
    st = lib().flock_tuning_set(ctypes.byref(t))
    if st:
        raise FlockError(st, last_error())
    return cur


def autotune(insts, pricing: str = "candidates") -> tuple[dict, float]:
    """flock_autotune on a sample of instances: installs and returns the fastest policy
    and its batch time in ms."""
    arr = (_Instance * len(insts))()
    keep = []
    for k, inst in enumerate(insts):
        ci, kp = _as_c(inst)
        arr[k] = ci
        keep.append(kp)
    t = _Tuning()
    ms = ctypes.c_double()
    o = options(pricing)
    st = lib().flock_autotune(len(insts), arr, ctypes.byref(o), ctypes.byref(t), ctypes.byref(ms))
    if st:
        raise FlockError(st, last_error())
    return get_tuning(), ms.value


def measure_sweep(inst, reps: int = 1000) -> dict:
    """flock_measure_sweep: `reps` back-to-back launches of the large-instance sweep grid
    over `inst`'s matrix, and of an empty kernel with the same grid."""
    ci, keep = _as_c(inst)
    sw, em, grid = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
    st = lib().flock_measure_sweep(ctypes.byref(ci), reps, ctypes.byref(sw), ctypes.byref(em), ctypes.byref(grid))
    if st:
        raise FlockError(st, last_error())
    return {"reps": reps, "sweep_ms": sw.value, "empty_ms": em.value, "grid": grid.value,
            "bytes_per_sweep": 4 * inst.m * inst.n + 4 * (inst.m + inst.n)}


def autotune_large(inst, pricing: str = "candidates") -> tuple[dict, float]:
    """flock_autotune_large on one instance of the large-instance path: installs and
    returns the fastest sweep tile shape and its solve time in ms."""
    ci, keep = _as_c(inst)
    t = _Tuning()
    ms = ctypes.c_double()
    o = options(pricing)
    st = lib().flock_autotune_large(ctypes.byref(ci), ctypes.byref(o), ctypes.byref(t), ctypes.byref(ms))
    if st:
        raise FlockError(st, last_error())
    return get_tuning(), ms.value


def last_error() -> str:
    return lib().flock_last_error().decode(errors="replace")


def options(pricing: str = "dantzig", initial: str = "vogel", max_iterations: int = 0,
            threads: int = 0, team: int = 0) -> _Options:
    o = _Options()
    lib().flock_default_options(ctypes.byref(o))
    o.pricing = PRICING[pricing]
    o.initial = INITIAL[initial]
    o.max_iterations = max_iterations
    o.threads = threads
    o.team = team
    return o


def _as_c(inst):
    cost = np.ascontiguousarray(inst.cost, dtype=np.int32)
    supply = np.ascontiguousarray(inst.supply, dtype=np.int32)
    demand = np.ascontiguousarray(inst.demand, dtype=np.int32)
    ci = _Instance(cost.shape[0], cost.shape[1], cost.ctypes.data, supply.ctypes.data, demand.ctypes.data)
    return ci, (cost, supply, demand)


@dataclass
class Solution:
    objective: int
    iterations: int
    status: int
    n_flows: int
    solve_ms: float
    rows: np.ndarray = field(repr=False)
    cols: np.ndarray = field(repr=False)
    flows: np.ndarray = field(repr=False)
    u: np.ndarray = field(repr=False)
    v: np.ndarray = field(repr=False)
    flow_list: np.ndarray = field(repr=False)   # (n_flows, 3): source, destination, qty
    sweeps: int = 0
    phase_us: tuple = ()
    start_us: float = 0.0
    done_us: float = 0.0

    @property
    def ok(self) -> bool:
        return self.status == OK


def _result_buffers(m: int, n: int):
    V = m + n
    bufs = dict(rows=np.zeros(V - 1, np.int32), cols=np.zeros(V - 1, np.int32), flows=np.zeros(V - 1, np.int32),
                u=np.zeros(m, np.int32), v=np.zeros(n, np.int32), fl=np.zeros((max(V - 1, 1), 3), np.int32))
    r = _Result()
    r.flows = bufs["fl"].ctypes.data
    r.basis_row = bufs["rows"].ctypes.data
    r.basis_col = bufs["cols"].ctypes.data
    r.basis_flow = bufs["flows"].ctypes.data
    r.u = bufs["u"].ctypes.data
    r.v = bufs["v"].ctypes.data
    return r, bufs


def _solution(r: _Result, b) -> Solution:
    return Solution(r.objective, r.iterations, r.status, r.n_flows, r.solve_ms, b["rows"], b["cols"], b["flows"],
                    b["u"], b["v"], b["fl"][:r.n_flows].copy(), r.sweeps,
                    tuple(r.phase_us), r.start_us, r.done_us)


def solve(inst, pricing: str = "dantzig", initial: str = "vogel", max_iterations: int = 0,
          threads: int = 0, team: int = 0, check: bool = True) -> Solution:
    """Plain solver: one instance per launch (flock_solve)."""
    ci, keep = _as_c(inst)
    r, b = _result_buffers(ci.m, ci.n)
    o = options(pricing, initial, max_iterations, threads, team)
    st = lib().flock_solve(ctypes.byref(ci), ctypes.byref(o), ctypes.byref(r))
    if check and st not in (OK, E_ITERATION_LIMIT, E_FOREST):
        raise FlockError(st, last_error())
    return _solution(r, b)


def solve_noalloc(inst, pricing: str = "dantzig", initial: str = "vogel", max_iterations: int = 0,
                  threads: int = 0, team: int = 0, check: bool = True) -> Solution:
    """flock_solve_noalloc: the plain solve with every device buffer in the library's
    fixed device array (no device allocation)."""
    ci, keep = _as_c(inst)
    r, b = _result_buffers(ci.m, ci.n)
    o = options(pricing, initial, max_iterations, threads, team)
    st = lib().flock_solve_noalloc(ctypes.byref(ci), ctypes.byref(o), ctypes.byref(r))
    if check and st not in (OK, E_ITERATION_LIMIT, E_FOREST):
        raise FlockError(st, last_error())
    return _solution(r, b)


class Batch:
    """Instances uploaded once and kept resident in HBM (flock_batch_*)."""

    def __init__(self, insts):
        self.insts = list(insts)
        arr = (_Instance * len(self.insts))()
        self._keep = []
        for k, inst in enumerate(self.insts):
            ci, keep = _as_c(inst)
            arr[k] = ci
            self._keep.append(keep)
        h = ctypes.c_void_p()
        st = lib().flock_batch_create(len(self.insts), arr, ctypes.byref(h))
        if st:
            raise FlockError(st, last_error())
        self._h = h

    def run(self, pricing: str = "dantzig", initial: str = "vogel", max_iterations: int = 0, threads: int = 0) -> float:
        o = options(pricing, initial, max_iterations, threads)
        st = lib().flock_batch_run(self._h, ctypes.byref(o))
        if st:
            raise FlockError(st, last_error())
        return lib().flock_batch_last_ms(self._h)

    def results(self, full: bool = True) -> list[Solution]:
        n = len(self.insts)
        res = (_Result * n)()
        bufs = []
        for k, inst in enumerate(self.insts):
            if full:
                r, b = _result_buffers(inst.m, inst.n)
            else:
                r, b = _Result(), None
            res[k] = r
            bufs.append(b)
        st = lib().flock_batch_results(self._h, res)
        if st:
            raise FlockError(st, last_error())
        out = []
        for k in range(n):
            if full:
                out.append(_solution(res[k], bufs[k]))
            else:
                r = res[k]
                out.append(Solution(r.objective, r.iterations, r.status, r.n_flows, r.solve_ms,
                                    None, None, None, None, None, None, r.sweeps, tuple(r.phase_us),
                                    r.start_us, r.done_us))
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().flock_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching solver (flock_server_*): submit instances at any time, poll
    for finished ones.  Results carry the objective, pivots, status and (optionally)
    basis, flows and duals."""

    def __init__(self, hbm_budget_bytes: int = 8 << 30, max_slots: int = 2048, step_pivots: int = 0,
                 pricing: str = "candidates", initial: str = "vogel", max_iterations: int = 0, slice_us: int = 0,
                 minor_limit: int = 0):
        cfg = _ServerConfig(hbm_budget_bytes, max_slots, step_pivots, slice_us, minor_limit)
        o = options(pricing, initial, max_iterations)
        h = ctypes.c_void_p()
        st = lib().flock_server_create(ctypes.byref(cfg), ctypes.byref(o), ctypes.byref(h))
        if st:
            raise FlockError(st, last_error())
        self._h = h
        self.sizes = {}
        self._bufs = {}

    def submit(self, inst) -> int:
        ci, keep = _as_c(inst)
        t = ctypes.c_int64()
        st = lib().flock_server_submit(self._h, ctypes.byref(ci), ctypes.byref(t))
        if st:
            raise FlockError(st, last_error())
        self.sizes[t.value] = (ci.m, ci.n)
        return t.value

    def _full_buffers(self, cap: int):
        """Result arrays for `cap` results, sized for the largest pending request and
        reused between polls (grown when a larger request is pending)."""
        mm = max((a for a, b in self.sizes.values()), default=1)
        nn = max((b for a, b in self.sizes.values()), default=1)
        fb = getattr(self, "_fb", None)
        if fb is None or fb["cap"] < cap or fb["m"] < mm or fb["n"] < nn:
            mm = max(mm, fb["m"] if fb else 1)
            nn = max(nn, fb["n"] if fb else 1)
            V = mm + nn
            fb = dict(cap=cap, m=mm, n=nn, tickets=(ctypes.c_int64 * cap)(), res=(_Result * cap)(),
                      rows=np.zeros((cap, V - 1), np.int32), cols=np.zeros((cap, V - 1), np.int32),
                      flows=np.zeros((cap, V - 1), np.int32), u=np.zeros((cap, mm), np.int32),
                      v=np.zeros((cap, nn), np.int32), fl=np.zeros((cap, max(V - 1, 1), 3), np.int32))
            for k in range(cap):
                r = fb["res"][k]
                r.flows = fb["fl"][k].ctypes.data
                r.basis_row = fb["rows"][k].ctypes.data
                r.basis_col = fb["cols"][k].ctypes.data
                r.basis_flow = fb["flows"][k].ctypes.data
                r.u = fb["u"][k].ctypes.data
                r.v = fb["v"][k].ctypes.data
            self._fb = fb
        return fb

    def poll(self, max_wait_ms: float = 0.0, cap: int = 4096, full: bool = False):
        """Advance the device and return [(ticket, Solution)] of finished requests.
        With full=True each Solution carries its basis (rows, cols, flows), duals and
        flow list."""
        if full:
            fb = self._full_buffers(cap)
            tickets, res = fb["tickets"], fb["res"]
        elif cap in self._bufs:  # reused between calls
            tickets, res = self._bufs[cap]
        else:
            tickets = (ctypes.c_int64 * cap)()
            res = (_Result * cap)()
            self._bufs[cap] = (tickets, res)
        nd = ctypes.c_int32()
        st = lib().flock_server_poll(self._h, ctypes.c_double(max_wait_ms), cap, tickets, res, ctypes.byref(nd))
        if st:
            raise FlockError(st, last_error())
        out = []
        for k in range(nd.value):
            r = res[k]
            t = tickets[k]
            m, n = self.sizes.pop(t)
            if full:
                V = m + n
                sol = Solution(r.objective, r.iterations, r.status, r.n_flows, r.solve_ms,
                               fb["rows"][k, :V - 1].copy(), fb["cols"][k, :V - 1].copy(),
                               fb["flows"][k, :V - 1].copy(), fb["u"][k, :m].copy(), fb["v"][k, :n].copy(),
                               fb["fl"][k, :r.n_flows].copy(), r.sweeps, tuple(r.phase_us))
            else:
                sol = Solution(r.objective, r.iterations, r.status, r.n_flows, r.solve_ms, None, None, None, None,
                               None, None, r.sweeps, tuple(r.phase_us))
            out.append((t, sol))
        return out

    def pending(self) -> int:
        return lib().flock_server_pending(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().flock_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_solve(insts, pricing: str = "dantzig", initial: str = "vogel", max_iterations: int = 0):
    b = Batch(insts)
    try:
        ms = b.run(pricing, initial, max_iterations)
        sols = b.results()
    finally:
        b.close()
    return sols, ms
