"""ctypes binding of the CPU oracle (oracle/transport_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, always as the checker.  The product package
``flock_amd`` never imports this module.

The oracle restates the reference's sequential UV / stepping-stone path
(``pure_sequencial``: src/parallel_simplex_method/IBFS_vogel.cu:546-769,
parallel_uv_method/DUAL_solver.cu:34-63, uv_model_parallel.cu:178-200,
PIVOT_uv.cu:156-308,318-401) in exact int64 arithmetic.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

STATUS = {0: "ok", 1: "iteration_limit", 2: "forest", 3: "unbalanced", 4: "invalid"}


class _Options(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int64),
                ("initial", ctypes.c_int32),
                ("record_limit", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [("objective", ctypes.c_int64),
                ("iterations", ctypes.c_int64),
                ("status", ctypes.c_int32),
                ("nonzero_flows", ctypes.c_int32),
                ("solve_ms", ctypes.c_double),
                ("vogel_ms", ctypes.c_double)]


def build() -> str:
    """Compile the oracle shared library (g++, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
        vp = ctypes.c_void_p
        L.oracle_vogel.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p, i32p, i64p]
        L.oracle_vogel.restype = ctypes.c_int
        L.oracle_nwc.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, i32p, i64p]
        L.oracle_nwc.restype = ctypes.c_int
        L.oracle_solve.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p,
                                   ctypes.POINTER(_Options), ctypes.POINTER(_Result),
                                   vp, vp, vp, vp, vp, vp, vp]
        L.oracle_solve.restype = ctypes.c_int
        L.oracle_certify.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, ctypes.c_int,
                                     i32p, i32p, i64p, vp, vp, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_certify.restype = ctypes.c_int
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


@dataclass
class OracleSolution:
    objective: int
    iterations: int
    status: int
    nonzero_flows: int
    solve_ms: float
    vogel_ms: float
    rows: np.ndarray
    cols: np.ndarray
    flows: np.ndarray
    u: np.ndarray
    v: np.ndarray
    trace_cell: np.ndarray
    trace_r: np.ndarray


def vogel(cost, supply, demand):
    cost = _c(cost, np.int32); m, n = cost.shape
    r = np.zeros(m + n - 1, np.int32); c = np.zeros(m + n - 1, np.int32); f = np.zeros(m + n - 1, np.int64)
    st = lib().oracle_vogel(m, n, cost.ravel(), _c(supply, np.int32), _c(demand, np.int32), r, c, f)
    if st:
        raise RuntimeError(f"oracle_vogel status {st}")
    return r, c, f


def nwc(supply, demand):
    supply = _c(supply, np.int32); demand = _c(demand, np.int32)
    m, n = len(supply), len(demand)
    r = np.zeros(m + n - 1, np.int32); c = np.zeros(m + n - 1, np.int32); f = np.zeros(m + n - 1, np.int64)
    st = lib().oracle_nwc(m, n, supply, demand, r, c, f)
    if st:
        raise RuntimeError(f"oracle_nwc status {st}")
    return r, c, f


def solve(cost, supply, demand, max_iterations: int = 0, initial: str = "vogel",
          record: int = 0) -> OracleSolution:
    cost = _c(cost, np.int32); m, n = cost.shape
    supply = _c(supply, np.int32); demand = _c(demand, np.int32)
    V = m + n
    rows = np.zeros(V - 1, np.int32); cols = np.zeros(V - 1, np.int32); flows = np.zeros(V - 1, np.int64)
    u = np.zeros(m, np.int64); v = np.zeros(n, np.int64)
    tc = np.zeros(max(record, 1), np.int64); tr = np.zeros(max(record, 1), np.int64)
    opt = _Options(max_iterations, 1 if initial == "nwc" else 0, record)
    res = _Result()
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib().oracle_solve(m, n, cost.ravel(), supply, demand, ctypes.byref(opt), ctypes.byref(res),
                       ptr(rows), ptr(cols), ptr(flows), ptr(u), ptr(v), ptr(tc), ptr(tr))
    k = min(record, res.iterations)
    return OracleSolution(res.objective, res.iterations, res.status, res.nonzero_flows,
                          res.solve_ms, res.vogel_ms, rows, cols, flows, u, v, tc[:k], tr[:k])


def certify(cost, supply, demand, rows, cols, flows, u=None, v=None):
    """Return (bitmask, objective); bitmask 0 means a certified optimal basis."""
    cost = _c(cost, np.int32); m, n = cost.shape
    rows = _c(rows, np.int32); cols = _c(cols, np.int32); flows = _c(flows, np.int64)
    obj = ctypes.c_int64(0)
    uu = _c(u, np.int64) if u is not None else None
    vv = _c(v, np.int64) if v is not None else None
    bad = lib().oracle_certify(m, n, cost.ravel(), _c(supply, np.int32), _c(demand, np.int32),
                               len(rows), rows, cols, flows,
                               uu.ctypes.data_as(ctypes.c_void_p) if uu is not None else None,
                               vv.ctypes.data_as(ctypes.c_void_p) if vv is not None else None,
                               ctypes.byref(obj))
    return bad, obj.value
