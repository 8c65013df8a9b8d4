"""ctypes wrapper of the CPU emulation of the device engine (test infrastructure)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def build(targets=("_build/libflock_emu.so",)):
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def load(asan: bool = False):
    return load_variant("asan" if asan else "plain")


def load_variant(variant: str):
    name = {"plain": "libflock_emu.so", "asan": "libflock_emu_asan.so", "tsan": "libflock_emu_tsan.so"}[variant]
    L = ctypes.CDLL(os.path.join(HERE, "_build", name))
    i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.emu_solve.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int64, ctypes.c_int, i32p, i32p, i32p, i32p,
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    return L


def solve(L, cost, supply, demand, group=64, groups_per_wg=4, pricing=0, initial=0, max_it=0, step_pivots=0):
    cost = np.ascontiguousarray(cost, np.int32); m, n = cost.shape
    V = m + n
    br = np.zeros(V - 1, np.int32); bc = np.zeros(V - 1, np.int32); bf = np.zeros(V - 1, np.int32)
    du = np.zeros(V, np.int32)
    obj, it = ctypes.c_int64(), ctypes.c_int64()
    st, steps = ctypes.c_int32(), ctypes.c_int32()
    r = L.emu_solve(m, n, cost.ravel(), np.ascontiguousarray(supply, np.int32), np.ascontiguousarray(demand, np.int32),
                    group, groups_per_wg, pricing, initial, max_it, step_pivots, br, bc, bf, du,
                    ctypes.byref(obj), ctypes.byref(it), ctypes.byref(st), ctypes.byref(steps))
    return dict(ret=r, objective=obj.value, iterations=it.value, status=st.value, steps=steps.value,
                rows=br, cols=bc, flows=bf, u=du[:m], v=du[m:])
