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
    L.emu_large.argtypes = [ctypes.c_int, ctypes.c_int, i32p, i32p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            i32p, i32p, i32p, i32p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int32)]
    L.emu_group_bytes.restype = ctypes.c_int64
    L.emu_group_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64]
    return L


def queue(L, insts, T=256, G=256, wgs=2, pricing=0):
    """One emulated launch of the work-queue kernel: `wgs` workgroups running together
    over `insts`.  Returns (ret, objectives, pivots, statuses)."""
    n = len(insts)
    cap = 160 * 1024 // (T // G)
    region = max(L.emu_group_bytes(x.m, x.n, cap) for x in insts)
    keep = [(np.ascontiguousarray(x.cost, np.int32), np.ascontiguousarray(x.supply, np.int32),
             np.ascontiguousarray(x.demand, np.int32)) for x in insts]
    P = ctypes.POINTER(ctypes.c_int32)
    ms = (ctypes.c_int32 * n)(*[x.m for x in insts])
    ns = (ctypes.c_int32 * n)(*[x.n for x in insts])
    costs = (P * n)(*[k[0].ctypes.data_as(P) for k in keep])
    sups = (P * n)(*[k[1].ctypes.data_as(P) for k in keep])
    dems = (P * n)(*[k[2].ctypes.data_as(P) for k in keep])
    obj = (ctypes.c_int64 * n)()
    it = (ctypes.c_int64 * n)()
    st = (ctypes.c_int32 * n)()
    r = L.emu_queue(n, ms, ns, costs, sups, dems, T, G, wgs, pricing, ctypes.c_int64(region), obj, it, st)
    return r, list(obj), list(it), list(st)


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


def large(L, cost, supply, demand, pricing=0, vpt=0, tree_global=0):
    """The large-instance kernel sequence (setup, sweep tiles, rounds) emulated;
    tree_global 1 / 2 puts the tree (and with 2 also Vogel's line state) in global
    memory, as the runtime does for instances whose tree does not fit LDS.
    pricing -1 stops after the set-up kernel (rows/cols/flows are Vogel's basis)."""
    cost = np.ascontiguousarray(cost, np.int32); m, n = cost.shape
    V = m + n
    br = np.zeros(V - 1, np.int32); bc = np.zeros(V - 1, np.int32); bf = np.zeros(V - 1, np.int32)
    du = np.zeros(V, np.int32)
    obj, it, st = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
    rounds = L.emu_large(m, n, cost.ravel(), np.ascontiguousarray(supply, np.int32),
                         np.ascontiguousarray(demand, np.int32), pricing, vpt, tree_global, br, bc, bf, du, ctypes.byref(obj),
                         ctypes.byref(it), ctypes.byref(st))
    return dict(rounds=rounds, objective=obj.value, iterations=it.value, status=st.value, rows=br, cols=bc, flows=bf,
                u=du[:m], v=du[m:])
