"""Host-side checks that need no GPU: the native library loads and exports every entry
point the C header declares, input validation fails loudly before touching a device,
the float-cost adapter is exact, and the instance generators/readers are correct."""
import ctypes
import os
import re

import numpy as np
import pytest

import flock_amd as F
from flock_amd import instances as I

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "flock_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(flock_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = F.lib()
    syms = declared_symbols()
    assert len(syms) >= 22 and set(F.EXPORTS) <= set(syms)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_cli_binary_built():
    assert os.path.exists(os.path.join(ROOT, "flock_amd", "lib", "flock"))


def test_dropin_caller_built_with_cpp14():
    """tests/cpp/dropin_caller.cpp copy-initialises every model class (as the reference's
    main.cpp:78,104,129 does), which compiles under -std=c++14 only with a move
    constructor; build() makes it and the GPU tests run it."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(ROOT, "tests", "cpp", "_build", "dropin_caller"))


REF_MAIN = "/root/reference/src/main.cpp"


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="needs the reference checkout (read where it lies)")
def test_reference_main_compiles_unchanged(tmp_path):
    """The reference's own driver, unchanged, compiles and links against include/flock and
    libflock_amd.so under its own standard (-std=c++14, Makefile_without_gurobi:1).  The
    file is compiled from a temporary copy so that its quote-includes resolve to this
    library's headers instead of the reference's CUDA ones; nothing is kept."""
    import shutil
    import subprocess
    src = str(tmp_path / "main.cpp")
    shutil.copyfile(REF_MAIN, src)
    out = str(tmp_path / "flock_ref_main")
    r = subprocess.run(["g++", "-std=c++14", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include", "flock"), src,
                        "-L", os.path.join(ROOT, "flock_amd", "lib"), "-lflock_amd", "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.getsize(out) > 0


def _inst(cost, supply, demand):
    return I.Instance(np.asarray(cost, np.int32), np.asarray(supply, np.int32), np.asarray(demand, np.int32))


@pytest.mark.parametrize("case,status", [
    (([[1, 2], [3, 4]], [1, 2], [2, 2]), F.E_UNBALANCED),
    (([[1, 2], [3, 4]], [-1, 4], [1, 2]), F.E_INVALID),
    (([[1, -2], [3, 4]], [1, 2], [1, 2]), F.E_RANGE),
    (([[2 ** 29, 2], [3, 4]], [1, 2], [1, 2]), F.E_RANGE),
])
def test_validation_errors_before_any_device_call(case, status):
    with pytest.raises(F.FlockError) as ei:
        F.solve(_inst(*case))
    assert ei.value.status == status


def test_too_large_rejected():
    """Plain solves take m+n up to 16383 (the tree in global memory); batches, the server
    and the allocation-free solve keep the tree in LDS (m+n up to 8000).  Both limits
    are checked before any device call."""
    inst = _inst(np.ones((1, 16383), np.int32), [16383], np.ones(16383, np.int32))
    with pytest.raises(F.FlockError) as ei:
        F.solve(inst)
    assert ei.value.status == F.E_TOO_LARGE
    inst = _inst(np.ones((1, 8000), np.int32), [8000], np.ones(8000, np.int32))
    with pytest.raises(F.FlockError) as ei:
        F.Batch([inst])
    assert ei.value.status == F.E_TOO_LARGE


def test_float_cost_adapter_exact():
    L = F.lib()
    costs = np.array([123.45, 0.01, 999.99, 100.0], np.float32)
    out = np.zeros(4, np.int32)
    k = ctypes.c_int32()
    st = L.flock_scale_float_costs(4, costs.ctypes.data, out.ctypes.data, ctypes.byref(k))
    assert st == 0 and k.value == 2 and out.tolist() == [12345, 1, 99999, 10000]
    costs = np.array([7.0, 3.0], np.float32)
    assert L.flock_scale_float_costs(2, costs.ctypes.data, out.ctypes.data, ctypes.byref(k)) == 0 and k.value == 0
    costs = np.array([1.0 / 3.0], np.float32)   # no decimal with <= 6 digits maps to this float
    assert L.flock_scale_float_costs(1, costs.ctypes.data, out.ctypes.data, ctypes.byref(k)) == F.E_NONINTEGRAL
    # the Python reader uses the same rule
    c, kk = I.scale_float32_costs(np.array([123.45, 0.01], np.float32))
    assert kk == 2 and c.tolist() == [12345, 1]


def test_generators_balanced_and_sized():
    for inst in (I.plain_small(1), I.plain_small(2)):
        assert inst.cost.shape == (100, 100)
        assert inst.cost.min() >= 1 and inst.cost.max() <= 100
        assert inst.supply.sum() == inst.demand.sum() and inst.demand.min() >= 1
    big = I.plain_large(1)
    assert big.cost.shape == (1000, 1000) and big.cost.max() <= 1000 and big.supply.sum() == big.demand.sum()
    d = I.degenerate(1, 40, 20)
    assert set(np.unique(d.cost)) <= set(range(1, 9)) and (d.supply == 50).all() and (d.demand == 100).all()
    b = I.batch(seed=3, count=64)
    ms = np.array([x.m for x in b]); ns = np.array([x.n for x in b])
    assert ms.min() >= 16 and ms.max() <= 512 and ns.min() >= 16 and ns.max() <= 512
    assert all(x.supply.sum() == x.demand.sum() for x in b)
    st = list(I.stream(seed=1, count=50, rate=100.0))
    ts = [t for t, _ in st]
    assert ts == sorted(ts) and all(32 <= x.m <= 2000 and 32 <= x.n <= 2000 for _, x in st)


def test_dat_roundtrip(tmp_path):
    inst = I.plain_small(4)
    p = str(tmp_path / "x.dat")
    I.write_dat(inst, p)
    back = I.read_dat(p)
    assert np.array_equal(back.cost, inst.cost) and np.array_equal(back.supply, inst.supply)
    assert np.array_equal(back.demand, inst.demand) and back.scale == 0
    dec = I.Instance(inst.cost, inst.supply, inst.demand, scale=2)
    I.write_dat(dec, p)
    back = I.read_dat(p)
    assert back.scale <= 2 and np.array_equal(back.cost * 10 ** (2 - back.scale), inst.cost)
