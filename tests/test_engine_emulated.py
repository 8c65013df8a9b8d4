"""The device engine (flock_amd/csrc/engine.h) compiled for the CPU and run lane by
lane on OS threads (tests/emu), checked against the oracle -- no GPU needed.

  * plain build: bit-exact pivot sequence vs the oracle for wavefront (64-lane) and
    workgroup (256-lane) groups, Dantzig and candidate pricing, NWC start, resumable
    step-bounded solves (the server's mode);
  * AddressSanitizer build: no out-of-bounds access of the engine's buffers;
  * ThreadSanitizer build: no data race between lanes (every cross-lane dependence
    goes through a group sync), and the emulator's call-site-checked barriers prove
    that all lanes of a group pass the same sequence of sync points.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from flock_amd import instances as I

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="host clang++ needed for the engine emulation")


def _runtime(kind):
    return subprocess.run([CLANG, f"-print-file-name=libclang_rt.{kind}-x86_64.so"], capture_output=True,
                          text=True).stdout.strip()


def run(variant, cases, timeout=600):
    target = {"plain": "_build/libflock_emu.so", "asan": "_build/libflock_emu_asan.so",
              "tsan": "_build/libflock_emu_tsan.so"}[variant]
    subprocess.run(["make", "-s", "-C", EMU, target], check=True)
    env = dict(os.environ)
    if variant == "asan":
        env["LD_PRELOAD"] = _runtime("asan")
        env["ASAN_OPTIONS"] = "detect_leaks=0"
    if variant == "tsan":
        env["LD_PRELOAD"] = _runtime("tsan")
        env["TSAN_OPTIONS"] = "report_signal_unsafe=0 halt_on_error=1 exitcode=66"
    p = subprocess.run([sys.executable, os.path.join(EMU, "run_cases.py"), variant, json.dumps(cases)],
                       capture_output=True, text=True, env=env, timeout=timeout)
    assert p.returncode == 0, p.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in p.stderr and "WARNING: ThreadSanitizer" not in p.stderr, p.stderr[-4000:]
    return [json.loads(l) for l in p.stdout.splitlines() if l.startswith("{")]


def instance_of(c):
    if c.get("degenerate"):
        return I.degenerate(c["seed"], c["m"], c["n"])
    return I.random_instance(np.random.default_rng(c["seed"]), c["m"], c["n"], c.get("cmax", 50), c.get("smax", 100))


def check(results, exact=True):
    for r in results:
        c = r["case"]
        inst = instance_of(c)
        o = oracle.solve(inst.cost, inst.supply, inst.demand, initial="nwc" if c.get("initial") else "vogel")
        assert r["status"] == 0, r
        assert r["objective"] == o.objective, (c, r["objective"], o.objective)
        bad, _ = oracle.certify(inst.cost, inst.supply, inst.demand, np.array(r["rows"], np.int32),
                                np.array(r["cols"], np.int32), np.array(r["flows"], np.int64),
                                np.array(r["u"], np.int64), np.array(r["v"], np.int64))
        assert bad == 0, (c, bad)
        if exact and c.get("pricing", 0) == 0:
            assert r["iterations"] == o.iterations, c
            assert r["rows"] == o.rows.tolist() and r["cols"] == o.cols.tolist(), c
            assert r["flows"] == o.flows.tolist(), c


def test_emulated_engine_matches_oracle():
    cases = []
    # 64-lane groups at moderate sizes; 256-lane groups (one OS thread per lane, so
    # slow to emulate) on small ones
    for g, shapes, size in ((64, [(1, 5), (6, 1), (4, 4), (17, 31), (40, 25), (70, 90)], 50),
                            (256, [(1, 5), (6, 1), (4, 4), (17, 31)], 16)):
        for k, (m, n) in enumerate(shapes):
            cases.append({"m": m, "n": n, "seed": 10 * k + g, "group": g})
        cases.append({"m": size + 10, "n": size + 20, "seed": 5 + g, "group": g, "pricing": 1})
        cases.append({"m": size + 12, "n": size + 3, "seed": 12 + g, "group": g, "pricing": 2})
        # the switch rule, also cut into steps (its mode survives a saved tree)
        cases.append({"m": size + 14, "n": size + 9, "seed": 13 + g, "group": g, "pricing": 4})
        cases.append({"m": size + 6, "n": size + 11, "seed": 14 + g, "group": g, "pricing": 4, "steps": 3})
        cases.append({"m": size, "n": size, "seed": 6 + g, "group": g, "initial": 1})
        cases.append({"m": size - 5, "n": size - 15, "seed": 7 + g, "group": g, "steps": 7})
        # time slices (negative steps: slice length in clock ticks), one pivot per call
        cases.append({"m": size - 3, "n": size - 8, "seed": 11 + g, "group": g, "steps": -1})
        cases.append({"m": 2 * (size // 2), "n": size // 2, "seed": 8 + g, "group": g, "degenerate": True})
    check(run("plain", cases))



def test_emulated_windowed_reroot_beyond_one_pass():
    """64-lane groups at m+n = 310 .. 420, above the 256 window slots of one pass: pivots
    whose re-root window fits take the windowed form, the others the pass over every
    vertex; Dantzig stays pivot for pivot the oracle's (also cut into steps), and the
    candidate rule reaches its optimum."""
    cases = [{"m": 150, "n": 160, "seed": 71, "group": 64},
             {"m": 180, "n": 240, "seed": 73, "group": 64, "pricing": 1},
             {"m": 160, "n": 150, "seed": 74, "group": 64, "steps": 5},
             {"m": 220, "n": 110, "seed": 75, "group": 64, "degenerate": True}]
    check(run("plain", cases))

def test_emulated_engine_address_sanitizer():
    cases = [{"m": 23, "n": 37, "seed": 1, "group": 64}, {"m": 13, "n": 11, "seed": 2, "group": 256},
             {"m": 30, "n": 30, "seed": 3, "group": 64, "pricing": 1, "steps": 5},
             {"m": 12, "n": 6, "seed": 4, "group": 256, "degenerate": True},
             {"m": 14, "n": 9, "seed": 5, "group": 64, "steps": -1},
             # the large-instance kernels with the tree (and Vogel's line state) in global memory
             {"m": 11, "n": 23, "seed": 6, "group": 256, "large": True, "vpt": 4, "tree_global": 2},
             {"m": 17, "n": 9, "seed": 7, "group": 256, "large": True, "vpt": 8, "tree_global": 1, "pricing": 1}]
    check(run("asan", cases))


def test_emulated_engine_thread_sanitizer():
    cases = [{"m": 6, "n": 9, "seed": 1, "group": 64}, {"m": 12, "n": 20, "seed": 2, "group": 256},
             {"m": 12, "n": 20, "seed": 3, "group": 256, "pricing": 1},
             {"m": 10, "n": 14, "seed": 4, "group": 64, "pricing": 1, "steps": 3},
             {"m": 13, "n": 16, "seed": 5, "group": 256, "pricing": 2},
             # 150 vertices on one wavefront: four window slots per rank in the re-root
             {"m": 70, "n": 80, "seed": 11, "group": 64, "pricing": 1}]
    check(run("tsan", cases))


def test_emulated_queue_concurrent_workgroups():
    """Three workgroups of 256 lanes run one launch together and pop instances from one
    queue head; every result equals the oracle's (Dantzig: pivot for pivot), for every
    pricing rule including the switch rule (multi-pivot rounds, then Dantzig)."""
    sys.path.insert(0, EMU)
    import emu
    subprocess.run(["make", "-s", "-C", EMU, "_build/libflock_emu.so"], check=True)
    L = emu.load_variant("plain")
    rng = np.random.default_rng(5)
    insts = [I.random_instance(rng, 100, 82, 50, 100), I.random_instance(rng, 12, 9, 50, 100),
             I.random_instance(rng, 20, 33, 50, 100), I.random_instance(rng, 47, 40, 50, 100)]
    for pricing in (0, 1, 2, 4):
        r, obj, it, st = emu.queue(L, insts, T=256, G=256, wgs=3, pricing=pricing)
        assert r == 0
        for x, o_, i_, s_ in zip(insts, obj, it, st):
            o = oracle.solve(x.cost, x.supply, x.demand)
            assert s_ == 0 and o_ == o.objective, (x.m, x.n, pricing, s_, o_, o.objective)
            if pricing == 0:
                assert i_ == o.iterations


def test_emulated_large_instance_kernel_sequence():
    """The large-instance path as separate kernels (setup; per round: one sweep kernel
    per matrix tile, then one pivot round), emulated with small tiles so that many
    tiles and candidate lists are exercised, with the tree in LDS and, as for instances
    whose tree does not fit LDS, in global memory (Vogel's line state in LDS or global):
    Dantzig equals the oracle pivot for pivot, the other rules reach the same optimum."""
    sys.path.insert(0, EMU)
    import emu
    subprocess.run(["make", "-s", "-C", EMU, "_build/libflock_emu.so"], check=True)
    L = emu.load_variant("plain")
    for m, n, vpt, tg in [(30, 37, 16, 0), (7, 90, 8, 0), (64, 3, 4, 0), (25, 25, 0, 0), (31, 29, 16, 1), (9, 70, 8, 2),
                          (40, 6, 4, 2)]:
        inst = I.random_instance(np.random.default_rng(m * n), m, n, 50, 100)
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        for pricing in (0, 1, 2, 4):
            r = emu.large(L, inst.cost, inst.supply, inst.demand, pricing=pricing, vpt=vpt, tree_global=tg)
            assert r["status"] == 0 and r["objective"] == o.objective, (m, n, pricing, r["objective"], o.objective)
            bad, _ = oracle.certify(inst.cost, inst.supply, inst.demand, r["rows"], r["cols"],
                                    r["flows"].astype(np.int64), r["u"].astype(np.int64), r["v"].astype(np.int64))
            assert bad == 0
            if pricing == 0:
                assert r["iterations"] == o.iterations and r["rounds"] == o.iterations + 1
                assert np.array_equal(r["rows"], o.rows) and np.array_equal(r["flows"], o.flows)
                assert np.array_equal(r["u"], o.u) and np.array_equal(r["v"], o.v)


def test_emulated_large_setup_vogel_register_modes():
    """Vogel in the large-instance set-up kernel (256-thread groups in the emulation)
    equals the oracle's basis slot by slot with each rank holding 8 lines in registers
    (m+n up to 2048 here), the tracked keys in LDS or in global scratch, and with the
    line state in global memory (no register lines)."""
    sys.path.insert(0, EMU)
    import emu
    subprocess.run(["make", "-s", "-C", EMU, "_build/libflock_emu.so"], check=True)
    L = emu.load_variant("plain")
    for m, n, tg in [(40, 1000, 0), (1030, 12, 1), (20, 1100, 2)]:
        inst = I.random_instance(np.random.default_rng(m * n), m, n, 50, 100000)
        vr, vc, vf = oracle.vogel(inst.cost, inst.supply, inst.demand)
        r = emu.large(L, inst.cost, inst.supply, inst.demand, pricing=-1, tree_global=tg)
        assert r["status"] == 0, (m, n, r["status"])
        assert np.array_equal(r["rows"], vr) and np.array_equal(r["cols"], vc), (m, n)
        assert np.array_equal(r["flows"], vf), (m, n)
