"""Regenerate the golden fixtures under tests/golden/ from the reference checkout.

Run in a container where the reference is mounted at /root/reference:

    python tests/golden/make_golden.py

It copies (gzip) the reference's own instance files -- every file of its tests/ and
data/ folders -- and writes expected.json with, per instance:
  * optimum: the exact integer optimum of the integer-scaled instance, computed here
    with scipy's HiGHS LP solver (an independent solver, not the reference and not the
    oracle; the transportation LP is totally unimodular, so its optimum is integral);
  * gurobi: the 10-digit optimum printed in the reference's results/gurobi log;
  * seq_log: the 6-significant-digit objective printed in results/pure_sequencial.
Generated instances (the 1000x1000 workload, seed 1) get their optimum recorded the same
way without the file.  --only-missing keeps the entries already in expected.json.
The reference's sources are not copied; only its data files and logged numbers.
"""
from __future__ import annotations

import gzip
import json
import os
import re
import sys

import numpy as np
from scipy.optimize import linprog
from scipy.sparse import coo_matrix

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from flock_amd.instances import plain_large, read_dat  # noqa: E402

REF = "/root/reference"
TEST_FILES = sorted(f for f in os.listdir(os.path.join(REF, "tests")) if f.endswith(".dat")) \
    if os.path.isdir(os.path.join(REF, "tests")) else []
DATA_FILES = sorted(f for f in os.listdir(os.path.join(REF, "data")) if f.endswith(".dat")) \
    if os.path.isdir(os.path.join(REF, "data")) else []
# generated instances whose exact optimum is recorded without the file (the generator
# rebuilds them): the 1000 x 1000 single-instance workload, seed 1
GENERATED = {"plain1000_s1": lambda: plain_large(1)}


def lp_optimum(inst) -> int:
    m, n = inst.m, inst.n
    rows = np.concatenate([np.repeat(np.arange(m), n), m + np.tile(np.arange(n), m)])
    cols = np.concatenate([np.arange(m * n), np.arange(m * n)])
    A = coo_matrix((np.ones(2 * m * n), (rows, cols)), shape=(m + n, m * n)).tocsr()
    b = np.concatenate([inst.supply, inst.demand]).astype(float)
    r = linprog(inst.cost.ravel().astype(float), A_eq=A, b_eq=b, bounds=(0, None), method="highs")
    assert r.status == 0, r.message
    x = np.rint(r.x).astype(np.int64)
    assert np.abs(r.x - x).max() < 1e-6, "LP vertex not integral"
    return int((x * inst.cost.ravel().astype(np.int64)).sum())


def parse_log(path: str, key: str) -> dict[str, str]:
    out, cur = {}, None
    with open(path) as fh:
        for line in fh:
            mfile = re.search(r"File Name = \S*/([^/\s]+\.dat)", line)
            if mfile:
                cur = mfile.group(1)
            mval = re.search(key, line)
            if mval and cur and cur not in out:
                out[cur] = mval.group(1)
    return out


def main() -> None:
    only_missing = "--only-missing" in sys.argv
    outdir = os.path.join(HERE, "dat")
    os.makedirs(outdir, exist_ok=True)
    gurobi = parse_log(os.path.join(REF, "results", "gurobi"), r"Optimal objective\s+(\S+)")
    seq = parse_log(os.path.join(REF, "results", "pure_sequencial"), r"BASIC STATISTICS \| Objective: (\S+) \|")
    path = os.path.join(HERE, "expected.json")
    expected = json.load(open(path)) if only_missing and os.path.exists(path) else {}
    for sub, files in (("tests", TEST_FILES), ("data", DATA_FILES)):
        for f in files:
            if f in expected and os.path.exists(os.path.join(outdir, f + ".gz")):
                continue
            src = os.path.join(REF, sub, f)
            with open(src, "rb") as fh:
                raw = fh.read()
            with gzip.open(os.path.join(outdir, f + ".gz"), "wb", compresslevel=9) as gz:
                gz.write(raw)
            inst = read_dat(src)
            opt = lp_optimum(inst)
            expected[f] = {"origin": f"{sub}/{f}", "m": inst.m, "n": inst.n, "scale": inst.scale,
                           "optimum": opt, "gurobi": gurobi.get(f), "seq_log": seq.get(f)}
            print(f, inst.m, inst.n, inst.scale, opt, gurobi.get(f), seq.get(f), flush=True)
            json.dump(expected, open(path, "w"), indent=1, sort_keys=True)
    for name, make in GENERATED.items():
        if name in expected:
            continue
        inst = make()
        opt = lp_optimum(inst)
        expected[name] = {"origin": "generated", "m": inst.m, "n": inst.n, "scale": 0, "optimum": opt,
                          "gurobi": None, "seq_log": None}
        print(name, inst.m, inst.n, opt, flush=True)
        json.dump(expected, open(path, "w"), indent=1, sort_keys=True)
    json.dump(expected, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
