"""Regenerate the golden fixtures under tests/golden/ from the reference checkout.

Run in a container where the reference is mounted at /root/reference:

    python tests/golden/make_golden.py

It copies (gzip) the reference's own instance files that the tests use -- the small
files of its tests/ folder and four benchmark instances of its data/ folder -- and
writes expected.json with, per instance:
  * optimum: the exact integer optimum of the integer-scaled instance, computed here
    with scipy's HiGHS LP solver (an independent solver, not the reference and not the
    oracle; the transportation LP is totally unimodular, so its optimum is integral);
  * gurobi: the 10-digit optimum printed in the reference's results/gurobi log;
  * seq_log: the 6-significant-digit objective printed in results/pure_sequencial.
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
from flock_amd.instances import read_dat  # noqa: E402

REF = "/root/reference"
TEST_FILES = ["test_simple1.dat", "test_simple1_degenerate.dat", "test_simple2.dat", "test_simple3.dat",
              "test_simple4.dat", "test_TransportModel_5_5_0.dat", "test_TransportModel_5_10_1.dat",
              "test_TransportModel_5_25_1.dat", "test_TransportModel_10_10_0.dat",
              "test_TransportModel_25_25_1.dat", "test_TransportModel_25_50_1.dat",
              "test_TransportModel_50_50_1.dat", "test_TransportModel_100_100_0.dat",
              "test_TransportModel_100_100_1.dat", "test_TransportModel_100_100_2.dat",
              "test_TransportModel_100_100_3.dat", "test_TransportModel_100_100_4.dat",
              "test_TransportModel_100_100_5.dat", "test_TransportModel_100_100_6.dat"]
DATA_FILES = ["TransportModel_100_100_1_equalityConstr.dat", "TransportModel_100_100_2_equalityConstr.dat",
              "TransportModel_100_100_3_equalityConstr.dat", "TransportModel_250_250_1_equalityConstr.dat"]


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
    outdir = os.path.join(HERE, "dat")
    os.makedirs(outdir, exist_ok=True)
    gurobi = parse_log(os.path.join(REF, "results", "gurobi"), r"Optimal objective\s+(\S+)")
    seq = parse_log(os.path.join(REF, "results", "pure_sequencial"), r"BASIC STATISTICS \| Objective: (\S+) \|")
    expected = {}
    for sub, files in (("tests", TEST_FILES), ("data", DATA_FILES)):
        for f in files:
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
    with open(os.path.join(HERE, "expected.json"), "w") as fh:
        json.dump(expected, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
