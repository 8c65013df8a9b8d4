"""The C++ drop-in on the GPU: a caller built with g++ -std=c++14 against include/flock
(tests/cpp/dropin_caller.cpp, built by build()) reads golden `.dat` files through
InputParser/make_problem, copy-initialises uvModel_parallel / ssModel_parallel /
switchModel_parallel, runs execute() and create_flows(), and must report the exact
optimum of every file (objVal in the file's units, objective_scaled exactly) with a
flow list that balances every supply and demand."""
import gzip
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
EXPECTED = json.load(open(os.path.join(GOLD, "expected.json")))
CALLER = os.path.join(HERE, "cpp", "_build", "dropin_caller")
FILES = ["test_simple1.dat", "test_simple1_degenerate.dat", "TransportModel_100_100_1_equalityConstr.dat",
         "TransportModel_250_250_1_equalityConstr.dat"]


@pytest.mark.parametrize("name", FILES)
@pytest.mark.parametrize("algo", ["parallel_uv", "parallel_ss", "switch_hybrid"])
def test_dropin_caller_reaches_golden_optimum(tmp_path, name, algo):
    assert os.path.exists(CALLER), "tests/cpp/_build/dropin_caller missing: run build()"
    path = str(tmp_path / name)
    with gzip.open(os.path.join(GOLD, "dat", name + ".gz"), "rb") as src, open(path, "wb") as dst:
        dst.write(src.read())
    print("dropin", algo, name, flush=True)
    p = subprocess.run([CALLER, "-i", path, "-a", algo], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT")][0]
    f = dict(zip(line.split()[2::2], line.split()[3::2]))
    e = EXPECTED[name]
    assert int(f["scaled"]) == e["optimum"] and int(f["scale"]) == e["scale"], (line, e)
    assert abs(float(f["objVal"]) - e["optimum"] / 10 ** e["scale"]) <= 1e-9 * e["optimum"] / 10 ** e["scale"]
    assert f["flows_ok"] == "1" and int(f["iterations"]) >= 1, line
    assert re.search(r"\[info\] => File Name = ", p.stdout), p.stdout[-500:]
