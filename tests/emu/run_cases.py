"""Run seeded instances through the CPU emulation of the device engine and print one
JSON line per case (test infrastructure; started by tests/test_engine_emulated.py in a
subprocess so a sanitizer runtime can be preloaded)."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import emu  # noqa: E402
from flock_amd import instances as I  # noqa: E402


def main():
    variant = sys.argv[1]
    cases = json.loads(sys.argv[2])
    L = emu.load_variant(variant)
    for c in cases:
        rng = np.random.default_rng(c["seed"])
        if c.get("degenerate"):
            inst = I.degenerate(c["seed"], c["m"], c["n"])
        else:
            inst = I.random_instance(rng, c["m"], c["n"], c.get("cmax", 50), c.get("smax", 100))
        if c.get("large"):  # the large-instance kernel sequence
            r = emu.large(L, inst.cost, inst.supply, inst.demand, pricing=c.get("pricing", 0), vpt=c.get("vpt", 0),
                          tree_global=c.get("tree_global", 0))
            r["steps"] = r["rounds"]
        else:
            r = emu.solve(L, inst.cost, inst.supply, inst.demand, group=c["group"], pricing=c.get("pricing", 0),
                          initial=c.get("initial", 0), step_pivots=c.get("steps", 0))
        print(json.dumps({"case": c, "objective": r["objective"], "iterations": r["iterations"],
                          "status": r["status"], "steps": r["steps"], "rows": r["rows"].tolist(),
                          "cols": r["cols"].tolist(), "flows": r["flows"].tolist(), "u": r["u"].tolist(),
                          "v": r["v"].tolist()}), flush=True)


if __name__ == "__main__":
    main()
