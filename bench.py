#!/usr/bin/env python3
"""Benchmark: optimal solves/sec of the MI355X transportation-simplex solver.

Default workload (one "step" = one pass of the hot path over one batch):
  batched -- 4096 independent balanced instances, m and n each log-uniform in
             [16, 512], costs U{1..100}, supplies/demands U{1..1000} balanced; all
             solved to certified optimality in one persistent work-queue launch per
             kernel size class, inputs resident in HBM before the timed region.
Other workloads (--workload): plain1000 (one 1000x1000 instance per step through the
plain solver), served (the Poisson request stream through the continuous-batching
server).

Prints one JSON line.  `value` is whole-job solves/s; `roofline` is measured for the
dominant kernel with HIP events on the library stream; `cpu_baseline` times the CPU
oracle (single thread, the reference's sequential algorithm restated) on a bounded
sample of the same workload and checks its objectives against the GPU's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import flock_amd as F  # noqa: E402
from flock_amd import instances as I  # noqa: E402

METRIC = "optimal solves/sec + p50/p99 latency: served/batched vs plain GPU, obj bit-exact"
HBM_PEAK_GBS = 8000.0


def sweep_bytes(m: int, n: int) -> int:
    """Algorithmic bytes of one reduced-cost sweep: int32 cost per cell + the duals."""
    return 4 * m * n + 4 * (m + n)


def cpu_baseline(insts, gpu_objs, budget_s: float = 12.0):
    """Oracle (port of the reference's sequential path), one thread, bounded sample."""
    import oracle
    t0 = time.perf_counter()
    done, mism = 0, 0
    solve_s = 0.0
    for inst, g in zip(insts, gpu_objs):
        s = oracle.solve(inst.cost, inst.supply, inst.demand)
        solve_s += s.solve_ms / 1000.0
        done += 1
        if s.objective != g:
            mism += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return {"value": done / solve_s if solve_s > 0 else None, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of {len(insts)} instances of the same workload, solve time only "
                      f"(std::chrono around Vogel + simplex loop)",
            "objective_mismatches": mism}


def bench_batched(args):
    insts = I.batch(seed=args.seed, count=args.count)
    b = F.Batch(insts)
    pricing = args.pricing
    for _ in range(args.warmup):
        b.run(pricing=pricing)
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        times.append(b.run(pricing=pricing))
    wall = time.perf_counter() - t0
    sols = b.results(full=True)
    b.close()
    bad = [k for k, s in enumerate(sols) if s.status != 0]
    objs = [s.objective for s in sols]
    ms_dev = float(np.mean(times))
    alg = sum(sweep_bytes(x.m, x.n) * s.sweeps for x, s in zip(insts, sols))
    ph = np.sum([s.phase_us for s in sols], axis=0)
    res_extra = {"pivots": int(sum(s.iterations for s in sols)), "sweeps": int(sum(s.sweeps for s in sols)),
                 "phase_ms_sum": dict(zip(["vogel", "tree", "sweep", "reprice", "cycle", "reroot"],
                                          [round(float(x) / 1000.0, 3) for x in ph]))}
    res = {
        "metric": METRIC, "value": len(insts) * args.steps / wall, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batched: {len(insts)} instances, m,n log-uniform [16,512], costs U{{1..100}}, "
                               f"supplies/demands U{{1..1000}} balanced", "count": len(insts), "seed": args.seed,
                   "pricing": pricing},
        "device_ms_per_step": ms_dev,
        "failed": len(bad),
        **res_extra,
    }
    if alg:
        achieved = alg / (ms_dev / 1000.0) / 1e9
        res["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(insts, objs, args.cpu_budget)
    return res


def bench_plain(args):
    inst = I.plain_large(args.seed)
    for _ in range(args.warmup):
        F.solve(inst, pricing=args.pricing)
    t0 = time.perf_counter()
    devs = []
    for _ in range(args.steps):
        s = F.solve(inst, pricing=args.pricing)
        devs.append(s.solve_ms)
    wall = time.perf_counter() - t0
    res = {
        "metric": METRIC, "value": args.steps / wall, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1000.0 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "plain1000: one 1000x1000 instance per launch, costs U{1..1000}, "
                               "supplies/demands U{1..10000} balanced", "seed": args.seed, "pricing": args.pricing},
        "device_ms_per_step": float(np.mean(devs)), "iterations": s.iterations, "objective": s.objective,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline([inst], [s.objective], args.cpu_budget)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain1000"])
    ap.add_argument("--pricing", default="dantzig")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the solver path here is "
              f"single-GPU and instances shard across GPUs without communication.", file=sys.stderr)
        sys.exit(2)
    res = bench_batched(args) if args.workload == "batched" else bench_plain(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
