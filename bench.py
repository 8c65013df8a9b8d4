#!/usr/bin/env python3
"""Benchmark: optimal solves/sec of the MI355X transportation-simplex solver.

Default workload (one "step" = one pass of the hot path over one batch):
  batched -- 4096 independent balanced instances, m and n each log-uniform in
             [16, 512], costs U{1..100}, supplies/demands U{1..1000} balanced, all
             solved to certified optimality in one persistent work-queue launch per
             engine configuration; inputs are resident in HBM before the timed region.

The JSON line also carries, for the same metric:
  latency_ms   -- p50/p99 of per-instance completion time inside a batch launch;
  plain_gpu    -- the plain solver (flock_solve: one instance per launch, host arrays in
                  and out) on a sample of the same instances, and the batched/plain ratio;
  served       -- (--served K) K requests of the Poisson stream (m,n log-uniform
                  [32,2000], costs U{1..1000}, supplies/demands U{1..10000}) through the
                  continuous-batching server: solves/s and p50/p99 arrival->result latency;
  roofline     -- algorithmic bytes of the dominant kernel per launch over its device
                  time (HIP events on the library stream) against the 8 TB/s HBM peak;
  cpu_baseline -- the CPU oracle (the reference's sequential path restated, one thread)
                  timed on a bounded sample of the batch, objectives checked vs the GPU.

Other workloads (--workload): plain1000 (one 1000x1000 instance per step through the
plain solver), degenerate (one 4000x2000 instance with costs in {1..8}), served (the
request stream alone, --count requests).
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
PHASES = ["vogel", "tree", "sweep", "reprice", "cycle", "reroot"]


def algorithmic_bytes(m: int, n: int, sweeps: int) -> int:
    """HBM bytes the batched engine cannot avoid for one instance: the transpose (int32
    matrix read + written), Vogel's first pass over both orientations (two matrix reads)
    and every full reduced-cost sweep (one matrix read plus the m+n duals)."""
    mn = m * n
    return 8 * mn + 8 * mn + sweeps * (4 * mn + 4 * (m + n))


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def work_order(insts):
    return sorted(range(len(insts)), key=lambda k: -(insts[k].m * insts[k].n * (insts[k].m + insts[k].n)))


def cpu_baseline(insts, gpu_objs, budget_s: float, what: str):
    """The oracle (port of the reference's sequential path), one thread, bounded sample."""
    import oracle
    t0 = time.perf_counter()
    done, mism, solve_s = 0, 0, 0.0
    for inst, g in zip(insts, gpu_objs):
        s = oracle.solve(inst.cost, inst.supply, inst.demand)
        solve_s += s.solve_ms / 1000.0
        done += 1
        mism += int(s.objective != g)
        if time.perf_counter() - t0 > budget_s:
            break
    return {"value": done / solve_s if solve_s > 0 else None, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{done} of {len(insts)} {what}; solve time only (Vogel + simplex loop), one thread",
            "objective_mismatches": mism}


def plain_gpu(insts, pricing, budget_s):
    """flock_solve per instance (upload, one launch, download), wall clock per solve."""
    t0 = time.perf_counter()
    lat, objs = [], []
    for inst in insts:
        a = time.perf_counter()
        s = F.solve(inst, pricing=pricing)
        lat.append((time.perf_counter() - a) * 1000.0)
        objs.append(s.objective)
        if time.perf_counter() - t0 > budget_s:
            break
    wall = time.perf_counter() - t0
    return {"value": len(objs) / wall, "unit": "solves/s", "sample": len(objs), "p50_ms": pct(lat, 50),
            "p99_ms": pct(lat, 99)}, objs


def served(count, rate, seed, pricing, budget_bytes, slots, slice_us, plain_sample=0):
    """Drive the server with Poisson arrivals in real time; latency = arrival -> result."""
    stream = list(I.stream(seed=seed, count=count, rate=rate))
    srv = F.Server(hbm_budget_bytes=budget_bytes, max_slots=slots, pricing=pricing, slice_us=slice_us)
    arrive, done_t, objs, ticket_of = {}, {}, {}, {}
    t0 = time.perf_counter()
    nxt = 0
    while len(done_t) < count:
        now = time.perf_counter() - t0
        while nxt < count and stream[nxt][0] <= now:
            tk = srv.submit(stream[nxt][1])
            arrive[tk] = t0 + stream[nxt][0]
            ticket_of[nxt] = tk
            nxt += 1
        for tk, s in srv.poll(max_wait_ms=0.0 if nxt < count else 2.0):
            done_t[tk] = time.perf_counter()
            objs[tk] = (s.objective, s.status)
    wall = time.perf_counter() - t0
    srv.close()
    lat = [(done_t[t] - arrive[t]) * 1000.0 for t in done_t]
    res = {"value": count / wall, "unit": "solves/s", "requests": count, "rate_per_s": rate,
           "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "failed": sum(1 for o in objs.values() if o[1] != 0),
           "wall_s": wall, "hbm_budget_gb": budget_bytes / 2 ** 30, "slots": slots, "slice_us": slice_us}
    if plain_sample:
        p, pobjs = plain_gpu([x for _, x in stream[:plain_sample]], pricing, 60.0)
        p["objective_mismatches_vs_served"] = sum(1 for k, o in enumerate(pobjs) if o != objs[ticket_of[k]][0])
        res["plain_gpu"] = p
        res["served_over_plain"] = res["value"] / p["value"]
    return res


def copy_calibration():
    """A plain device-to-device copy of known traffic (read 256 MiB, write 256 MiB; the
    runtime runs it as its copy kernel), in the same process so that a hardware-counter
    run (FETCH_SIZE / WRITE_SIZE) can scale the solver kernel's counters by what they
    read for this one."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        raise RuntimeError("hipMalloc failed")
    rc = hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # device to device
    rc |= hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc:
        raise RuntimeError("calibration copy failed")
    return {"kernel": "hipMemcpy device to device (runtime copy kernel)", "read_bytes": n, "write_bytes": n}


def measured_traffic(kernel_key):
    """HBM bytes per launch of the dominant kernel from the committed counter summary
    (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by the copy kernel's
    counts), or None when it has not been measured."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get(kernel_key)


def bench_batched(args):
    insts = I.batch(seed=args.seed, count=args.count)
    tuned = None
    if args.minor_limit > 0:
        F.set_tuning(minor_limit=args.minor_limit)
    if not args.no_autotune:  # the project's autotuner on this batch, before the timed region
        t0 = time.perf_counter()
        tuning, tuned_ms = F.autotune(insts, pricing=args.pricing)
        tuned = {"tuning": tuning, "best_batch_ms": tuned_ms, "search_s": time.perf_counter() - t0}
    b = F.Batch(insts)
    for _ in range(args.warmup):
        b.run(pricing=args.pricing)
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        times.append(b.run(pricing=args.pricing))
    wall = time.perf_counter() - t0
    sols = b.results(full=False)
    b.close()
    objs = [s.objective for s in sols]
    ms_dev = float(np.mean(times))
    minor = F.get_tuning()["minor_limit"]
    alg = sum(algorithmic_bytes(x.m, x.n, s.sweeps) for x, s in zip(insts, sols))
    ph = np.sum([s.phase_us for s in sols], axis=0)
    achieved = alg / (ms_dev / 1000.0) / 1e9
    res = {
        "metric": METRIC, "value": len(insts) * args.steps / wall, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1000.0 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batched: {len(insts)} independent instances per launch, m,n log-uniform "
                               f"[16,512], costs U{{1..100}}, supplies/demands U{{1..1000}} balanced",
                   "count": len(insts), "seed": args.seed, "pricing": args.pricing, "minor_limit": minor},
        "device_ms_per_step": ms_dev,
        "autotune": tuned,
        "failed": sum(1 for s in sols if s.status != 0),
        "latency_ms": {"p50": pct([s.done_us / 1000.0 for s in sols], 50),
                       "p99": pct([s.done_us / 1000.0 for s in sols], 99),
                       "max": max(s.done_us for s in sols) / 1000.0,
                       "definition": "device time from the start of the batch launch to the instance's "
                                     "completion (last timed step)"},
        "pivots": int(sum(s.iterations for s in sols)), "sweeps": int(sum(s.sweeps for s in sols)),
        "phase_ms_sum": dict(zip(PHASES, [round(float(x) / 1000.0, 3) for x in ph])),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (measured_traffic(f"k_solve_queue batched{len(insts)} seed{args.seed} minor{minor}")
                                 if args.pricing == "candidates" else None),
                     "kernel": "k_solve_queue (workgroup per instance)", "algorithmic_bytes_per_launch": alg},
    }
    order = work_order(insts)
    if not args.no_plain:
        pick = order[::8]  # every 8th in work order keeps the batch's size mix
        p, pobjs = plain_gpu([insts[k] for k in pick], args.pricing, args.plain_budget)
        p["objective_mismatches_vs_batched"] = int(sum(a != objs[k] for a, k in zip(pobjs, pick)))
        res["plain_gpu"] = p
        res["batched_over_plain"] = res["value"] / p["value"]
    if args.served > 0:
        res["served"] = served(args.served, args.rate, args.seed, args.pricing, args.budget_gb << 30, args.slots,
                               args.slice_us, plain_sample=args.served_plain)
    if args.copy_calibration:
        res["calibration"] = copy_calibration()
    if not args.no_cpu:
        pick = order[::4]
        res["cpu_baseline"] = cpu_baseline([insts[k] for k in pick], [objs[k] for k in pick], args.cpu_budget,
                                           "instances (every 4th of the batch, largest first)")
    return res


def bench_single(args, inst, label):
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
        "config": {"workload": label, "seed": args.seed, "pricing": args.pricing},
        "device_ms_per_step": float(np.mean(devs)), "iterations": s.iterations, "sweeps": s.sweeps,
        "objective": s.objective, "status": s.status,
        "phase_ms": dict(zip(PHASES, [round(float(x) / 1000.0, 3) for x in s.phase_us])),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline([inst], [s.objective], args.cpu_budget, "instance")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain1000", "degenerate", "served"])
    ap.add_argument("--pricing", default="candidates")
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--minor-limit", type=int, default=0,
                    help="candidate-list minor iterations per sweep (0: the library default; autotune overrides)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run an elementwise kernel of known traffic (for hardware-counter runs)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--plain-budget", type=float, default=6.0)
    ap.add_argument("--served", type=int, default=0, help="requests of the served stream to run after the batch")
    ap.add_argument("--served-plain", type=int, default=0, help="stream requests also solved by the plain solver")
    ap.add_argument("--rate", type=float, default=400.0)
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--slots", type=int, default=2048)
    ap.add_argument("--slice-us", type=int, default=4000, help="server time slice")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): instances are independent, so the "
              f"path would shard across GPUs without communication; no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        res = bench_batched(args)
    elif args.workload == "plain1000":
        res = bench_single(args, I.plain_large(args.seed), "plain1000: one 1000x1000 instance per launch, "
                           "costs U{1..1000}, supplies/demands U{1..10000} balanced")
    elif args.workload == "degenerate":
        res = bench_single(args, I.degenerate(args.seed), "degenerate: one 4000x2000 instance per launch, "
                           "costs in {1..8}, supplies 50, demands 100")
    else:
        s = served(args.count, args.rate, args.seed, args.pricing, args.budget_gb << 30, args.slots,
                   args.slice_us, plain_sample=args.served_plain)
        res = {"metric": METRIC, "value": s["value"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
               "ms_per_step": s["wall_s"] * 1000.0, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int32", "data": "synthetic",
               "config": {"workload": f"served: {args.count} Poisson requests at {args.rate}/s, m,n log-uniform "
                                      f"[32,2000], costs U{{1..1000}}, supplies/demands U{{1..10000}}",
                          "seed": args.seed, "pricing": args.pricing}, "served": s}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
