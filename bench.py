#!/usr/bin/env python3
"""Benchmark: optimal solves/sec of the MI355X transportation-simplex solver.

Headline workload (one "step" = one pass of the hot path over one batch):
  batched -- 4096 independent balanced instances, m and n each log-uniform in
             [16, 512], costs U{1..100}, supplies/demands U{1..1000} balanced, all
             solved to optimality in one persistent work-queue launch per engine
             configuration; inputs are resident in HBM before the timed region.  After
             the timed region every one of the 4096 bases is certified on the CPU
             (exact primal sums, an (m+n-1)-arc spanning tree, all reduced costs >= 0).

The JSON line also carries, for the same metric:
  latency_ms   -- p50/p99 of per-instance completion time inside a batch launch;
  plain_gpu    -- the plain solver (flock_solve: one instance per launch, host arrays in
                  and out) on a sample of the same instances, and the batched/plain ratio;
  served       -- the request stream (Poisson arrivals at --rate, m,n log-uniform
                  [32,2000], costs U{1..1000}, supplies/demands U{1..10000}) through the
                  continuous-batching server: solves/s, p50/p99 arrival->result latency,
                  every result certified, objectives of a sample checked against the
                  oracle, and the plain solver on the same stream;
  plain        -- single instances through flock_solve: 100x100 (costs U{1..100}),
                  1000x1000 (costs U{1..1000}) and the degenerate 4000x2000 (costs in
                  {1..8}, supplies 50, demands 100), each with its own CPU baseline; the
                  1000x1000 entry carries the sustained sweep rate over 1000 back-to-back
                  sweeps and the empty-grid launch floor;
  roofline     -- algorithmic HBM bytes of the batched launch (Vogel's first pass 4*m*n
                  plus 4*m*n + 4*(m+n) per full sweep, per instance) over its device time
                  (HIP events on the library stream) against the 8 TB/s HBM peak; the
                  design's extra traffic (device transpose, second Vogel orientation) is
                  reported beside it;
  cpu_baseline -- the CPU oracle (the reference's sequential path restated) timed on a
                  bounded sample of the batch, one thread and all host cores.

Other workloads (--workload): plain1000, degenerate, served (each alone).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import flock_amd as F  # noqa: E402
from flock_amd import instances as I  # noqa: E402

METRIC = "optimal solves/sec + p50/p99 latency: served/batched vs plain GPU, obj bit-exact"
HBM_PEAK_GBS = 8000.0
PHASES = ["vogel", "tree", "sweep", "reprice", "cycle", "reroot"]


def algorithmic_bytes(m: int, n: int, sweeps: int) -> int:
    """HBM bytes the algorithm must move for one instance: Vogel's first penalty pass
    reads the int32 matrix once (4*m*n); every full reduced-cost sweep reads the matrix
    and the m+n duals (4*m*n + 4*(m+n)).  Re-pricing a sweep's candidates reads no
    matrix cells (their costs are kept in registers)."""
    mn = m * n
    return 4 * mn + sweeps * (4 * mn + 4 * (m + n))


def design_bytes(m: int, n: int) -> int:
    """Traffic this design adds per instance: the device transpose (read and write of
    the matrix) and Vogel's first pass over the transposed copy."""
    return 8 * m * n + 4 * m * n


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def host_cpus():
    """Threads this process may use (the GPU box grants 16 per GPU) and the CPU model."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": aff, "threads_used": min(16, aff), "model": model}


def work_order(insts):
    return sorted(range(len(insts)), key=lambda k: -(insts[k].m * insts[k].n * (insts[k].m + insts[k].n)))


def cpu_baseline(insts, gpu_objs, budget_s: float, what: str, all_cores: bool = True):
    """The oracle (port of the reference's sequential path) on a bounded sample: one
    thread (solve time summed, as the reference's totalSolveTime), then one instance per
    thread on all host cores (wall clock), objectives checked against the GPU's."""
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
    res = {"value": done / solve_s if solve_s > 0 else None, "unit": "solves/s", "cores": 1, "kind": "port",
           "sample": f"{done} of {len(insts)} {what}; solve time only (Vogel + simplex loop), one thread",
           "objective_mismatches": mism}
    if all_cores and len(insts) > 1:
        cpus = host_cpus()
        nt = cpus["threads_used"]
        lock = threading.Lock()
        state = {"next": 0, "done": 0, "mism": 0}
        stop = time.perf_counter() + budget_s

        def worker():
            while time.perf_counter() < stop:
                with lock:
                    k = state["next"]
                    if k >= len(insts):
                        return
                    state["next"] += 1
                s = oracle.solve(insts[k].cost, insts[k].supply, insts[k].demand)
                with lock:
                    state["done"] += 1
                    state["mism"] += int(s.objective != gpu_objs[k])

        a = time.perf_counter()
        th = [threading.Thread(target=worker) for _ in range(nt)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        wall = time.perf_counter() - a
        res["all_cores"] = {"value": state["done"] / wall, "unit": "solves/s", "cores": nt,
                            "sample": f"{state['done']} instances, one per thread, wall clock", "host": cpus,
                            "objective_mismatches": state["mism"]}
    return res


def certify_all(insts, sols, threads=8):
    """LP-duality certificate of every returned basis (CPU, after the timed region)."""
    import oracle

    def one(k):
        x, s = insts[k], sols[k]
        bad, obj = oracle.certify(x.cost, x.supply, x.demand, s.rows, s.cols, s.flows.astype(np.int64))
        return bad == 0 and obj == s.objective and s.status == 0

    with ThreadPoolExecutor(max_workers=threads) as pool:
        ok = list(pool.map(one, range(len(insts))))
    return {"count": len(ok), "passed": int(sum(ok)),
            "checks": "exact row/column sums, (m+n-1)-arc spanning tree, all reduced costs >= 0, objective"}


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


def served(count, rate, seed, pricing, budget_bytes, slots, slice_us, plain_budget=10.0, oracle_budget=15.0,
           cpu_budget=10.0, no_cpu=False, minor_limit=0):
    """Drive the server with Poisson arrivals in real time; latency = arrival -> result.
    Requests are built by a thread pool ahead of their arrival (each from its own seed)
    and dropped once their result is certified, so the host holds only the requests in
    flight.  Every result is certified; a sample is checked against the oracle."""
    import oracle
    times, ms, ns = I.stream_plan(seed=seed, count=count, rate=rate)
    build = lambda k: I.stream_request(seed, k, ms[k], ns[k])  # noqa: E731
    gen = ThreadPoolExecutor(max_workers=6)
    cert = ThreadPoolExecutor(max_workers=6)
    ahead = 192
    futs = {k: gen.submit(build, k) for k in range(min(ahead, count))}
    sample = set(range(0, count, 50))   # oracle objective checks: every 50th request
    srv = F.Server(hbm_budget_bytes=budget_bytes, max_slots=slots, pricing=pricing, slice_us=slice_us,
                   minor_limit=minor_limit)
    arrive, done_t, inflight, certs, samples, obj_of = {}, {}, {}, [], [], {}
    failed = 0

    def certify_one(inst, sol):
        bad, obj = oracle.certify(inst.cost, inst.supply, inst.demand, sol.rows, sol.cols, sol.flows.astype(np.int64))
        return bad == 0 and obj == sol.objective

    t0 = time.perf_counter()
    nxt = 0
    while len(done_t) < count:
        now = time.perf_counter() - t0
        while nxt < count and times[nxt] <= now:
            inst = futs.pop(nxt).result()
            if nxt + ahead < count:
                futs[nxt + ahead] = gen.submit(build, nxt + ahead)
            tk = srv.submit(inst)
            arrive[tk] = t0 + times[nxt]
            inflight[tk] = (nxt, inst)
            nxt += 1
        for tk, s in srv.poll(max_wait_ms=0.0 if nxt < count else 2.0, cap=256, full=True):
            done_t[tk] = time.perf_counter()
            k, inst = inflight.pop(tk)
            failed += int(s.status != 0)
            certs.append(cert.submit(certify_one, inst, s))
            obj_of[k] = s.objective
            if k in sample:
                samples.append((k, s.objective))
    wall = time.perf_counter() - t0
    srv.close()
    certified = sum(1 for f in certs if f.result())
    gen.shutdown()
    cert.shutdown()
    lat = [(done_t[t] - arrive[t]) * 1000.0 for t in done_t]
    res = {"value": count / wall, "unit": "solves/s", "requests": count, "rate_per_s": rate,
           "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": max(lat), "failed": failed,
           "certified": certified, "wall_s": wall, "hbm_budget_gb": budget_bytes / 2 ** 30, "slots": slots,
           "slice_us": slice_us, "pricing": pricing,
           "minor_limit": minor_limit if minor_limit > 0 else F.get_tuning()["minor_limit"],
           "latency_definition": "host wall clock from the request's scheduled arrival to its result in poll"}
    # oracle objectives on a sample (every 50th request, in order, within a time budget)
    stop = time.perf_counter() + oracle_budget
    mism, checked = 0, 0
    with ThreadPoolExecutor(max_workers=host_cpus()["threads_used"]) as pool:
        def one(ko):
            k, obj = ko
            if time.perf_counter() > stop:
                return None
            x = build(k)
            return oracle.solve(x.cost, x.supply, x.demand).objective == obj
        for r in pool.map(one, sorted(samples)):
            if r is not None:
                checked += 1
                mism += int(not r)
    res["oracle_sample"] = {"checked": checked, "objective_mismatches": mism,
                            "sample": "every 50th request, in arrival order, until the time budget"}
    # the plain solver (one instance per launch) on the same stream, back to back
    plain_insts = (build(k) for k in range(count))
    t0 = time.perf_counter()
    lat_p, n_p, pm = [], 0, 0
    for k, x in enumerate(plain_insts):
        a = time.perf_counter()
        s = F.solve(x, pricing=pricing)
        lat_p.append((time.perf_counter() - a) * 1000.0)
        n_p += 1
        pm += int(s.objective != obj_of[k])
        if time.perf_counter() - t0 > plain_budget:
            break
    wp = time.perf_counter() - t0
    res["plain_gpu"] = {"value": n_p / wp, "unit": "solves/s", "sample": n_p, "p50_ms": pct(lat_p, 50),
                        "p99_ms": pct(lat_p, 99), "objective_mismatches_vs_served": pm,
                        "definition": "flock_solve on the stream's first requests, one after another"}
    res["served_over_plain"] = res["value"] / res["plain_gpu"]["value"]
    if not no_cpu:
        ks = list(range(0, count, 97))[:400]
        res["cpu_baseline"] = cpu_baseline([build(k) for k in ks], [obj_of[k] for k in ks], cpu_budget,
                                           "stream requests (every 97th)")
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
    counts), or None when it has not been measured for this configuration."""
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
    if args.prio_permille >= 0:
        F.set_tuning(prio_permille=args.prio_permille)
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
    sols = b.results(full=True)
    b.close()
    objs = [s.objective for s in sols]
    ms_dev = float(np.mean(times))
    minor = F.get_tuning()["minor_limit"]
    alg = sum(algorithmic_bytes(x.m, x.n, s.sweeps) for x, s in zip(insts, sols))
    des = sum(design_bytes(x.m, x.n) for x in insts)
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
        "certified": certify_all(insts, sols),
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
                     "kernel": "k_solve_queue (workgroup per instance)", "kernel_ms": ms_dev,
                     "algorithmic_bytes_per_launch": alg,
                     "design_bytes_per_launch": des,
                     "algorithmic_definition": "per instance 4*m*n (Vogel's first pass) + sweeps*(4*m*n + 4*(m+n))",
                     "design_definition": "per instance 8*m*n (device transpose) + 4*m*n (Vogel over the "
                                          "transposed copy), not counted as algorithmic"},
    }
    order = work_order(insts)
    if not args.no_plain:
        pick = order[::8]  # every 8th in work order keeps the batch's size mix
        p, pobjs = plain_gpu([insts[k] for k in pick], args.pricing, args.plain_budget)
        p["objective_mismatches_vs_batched"] = int(sum(a != objs[k] for a, k in zip(pobjs, pick)))
        res["plain_gpu"] = p
        res["batched_over_plain"] = res["value"] / p["value"]
    if args.copy_calibration:
        res["calibration"] = copy_calibration()
    if not args.no_cpu:
        pick = order[::4]
        res["cpu_baseline"] = cpu_baseline([insts[k] for k in pick], [objs[k] for k in pick], args.cpu_budget,
                                           "instances (every 4th of the batch, largest first)")
    return res


def plain_case(args, inst, label, steps, cpu_budget, sweep_reps=0, cpu_thread=None, tune_large=False):
    """flock_solve on one instance: solves/s over `steps` solves, the result certified,
    the oracle's objective, and the CPU baseline (one thread).  tune_large: the
    large-instance path's autotuner picks the sweep tile shape first (outside the timing)."""
    import oracle
    tuned = None
    if tune_large and not args.no_autotune:
        before = F.get_tuning()
        t, ms = F.autotune_large(inst, pricing=args.pricing)
        tuned = {"sweep_vectors": t["sweep_vectors"], "large_threads": t["large_threads"],
                 "large_minor_limit": t["large_minor_limit"], "best_solve_ms": ms}
    F.solve(inst, pricing=args.pricing)  # warm-up (graph capture, first-use set-up)
    devs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = F.solve(inst, pricing=args.pricing)
        devs.append(s.solve_ms)
    wall = time.perf_counter() - t0
    bad, obj = oracle.certify(inst.cost, inst.supply, inst.demand, s.rows, s.cols, s.flows.astype(np.int64))
    if tuned is not None:
        F.set_tuning(**before)
    res = {"workload": label, "value": steps / wall, "unit": "solves/s", "ms_per_solve": wall * 1000.0 / steps,
           "autotune_large": tuned,
           "device_ms": float(np.mean(devs)), "objective": s.objective, "status": s.status,
           "certified": bad == 0 and obj == s.objective, "pivots": s.iterations, "sweeps": s.sweeps,
           "phase_ms": dict(zip(PHASES, [round(float(x) / 1000.0, 3) for x in s.phase_us]))}
    if sweep_reps:  # the sweep grid alone, back to back, at every tile shape
        base = F.get_tuning()
        shapes = []
        for v in (1, 2, 4, 8):
            F.set_tuning(sweep_vectors=v)
            ms = F.measure_sweep(inst, sweep_reps)
            ms["sweep_vectors"] = v
            ms["sustained_gbs"] = ms["bytes_per_sweep"] / (ms["sweep_ms"] / 1000.0) / 1e9
            ms["frac_of_hbm_peak"] = ms["sustained_gbs"] / HBM_PEAK_GBS
            shapes.append(ms)
        F.set_tuning(**base)
        best = max(shapes, key=lambda x: x["sustained_gbs"])
        res["sweep"] = dict(best, by_tile_shape=[{k: x[k] for k in ("sweep_vectors", "grid", "sweep_ms", "empty_ms",
                                                                      "sustained_gbs")} for x in shapes],
                            definition="HIP events around `reps` back-to-back launches of the sweep grid (gaps "
                                       "included), best tile shape; empty_ms: an empty kernel with the same grid")
    if cpu_thread is not None:
        res["cpu_baseline"] = cpu_thread
    elif not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline([inst], [s.objective], cpu_budget, "instance", all_cores=False)
    return res


def plain_suite(args):
    """The three single-instance configurations.  The degenerate instance's CPU solve
    (tens of seconds: Vogel rescans) runs on a host thread while the GPU parts run."""
    import oracle
    deg = I.degenerate(args.seed)
    out = {}
    c3_cpu = {}

    def c3_oracle():
        s = oracle.solve(deg.cost, deg.supply, deg.demand)
        c3_cpu.update(s=s)

    th = None
    if not args.no_cpu:
        th = threading.Thread(target=c3_oracle)
        th.start()
    small = [I.plain_small(args.seed + k) for k in range(40)]
    r = plain_case(args, small[0], "one 100x100 instance per launch, costs U{1..100}, supplies/demands "
                   "U{1..1000} balanced", 40, 0, cpu_thread={})
    if not args.no_cpu:
        objs = [F.solve(x, pricing=args.pricing).objective for x in small]
        r["cpu_baseline"] = cpu_baseline(small, objs, 3.0, "100x100 instances (seeds in order)", all_cores=False)
    out["plain100"] = r
    print("plain 100x100 done", file=sys.stderr, flush=True)
    out["plain1000"] = plain_case(args, I.plain_large(args.seed), "one 1000x1000 instance per launch, costs "
                                  "U{1..1000}, supplies/demands U{1..10000} balanced", 5, 20.0, sweep_reps=1000,
                                  tune_large=True)
    print("plain 1000x1000 done", file=sys.stderr, flush=True)
    r = plain_case(args, deg, "one 4000x2000 instance per launch, costs in {1..8}, supplies 50, demands 100", 5, 0,
                   cpu_thread={}, tune_large=True)
    print("plain 4000x2000 done", file=sys.stderr, flush=True)
    if th is not None:
        while th.is_alive():  # the CPU solve of the degenerate instance
            th.join(timeout=30.0)
            print("waiting for the CPU solve of the 4000x2000 instance", file=sys.stderr, flush=True)
        s = c3_cpu["s"]
        r["cpu_baseline"] = {"value": 1000.0 / s.solve_ms, "unit": "solves/s", "cores": 1, "kind": "port",
                             "sample": "the instance, one solve, solve time only (Vogel + simplex loop), one thread "
                                       "(run on a host thread while the GPU cases ran)",
                             "objective_mismatches": int(s.objective != r["objective"]),
                             "vogel_ms": s.vogel_ms, "pivots": s.iterations}
    out["degenerate4000x2000"] = r
    return out


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
    ap.add_argument("--no-singles", action="store_true", help="skip the single-instance configurations")
    ap.add_argument("--minor-limit", type=int, default=0,
                    help="candidate-list minor iterations per sweep (0: the library default; autotune overrides)")
    ap.add_argument("--prio-permille", type=int, default=-1,
                    help="raised wave priority for this share of the batch's largest instances (-1: library default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a device copy of known traffic (for hardware-counter runs)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--plain-budget", type=float, default=6.0)
    ap.add_argument("--served", type=int, default=20000, help="requests of the served stream (0: skip)")
    ap.add_argument("--rate", type=float, default=1600.0, help="served stream arrival rate per second")
    ap.add_argument("--budget-gb", type=int, default=64)
    ap.add_argument("--slots", type=int, default=2048)
    ap.add_argument("--slice-us", type=int, default=4000, help="server time slice")
    ap.add_argument("--served-minor-limit", type=int, default=128,
                    help="server's candidate-pricing minor limit (0: the library's tuning)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): instances are independent, so the "
              f"path would shard across GPUs without communication; no multi-GPU figure is produced.",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        res = bench_batched(args)
        print(f"batched done: {res['value']:.0f} solves/s", file=sys.stderr, flush=True)
        if args.served > 0:
            res["served"] = served(args.served, args.rate, args.seed, args.pricing, args.budget_gb << 30, args.slots,
                                   args.slice_us, no_cpu=args.no_cpu, minor_limit=args.served_minor_limit)
            print(f"served done: {res['served']['value']:.0f} solves/s", file=sys.stderr, flush=True)
        if not args.no_singles:
            res["plain"] = plain_suite(args)
            print("single instances done", file=sys.stderr, flush=True)
    elif args.workload in ("plain1000", "degenerate"):
        inst = I.plain_large(args.seed) if args.workload == "plain1000" else I.degenerate(args.seed)
        label = ("one 1000x1000 instance per launch, costs U{1..1000}, supplies/demands U{1..10000} balanced"
                 if args.workload == "plain1000" else
                 "one 4000x2000 instance per launch, costs in {1..8}, supplies 50, demands 100")
        r = plain_case(args, inst, label, args.steps, args.cpu_budget,
                       sweep_reps=1000 if args.workload == "plain1000" else 0, tune_large=True)
        res = {"metric": METRIC, "value": r["value"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
               "warmup": 1, "ms_per_step": r["ms_per_solve"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int32", "data": "synthetic",
               "config": {"workload": label, "seed": args.seed, "pricing": args.pricing}, **{"result": r}}
    else:
        s = served(args.count, args.rate, args.seed, args.pricing, args.budget_gb << 30, args.slots,
                   args.slice_us, no_cpu=args.no_cpu, minor_limit=args.served_minor_limit)
        res = {"metric": METRIC, "value": s["value"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
               "ms_per_step": s["wall_s"] * 1000.0, "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "int32", "data": "synthetic",
               "config": {"workload": f"served: {args.count} Poisson requests at {args.rate}/s, m,n log-uniform "
                                      f"[32,2000], costs U{{1..1000}}, supplies/demands U{{1..10000}}",
                          "seed": args.seed, "pricing": args.pricing}, "served": s}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
