"""The served stream's own distribution through the continuous-batching server: requests
with m, n log-uniform in [32, 2000], costs U{1..1000}, supplies/demands U{1..10000}
balanced, plus four with m, n >= 1500.  Requests arrive in waves while earlier ones are
on the device.

  * candidate pricing (the server's default): every result's int64 objective equals
    the CPU oracle's and its basis passes the optimality certificate;
  * Dantzig pricing: every result equals the same request solved alone (flock_solve,
    which takes the graph path for the large ones) and the oracle, bit for bit --
    objective, pivot count, basis slot by slot, flows and duals.

The oracle runs in a thread pool while the device works (its ctypes calls release
the GIL)."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import flock_amd as F
import oracle
from flock_amd import instances as I

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def stream_requests(seed=7, count=64, big=4):
    reqs = [x for _, x in I.stream(seed=seed, count=count, rate=1000.0, lo=32, hi=2000, cmax=1000, smax=10000)]
    rng = np.random.default_rng(seed + 100)
    for k in range(big):
        m, n = (int(v) for v in rng.integers(1500, 2001, size=2))
        reqs.insert(8 + 12 * k, I.random_instance(rng, m, n, 1000, 10000, name=f"big{k}"))
    return reqs


@pytest.fixture(scope="module")
def requests_and_oracle():
    reqs = stream_requests()
    assert sum(1 for x in reqs if x.m >= 1500 and x.n >= 1500) >= 4 and len(reqs) >= 68
    pool = ThreadPoolExecutor(max_workers=8)
    # largest first, so the long oracle solves start early
    futs = {k: pool.submit(oracle.solve, x.cost, x.supply, x.demand)
            for k, x in sorted(enumerate(reqs), key=lambda kv: -kv[1].m * kv[1].n)}
    yield reqs, futs
    pool.shutdown(wait=True)


def serve(reqs, pricing):
    # the minor limit the bench's served stream uses (the server's own setting)
    srv = F.Server(hbm_budget_bytes=8 << 30, max_slots=256, pricing=pricing, minor_limit=128)
    got, tickets = {}, []
    for w in range(0, len(reqs), 12):  # waves of arrivals while earlier requests run
        tickets += [srv.submit(x) for x in reqs[w:w + 12]]
        for t, s in srv.poll(max_wait_ms=0.0, cap=64, full=True):
            got[t] = s
    polls = 0
    while len(got) < len(reqs) and polls < 200000:
        for t, s in srv.poll(max_wait_ms=20.0, cap=64, full=True):
            got[t] = s
        polls += 1
    srv.close()
    assert len(got) == len(reqs)
    return [got[t] for t in tickets]


def test_served_stream_candidates_objective_and_certificate(requests_and_oracle):
    reqs, futs = requests_and_oracle
    say("served stream, candidates:", len(reqs), "requests, largest", max(x.m * x.n for x in reqs), "cells")
    sols = serve(reqs, "candidates")
    for k, (x, s) in enumerate(zip(reqs, sols)):
        assert s.status == 0, (k, x.m, x.n, s.status)
        bad, obj = oracle.certify(x.cost, x.supply, x.demand, s.rows, s.cols, s.flows.astype(np.int64),
                                  s.u.astype(np.int64), s.v.astype(np.int64))
        assert bad == 0 and obj == s.objective, (k, x.m, x.n, bad)
    for k, (x, s) in enumerate(zip(reqs, sols)):
        o = futs[k].result(timeout=600)
        assert o.status == 0 and s.objective == o.objective, (k, x.m, x.n, s.objective, o.objective)


def test_served_stream_dantzig_alone_equals_mix(requests_and_oracle):
    reqs, futs = requests_and_oracle
    say("served stream, dantzig:", len(reqs), "requests")
    sols = serve(reqs, "dantzig")
    for k, (x, s) in enumerate(zip(reqs, sols)):
        o = futs[k].result(timeout=600)
        assert s.status == 0 and s.objective == o.objective and s.iterations == o.iterations, (k, x.m, x.n)
        assert np.array_equal(s.rows, o.rows) and np.array_equal(s.cols, o.cols), (k, x.m, x.n)
        assert np.array_equal(s.flows.astype(np.int64), o.flows), (k, x.m, x.n)
        assert np.array_equal(s.u.astype(np.int64), o.u) and np.array_equal(s.v.astype(np.int64), o.v), (k, x.m, x.n)
    # alone: the plain solver on a sample spanning the sizes, the four large ones included
    for k in sorted(range(len(reqs)), key=lambda k: -reqs[k].m * reqs[k].n)[:8] + list(range(0, len(reqs), 9)):
        x, s = reqs[k], sols[k]
        say("alone", k, x.m, x.n)
        a = F.solve(x, pricing="dantzig")
        assert a.status == 0 and a.objective == s.objective and a.iterations == s.iterations, (k, x.m, x.n)
        assert np.array_equal(a.rows, s.rows) and np.array_equal(a.flows, s.flows), (k, x.m, x.n)
        assert np.array_equal(a.u, s.u) and np.array_equal(a.v, s.v), (k, x.m, x.n)
