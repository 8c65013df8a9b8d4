"""Continuous-batching server on the GPU: every request's result is the same alone and
in the mix; the HBM budget bounds admission without losing requests; oversize requests
are refused."""
import sys

import numpy as np
import pytest

import flock_amd as F
import oracle
from flock_amd import instances as I

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mix(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    sizes = I.log_uniform_sizes(rng, 2 * count, lo, hi)
    return [I.random_instance(rng, int(sizes[2 * k]), int(sizes[2 * k + 1]), 100, 1000) for k in range(count)]


def drain(srv, want, timeout_polls=100000):
    got = {}
    polls = 0
    while len(got) < want and polls < timeout_polls:
        for t, s in srv.poll(max_wait_ms=50.0, cap=64, full=True):
            got[t] = s
        polls += 1
    return got


def test_server_alone_equals_mix_dantzig():
    insts = _mix(3, 40, 10, 300)
    say("server mix 40")
    srv = F.Server(hbm_budget_bytes=1 << 30, max_slots=16, step_pivots=7, pricing="dantzig")
    tickets = [srv.submit(x) for x in insts]
    got = drain(srv, len(insts))
    srv.close()
    assert len(got) == len(insts)
    for t, x in zip(tickets, insts):
        s = got[t]
        alone = F.solve(x, pricing="dantzig")
        assert s.status == 0 and s.objective == alone.objective and s.iterations == alone.iterations
        assert np.array_equal(s.rows, alone.rows) and np.array_equal(s.flows, alone.flows)
        assert np.array_equal(s.u, alone.u) and np.array_equal(s.v, alone.v)


def test_server_short_slices_equal_alone():
    """Slices of 300 us cut every solve at pivot boundaries (and the engine
    configuration changes between slices as the largest active request changes); each
    result still equals the request solved alone, pivot for pivot."""
    insts = _mix(5, 24, 8, 700)
    say("server short slices 24")
    srv = F.Server(hbm_budget_bytes=1 << 30, max_slots=64, pricing="dantzig", slice_us=300)
    tickets = [srv.submit(x) for x in insts]
    got = drain(srv, len(insts))
    srv.close()
    assert len(got) == len(insts)
    for t, x in zip(tickets, insts):
        s = got[t]
        alone = F.solve(x, pricing="dantzig")
        assert s.status == 0 and s.objective == alone.objective and s.iterations == alone.iterations
        assert np.array_equal(s.rows, alone.rows) and np.array_equal(s.flows, alone.flows)
        assert np.array_equal(s.u, alone.u) and np.array_equal(s.v, alone.v)


def test_server_candidates_tight_budget_and_arrivals_during_work():
    insts = _mix(4, 30, 16, 400)
    # budget below the sum of the requests: admission waits for memory to free up
    srv = F.Server(hbm_budget_bytes=24 << 20, max_slots=8, step_pivots=32, pricing="candidates")
    say("server tight budget")
    tickets = [srv.submit(x) for x in insts[:10]]
    got = {}
    for t, s in srv.poll(max_wait_ms=1.0, cap=64, full=True):   # some work is in flight now
        got[t] = s
    tickets += [srv.submit(x) for x in insts[10:]]              # arrivals while others run
    got.update(drain(srv, len(insts) - len(got)))
    srv.close()
    assert len(got) == len(insts)
    for t, x in zip(tickets, insts):
        s = got[t]
        o = oracle.solve(x.cost, x.supply, x.demand)
        assert s.status == 0 and s.objective == o.objective
        bad, _ = oracle.certify(x.cost, x.supply, x.demand, s.rows, s.cols, s.flows.astype(np.int64))
        assert bad == 0



def test_server_range_edges_resume_exact():
    """Requests at the edges of the int32 range contract (the largest costs it admits with
    padding cells in every row, one cost everywhere, a total supply of 2^31 - 1), cut
    every 3 pivots so that their trees, flows and duals go through the slot's saved
    state many times: each result equals the oracle's pivot for pivot."""
    rng = np.random.default_rng(4545)
    m, n = 37, 53
    cmax = (2 ** 31 - 2) // (2 * (m + n) + 1)
    supply = rng.integers(1, 1_000_001, size=m).astype(np.int32)
    demand = I.balance_demands(rng, supply, 1, 1_000_000, n)
    insts = [I.Instance(rng.integers(cmax // 2, cmax + 1, size=(m, n)).astype(np.int32), supply, demand),
             I.Instance(np.where(rng.random((m, n)) < 0.5, 0, cmax).astype(np.int32), supply, demand),
             I.Instance(np.full((m, n), 7, np.int32), supply, demand),
             I.Instance(rng.integers(0, (2 ** 31 - 2) // 17 + 1, size=(3, 5)).astype(np.int32),
                        np.array([715827882, 715827882, 715827883], np.int32),
                        np.array([429496729] * 4 + [429496731], np.int32))]
    say("server range edges")
    srv = F.Server(hbm_budget_bytes=1 << 30, max_slots=8, step_pivots=3, pricing="dantzig")
    tickets = [srv.submit(x) for x in insts]
    got = drain(srv, len(insts))
    srv.close()
    assert len(got) == len(insts)
    for t, x in zip(tickets, insts):
        s = got[t]
        o = oracle.solve(x.cost, x.supply, x.demand)
        assert s.status == 0 and s.objective == o.objective and s.iterations == o.iterations
        assert np.array_equal(s.rows, o.rows) and np.array_equal(s.cols, o.cols)
        assert np.array_equal(s.flows.astype(np.int64), o.flows)
        assert np.array_equal(s.u.astype(np.int64), o.u) and np.array_equal(s.v.astype(np.int64), o.v)

def test_server_refuses_oversize_request():
    srv = F.Server(hbm_budget_bytes=64 << 10, max_slots=4)
    big = I.random_instance(np.random.default_rng(1), 200, 200, 10, 100)
    with pytest.raises(F.FlockError) as ei:
        srv.submit(big)
    assert ei.value.status == F.E_BUDGET
    srv.close()
