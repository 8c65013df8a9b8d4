"""flock_solve_noalloc (no device allocation; every buffer in the library's fixed device
array) against flock_solve, through the C ABI: two calls in a row, the second on a
smaller instance (the array keeps the first call's contents), then a third on a
2000 x 2000 instance near the array's size; each bit for bit equal to flock_solve with
Dantzig pricing (objective, pivots, basis slot by slot, flows, duals, flow list) and to
the oracle.  An instance beyond the array is refused with FLOCK_E_TOO_LARGE before any
device work."""
import ctypes

import numpy as np
import pytest

import flock_amd as F
import oracle
from flock_amd import instances as I

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.status == 0 and b.status == 0, (a.status, b.status, F.last_error())
    assert a.objective == b.objective and a.iterations == b.iterations
    assert np.array_equal(a.rows, b.rows) and np.array_equal(a.cols, b.cols) and np.array_equal(a.flows, b.flows)
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v)
    assert a.n_flows == b.n_flows and np.array_equal(a.flow_list, b.flow_list)


def test_noalloc_equals_flock_solve_twice_then_smaller():
    rng = np.random.default_rng(21)
    cases = [I.random_instance(rng, 600, 500, 1000, 10000), I.random_instance(rng, 37, 91, 100, 1000),
             I.random_instance(rng, 2000, 2000, 1000, 10000)]
    for k, inst in enumerate(cases):
        print("noalloc case", k, inst.m, inst.n, flush=True)
        a = F.solve_noalloc(inst, pricing="dantzig")
        b = F.solve(inst, pricing="dantzig")
        same(a, b)
        if inst.m * inst.n <= 300000:
            o = oracle.solve(inst.cost, inst.supply, inst.demand)
            assert a.objective == o.objective and a.iterations == o.iterations
        # the flow list is the positive-flow basic arcs in slot order
        pos = b.flows != 0
        assert np.array_equal(a.flow_list, np.stack([b.rows[pos], b.cols[pos], b.flows[pos]], axis=1))


def test_noalloc_candidates_objective_and_certificate():
    rng = np.random.default_rng(22)
    for m, n in [(300, 200), (64, 64)]:
        inst = I.random_instance(rng, m, n, 100, 1000)
        a = F.solve_noalloc(inst, pricing="candidates")
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        assert a.status == 0 and a.objective == o.objective
        bad, obj = oracle.certify(inst.cost, inst.supply, inst.demand, a.rows, a.cols, a.flows.astype(np.int64))
        assert bad == 0 and obj == a.objective


def test_noalloc_refuses_instances_beyond_its_array():
    # 3000 x 3000: the matrix and its transpose alone are 72 MB, beyond the 48 MiB array
    inst = I.Instance(np.ones((3000, 3000), np.int32), np.full(3000, 4, np.int32), np.full(3000, 4, np.int32))
    with pytest.raises(F.FlockError) as ei:
        F.solve_noalloc(inst)
    assert ei.value.status == F.E_TOO_LARGE and "scratch" in str(ei.value)
    L = F.lib()
    assert hasattr(L, "flock_solve_noalloc") and isinstance(L.flock_solve_noalloc, ctypes._CFuncPtr)
