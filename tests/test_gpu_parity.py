"""GPU parity: the HIP engine against the CPU oracle, through the C ABI.

Bar: bit-exact.  With Dantzig pricing the device follows the sequential reference's
rules, so objective, pivot count, the basis slot by slot and the duals must all equal
the oracle's.  With the faster pricing rules the pivot sequence differs; the int64
objective must still equal the oracle's and every basis must pass the LP optimality
certificate (exact primal sums, an (m+n-1)-arc spanning tree, all reduced costs >= 0).
"""
import json
import os
import sys

import numpy as np
import pytest

import flock_amd as F
import oracle
from flock_amd import instances as I

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EXPECTED = json.load(open(os.path.join(GOLD, "expected.json")))
GPU_OK = {k: v for k, v in EXPECTED.items()}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def certify(inst, s):
    bad, obj = oracle.certify(inst.cost, inst.supply, inst.demand, s.rows, s.cols, s.flows.astype(np.int64),
                              s.u.astype(np.int64), s.v.astype(np.int64))
    return bad, obj


def assert_same_as_oracle(inst, s, o, exact_sequence):
    assert s.status == 0, (s.status, F.last_error())
    assert s.objective == o.objective
    bad, obj = certify(inst, s)
    assert bad == 0 and obj == s.objective
    if exact_sequence:
        assert s.iterations == o.iterations
        assert np.array_equal(s.rows, o.rows) and np.array_equal(s.cols, o.cols)
        assert np.array_equal(s.flows.astype(np.int64), o.flows)
        assert np.array_equal(s.u.astype(np.int64), o.u) and np.array_equal(s.v.astype(np.int64), o.v)
        assert s.n_flows == o.nonzero_flows


def fits_int32_engine(inst):
    return (2 * (inst.m + inst.n) + 1) * int(inst.cost.max()) < 2 ** 31 - 1


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_golden_fixtures_dantzig(name):
    from conftest import load_golden
    inst = load_golden(name)
    say("golden", name, inst.m, inst.n)
    if not fits_int32_engine(inst):
        with pytest.raises(F.FlockError) as ei:
            F.solve(inst)
        assert ei.value.status == F.E_RANGE
        return
    s = F.solve(inst, pricing="dantzig")
    o = oracle.solve(inst.cost, inst.supply, inst.demand)
    assert s.objective == EXPECTED[name]["optimum"]
    assert_same_as_oracle(inst, s, o, exact_sequence=True)


def _rand(rng, m, n, cmax=50, smax=100, zeros=False):
    cost = rng.integers(1, cmax + 1, size=(m, n)).astype(np.int32)
    supply = rng.integers(0 if zeros else 1, smax + 1, size=m).astype(np.int32)
    supply[0] += 1
    if zeros:
        cuts = np.sort(rng.integers(0, int(supply.sum()) + 1, size=n - 1))
        demand = np.diff(np.concatenate([[0], cuts, [supply.sum()]])).astype(np.int32)
    else:
        demand = I.balance_demands(rng, supply, 1, smax, n) if supply.sum() >= n else None
        if demand is None:
            return _rand(rng, m, n, cmax, smax, zeros=True)
    return I.Instance(cost, supply, demand)


SHAPES = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 5), (16, 16), (17, 33), (64, 3), (3, 64), (100, 100),
          (128, 257), (300, 40), (511, 512)]


@pytest.mark.parametrize("shape", SHAPES)
def test_random_shapes_dantzig_exact_sequence(shape):
    rng = np.random.default_rng(1000 + shape[0] * 7 + shape[1])
    for zeros in (False, True):
        inst = _rand(rng, *shape, zeros=zeros)
        say("shape", shape, "zeros", zeros)
        s = F.solve(inst, pricing="dantzig")
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        assert_same_as_oracle(inst, s, o, exact_sequence=True)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_plain_100x100_dantzig(seed):
    inst = I.plain_small(seed)
    say("plain100 seed", seed)
    s = F.solve(inst, pricing="dantzig")
    o = oracle.solve(inst.cost, inst.supply, inst.demand)
    assert_same_as_oracle(inst, s, o, exact_sequence=True)


@pytest.mark.parametrize("pricing", ["candidates", "multi"])
def test_fast_pricing_objective_and_certificate(pricing):
    rng = np.random.default_rng(77)
    for shape in [(5, 9), (40, 60), (100, 100), (257, 130)]:
        inst = _rand(rng, *shape)
        say("pricing", pricing, shape)
        s = F.solve(inst, pricing=pricing)
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        assert_same_as_oracle(inst, s, o, exact_sequence=False)


@pytest.mark.parametrize("cost,supply,demand", [
    ([[2, 2], [6, 4], [5, 8]], [5, 7, 2], [6, 8]),          # argmax tie: the last maximum wins
    ([[8, 1, 9], [2, 2, 4]], [6, 4], [6, 1, 3]),            # stale penalties of the other kind
])
def test_vogel_tie_and_stale_rules_on_device(cost, supply, demand):
    """The two hand-derived Vogel pins of the oracle tests: the device's initial basis
    (one pivot budget: the same basis must come out) and full solve equal the oracle's."""
    inst = I.Instance(np.array(cost, np.int32), np.array(supply, np.int32), np.array(demand, np.int32))
    r, c, f = oracle.vogel(inst.cost, inst.supply, inst.demand)
    for threads in (64, 256):
        s = F.solve(inst, pricing="dantzig", max_iterations=1, threads=threads, team=1 if threads == 64 else 0)
        o = oracle.solve(inst.cost, inst.supply, inst.demand, max_iterations=1)
        assert np.array_equal(s.rows, o.rows) and np.array_equal(s.flows.astype(np.int64), o.flows)
        s = F.solve(inst, pricing="dantzig", threads=threads, team=1 if threads == 64 else 0)
        assert_same_as_oracle(inst, s, oracle.solve(inst.cost, inst.supply, inst.demand), exact_sequence=True)
    assert len(r) == inst.m + inst.n - 1 and int(f.sum()) == int(inst.supply.sum())


def test_nwc_initial_matches_oracle():
    for seed in (4, 5):
        inst = I.plain_small(seed)
        say("nwc seed", seed)
        s = F.solve(inst, pricing="dantzig", initial="nwc")
        o = oracle.solve(inst.cost, inst.supply, inst.demand, initial="nwc")
        assert_same_as_oracle(inst, s, o, exact_sequence=True)


def test_degenerate_family_small_exact():
    inst = I.degenerate(seed=3, m=200, n=100)
    say("degenerate 200x100")
    s = F.solve(inst, pricing="dantzig")
    o = oracle.solve(inst.cost, inst.supply, inst.demand)
    assert_same_as_oracle(inst, s, o, exact_sequence=True)


def test_iteration_budget_reported():
    inst = I.plain_small(1)
    s = F.solve(inst, pricing="dantzig", max_iterations=5)
    assert s.status == F.E_ITERATION_LIMIT and s.iterations == 5
    o = oracle.solve(inst.cost, inst.supply, inst.demand, max_iterations=5)
    # the budget stops both at the same basis
    assert np.array_equal(s.rows, o.rows) and np.array_equal(s.flows.astype(np.int64), o.flows)


def test_batch_equals_plain_solves():
    insts = I.batch(seed=5, count=48)
    say("batch 48")
    sols, ms = F.batch_solve(insts, pricing="dantzig")
    assert ms > 0
    for inst, s in zip(insts, sols):
        p = F.solve(inst, pricing="dantzig")
        assert s.status == 0 and s.objective == p.objective and s.iterations == p.iterations
        assert np.array_equal(s.rows, p.rows) and np.array_equal(s.flows, p.flows)
        assert certify(inst, s)[0] == 0


@pytest.mark.parametrize("threads,team", [(64, 1), (64, 2), (64, 4), (256, 0), (512, 0), (1024, 0)])
def test_every_kernel_configuration_exact(threads, team):
    """Wavefront-per-instance (1, 2 or 4 per workgroup) and workgroup-per-instance
    engines all follow the same pivot sequence as the oracle."""
    rng = np.random.default_rng(31 + threads + team)
    for shape in [(5, 9), (37, 61), (120, 130)]:
        inst = _rand(rng, *shape)
        say("cfg", threads, team, shape)
        s = F.solve(inst, pricing="dantzig", threads=threads, team=team)
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        assert_same_as_oracle(inst, s, o, exact_sequence=True)
        for pricing in ("candidates", "multi"):
            c = F.solve(inst, pricing=pricing, threads=threads, team=team)
            assert c.objective == o.objective and certify(inst, c)[0] == 0, pricing


def test_batch_mixed_configurations():
    """One batch mixing wavefront and workgroup classes, candidates pricing."""
    rng = np.random.default_rng(8)
    insts = [_rand(rng, int(a), int(b)) for a, b in [(20, 30), (700, 300), (64, 64), (1500, 900), (3, 400)]]
    say("mixed batch")
    sols, _ = F.batch_solve(insts, pricing="candidates")
    for inst, s in zip(insts, sols):
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        assert s.status == 0 and s.objective == o.objective and certify(inst, s)[0] == 0




def _random_batch(seed, count):
    """Seeded mixed instances: m, n log-uniform in [1, 200], cost ranges from 1 (ties
    everywhere) to 10^4, zero supplies in a third of them (and zero demands where the
    random cuts meet), n often not a multiple of 4."""
    rng = np.random.default_rng(seed)
    sizes = I.log_uniform_sizes(rng, 2 * count, 1, 200)
    out = []
    for k in range(count):
        m, n = int(sizes[2 * k]), int(sizes[2 * k + 1])
        cmax = int(rng.choice([1, 5, 100, 10_000]))
        cost = rng.integers(1, cmax + 1, size=(m, n)).astype(np.int32)
        supply = rng.integers(1, 1001, size=m).astype(np.int32)
        if k % 3 == 0:
            supply[rng.random(m) < 0.3] = 0
        supply[0] += 1
        cuts = np.sort(rng.integers(0, int(supply.sum()) + 1, size=n - 1))
        demand = np.diff(np.concatenate([[0], cuts, [supply.sum()]])).astype(np.int32)
        out.append(I.Instance(cost, supply, demand))
    return out


@pytest.mark.parametrize("threads", [0, 64, 1024])
def test_random_batch_dantzig_exact(threads):
    """256 seeded instances in one batch launch (the default size-class policy, or every
    instance on one-wavefront groups, or on 1024-thread workgroups):
    every one follows the oracle pivot for pivot -- objective, pivot count, basis slot by
    slot, flows and duals."""
    insts = _random_batch(900 + threads, 256)
    say("random batch of", len(insts), "threads", threads)
    b = F.Batch(insts)
    try:
        b.run("dantzig", threads=threads)
        sols = b.results()
    finally:
        b.close()
    for k, (inst, s) in enumerate(zip(insts, sols)):
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        try:
            assert_same_as_oracle(inst, s, o, exact_sequence=True)
        except AssertionError:
            say("instance", k, inst.m, inst.n)
            raise


@pytest.mark.parametrize("pricing", ["candidates", "multi", "switch"])
def test_random_batch_fast_pricing_certified(pricing):
    """The same kind of batch under the faster pricing rules: every objective equals the
    oracle's and every basis passes the exact certificate."""
    insts = _random_batch(1300, 256)
    say("random batch of", len(insts), "pricing", pricing)
    sols, _ = F.batch_solve(insts, pricing=pricing)
    for k, (inst, s) in enumerate(zip(insts, sols)):
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        bad, obj = certify(inst, s)
        assert s.status == 0 and s.objective == o.objective and bad == 0 and obj == s.objective, (k, inst.m, inst.n)

def _edge_instances():
    """Instances at the edges of the int32 range contract ((2(m+n)+1) * max cost < 2^31,
    total supply < 2^31): the largest costs it admits, costs of 0 beside them, one cost
    everywhere (every basis optimal, ties throughout), and a total supply of 2^31 - 1."""
    rng = np.random.default_rng(4242)
    m, n = 37, 53  # n is not a multiple of 4: every row has padding cells
    cmax = (2 ** 31 - 2) // (2 * (m + n) + 1)
    supply = rng.integers(1, 1_000_001, size=m).astype(np.int32)
    demand = I.balance_demands(rng, supply, 1, 1_000_000, n)
    top = rng.integers(cmax // 2, cmax + 1, size=(m, n)).astype(np.int32)
    top[0, 0] = cmax
    zero_top = np.where(rng.random((m, n)) < 0.5, 0, cmax).astype(np.int32)
    flat = np.full((m, n), 7, np.int32)
    out = [("largest costs", I.Instance(top, supply, demand)),
           ("costs 0 or largest", I.Instance(zero_top, supply, demand)),
           ("one cost everywhere", I.Instance(flat, supply, demand))]
    big = np.array([715827882, 715827882, 715827883], np.int32)  # sums to 2^31 - 1
    dem = np.array([429496729] * 4 + [429496731], np.int32)
    assert int(big.sum()) == int(dem.sum()) == 2 ** 31 - 1
    c5 = rng.integers(0, (2 ** 31 - 2) // 17 + 1, size=(3, 5)).astype(np.int32)
    out.append(("total supply 2^31 - 1", I.Instance(c5, big, dem)))
    return out


@pytest.mark.parametrize("threads,team", [(64, 1), (256, 0), (1024, 0)])
def test_range_contract_edges_exact(threads, team):
    """At the edges of the range contract every engine follows the oracle's pivot
    sequence with Dantzig pricing (the oracle keeps duals and flows in int64), and the
    candidate rule's basis passes the exact certificate with the oracle's objective."""
    for what, inst in _edge_instances():
        assert (2 * (inst.m + inst.n) + 1) * int(inst.cost.max()) < 2 ** 31 - 1
        say("range edge", what, "threads", threads)
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        s = F.solve(inst, pricing="dantzig", threads=threads, team=team)
        assert_same_as_oracle(inst, s, o, exact_sequence=True)
        c = F.solve(inst, pricing="candidates", threads=threads, team=team)
        assert c.status == 0 and c.objective == o.objective and certify(inst, c)[0] == 0, what
    # one past the contract is refused before any device work
    m, n = 37, 53
    over = np.full((m, n), (2 ** 31 - 2) // (2 * (m + n) + 1) + 1, np.int32)
    sup = np.full(m, n, np.int32)
    with pytest.raises(F.FlockError) as ei:
        F.solve(I.Instance(over, sup, np.full(n, m, np.int32)))
    assert ei.value.status == F.E_RANGE

@pytest.mark.parametrize("threads", [512, 1024])
def test_large_graph_path_1000x1000(threads):
    """1000x1000 runs as a graph of sweep-grid and pivot-round kernels (set-up and round
    workgroups of 512 or 1024 threads): Dantzig equals the oracle pivot for pivot; the
    candidate, multi-pivot and switch rules reach the same certified optimum."""
    inst = I.plain_large(2)
    say("graph path 1000x1000, threads", threads)
    assert inst.m * inst.n >= F.get_tuning()["graph_min_cells"] > 0
    before = F.get_tuning()
    try:
        F.set_tuning(large_threads=threads)
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        s = F.solve(inst, pricing="dantzig")
        assert_same_as_oracle(inst, s, o, exact_sequence=True)
        assert s.sweeps == s.iterations + 1
        for pricing in ("candidates", "multi", "switch"):
            c = F.solve(inst, pricing=pricing)
            assert c.status == 0 and c.objective == o.objective and certify(inst, c)[0] == 0, pricing
    finally:
        F.set_tuning(**before)



def test_large_graph_path_range_contract_edges():
    """The graph path at the edges of the range contract: 500x502 (padding cells in every
    row) with costs up to the largest it admits -- Dantzig equals the oracle pivot for
    pivot, candidates reach its optimum with a certified basis -- and with one cost
    everywhere (Vogel's basis is already optimal: no pivot)."""
    rng = np.random.default_rng(4343)
    m, n = 500, 502
    cmax = (2 ** 31 - 2) // (2 * (m + n) + 1)
    supply = rng.integers(1, 100_001, size=m).astype(np.int32)
    demand = I.balance_demands(rng, supply, 1, 100_000, n)
    top = rng.integers(cmax // 2, cmax + 1, size=(m, n)).astype(np.int32)
    top[m - 1, n - 1] = cmax
    assert m * n >= F.get_tuning()["graph_min_cells"] > 0
    inst = I.Instance(top, supply, demand)
    say("graph path, largest costs 500x502")
    o = oracle.solve(inst.cost, inst.supply, inst.demand)
    s = F.solve(inst, pricing="dantzig")
    assert_same_as_oracle(inst, s, o, exact_sequence=True)
    c = F.solve(inst, pricing="candidates")
    assert c.status == 0 and c.objective == o.objective and certify(inst, c)[0] == 0
    flat = I.Instance(np.full((m, n), 3, np.int32), supply, demand)
    say("graph path, one cost everywhere 500x502")
    o = oracle.solve(flat.cost, flat.supply, flat.demand)
    for pricing in ("dantzig", "candidates"):
        f = F.solve(flat, pricing=pricing)
        assert f.iterations == 0 and o.iterations == 0
        assert f.status == 0 and f.objective == o.objective == 3 * int(supply.sum())
        assert certify(flat, f)[0] == 0

def test_large_graph_path_degenerate_4000x2000():
    """The degenerate 4000x2000 family (costs in 1..8, ties everywhere) through the graph
    path: certified optimal (exact primal sums, spanning tree, reduced costs >= 0)."""
    inst = I.degenerate(seed=1)
    say("graph path degenerate 4000x2000")
    for pricing in ("candidates", "dantzig", "multi"):
        s = F.solve(inst, pricing=pricing)
        assert s.status == 0, (pricing, s.status, F.last_error())
        bad, obj = certify(inst, s)
        assert bad == 0 and obj == s.objective


def test_large_path_vogel_eight_lines_per_rank():
    """m+n between 4096 and 8192 on the large-instance path: Vogel's set-up kernel keeps
    8 lines per rank in registers, with the tracked keys in global scratch (they do not
    fit LDS beside the line state).  Stopped after one Dantzig pivot, the basis, flows
    and duals equal the oracle's (status 1: the iteration budget)."""
    inst = I.random_instance(np.random.default_rng(4200), 600, 3600, 1000, 10000)
    say("large path, Vogel 8 lines per rank, 600x3600")
    o = oracle.solve(inst.cost, inst.supply, inst.demand, max_iterations=1)
    s = F.solve(inst, pricing="dantzig", max_iterations=1)
    assert o.status == 1 and s.status == 1 and s.iterations == o.iterations == 1
    assert np.array_equal(s.rows, o.rows) and np.array_equal(s.cols, o.cols)
    assert np.array_equal(s.flows.astype(np.int64), o.flows)
    assert np.array_equal(s.u.astype(np.int64), o.u) and np.array_equal(s.v.astype(np.int64), o.v)


@pytest.mark.parametrize("shape", [(20, 8200), (8100, 6)])
def test_tree_beyond_lds_exact(shape):
    """m+n above 8000: the basis tree no longer fits one workgroup's LDS, so flock_solve
    takes the large-instance path with the tree kept in global memory (Vogel's line state
    too when it does not fit).  Dantzig equals the oracle pivot for pivot; the other
    rules reach the same certified optimum."""
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    inst = I.random_instance(rng, *shape, 1000, 10000)
    say("tree beyond LDS", shape)
    o = oracle.solve(inst.cost, inst.supply, inst.demand)
    s = F.solve(inst, pricing="dantzig")
    assert_same_as_oracle(inst, s, o, exact_sequence=True)
    for pricing in ("candidates", "multi", "switch"):
        c = F.solve(inst, pricing=pricing)
        assert c.status == 0 and c.objective == o.objective and certify(inst, c)[0] == 0, pricing


def test_reference_largest_shape_1000x10000():
    """The reference's largest logged instance shape (1000 x 10000, results/parallel_uv:1423),
    with costs U{1..1000} and supplies/demands U{1..10000}: certified optimal with the
    candidate and switch rules, and both reach the same objective."""
    inst = I.random_instance(np.random.default_rng(10000), 1000, 10000, 1000, 10000)
    say("1000x10000")
    a = F.solve(inst, pricing="candidates")
    assert a.status == 0 and certify(inst, a)[0] == 0
    b = F.solve(inst, pricing="switch")
    assert b.status == 0 and certify(inst, b)[0] == 0 and b.objective == a.objective


def test_instances_too_large_for_256_thread_lds_use_the_1024_engine():
    """A 256- or 512-thread engine needs the tree and the pivot snapshot in LDS; larger
    instances asked for with those sizes (explicitly, or by a size policy) are solved by
    the 1024-thread engine: certified optimal, same objective as the graph path."""
    big = I.degenerate(seed=2, m=2400, n=1600)
    small = I.plain_small(3)
    say("routing 2400x1600 with threads=256")
    ref = F.solve(big, pricing="candidates")
    assert ref.status == 0
    for threads in (256, 512):
        s = F.solve(big, pricing="candidates", threads=threads)
        assert s.status == 0, (threads, s.status, F.last_error())
        bad, obj = certify(big, s)
        assert bad == 0 and obj == s.objective == ref.objective
    before = F.get_tuning()
    try:
        F.set_tuning(wg256_max_v=8000)
        sols, _ = F.batch_solve([big, small, big], pricing="candidates")
        assert [x.status for x in sols] == [0, 0, 0]
        assert sols[0].objective == sols[2].objective == ref.objective
        o = oracle.solve(small.cost, small.supply, small.demand)
        assert sols[1].objective == o.objective
    finally:
        F.set_tuning(**before)


def test_autotune_installs_a_policy_with_identical_objectives():
    """flock_autotune measures policies on a sample, requires identical objectives from
    every one (it fails otherwise), and installs the fastest."""
    insts = I.batch(seed=9, count=96)
    before = F.get_tuning()
    say("autotune 96")
    try:
        tuning, ms = F.autotune(insts, pricing="candidates")
        assert ms > 0 and F.get_tuning() == tuning
        sols, _ = F.batch_solve(insts, pricing="candidates")
        for inst, s in zip(insts[:20], sols[:20]):
            o = oracle.solve(inst.cost, inst.supply, inst.demand)
            assert s.status == 0 and s.objective == o.objective
    finally:
        F.set_tuning(**before)


def test_autotune_large_picks_a_tile_shape_with_the_same_optimum():
    """flock_autotune_large tries every sweep tile shape on one large instance, requires the
    same objective from each (it fails otherwise) and installs the fastest; the installed
    shape solves to the oracle's optimum, and Dantzig stays pivot for pivot the oracle's
    whatever the shape."""
    inst = I.random_instance(np.random.default_rng(77), 600, 500, 1000, 10000)
    before = F.get_tuning()
    say("autotune_large 600x500")
    try:
        tuning, ms = F.autotune_large(inst, pricing="candidates")
        assert ms > 0 and tuning["sweep_vectors"] in (1, 2, 4, 8) and tuning["large_threads"] in (512, 1024)
        assert F.get_tuning() == tuning
        o = oracle.solve(inst.cost, inst.supply, inst.demand)
        s = F.solve(inst, pricing="candidates")
        assert s.status == 0 and s.objective == o.objective and certify(inst, s)[0] == 0
        for v in (1, 8):
            F.set_tuning(sweep_vectors=v)
            assert_same_as_oracle(inst, F.solve(inst, pricing="dantzig"), o, exact_sequence=True)
    finally:
        F.set_tuning(**before)
