"""Pins of the CPU oracle (oracle/transport_oracle.cpp) -- CPU only.

The oracle is checked against things that do not depend on it:
  * golden optima of the reference's own instance files (tests/golden/expected.json:
    exact integer optima from an independent LP solver, plus the reference's logged
    Gurobi and sequential-path objectives);
  * textbook answers for Vogel's approximation and the north-west corner rule;
  * brute force over all bases of tiny instances;
  * closed forms (one row / one column, constant costs, rank-one costs);
  * an independent LP solve (scipy HiGHS) on random small instances;
  * planted mistakes that the certificate checker must reject.
"""
import itertools
import json
import os

import numpy as np
import pytest
from scipy.optimize import linprog
from scipy.sparse import coo_matrix

import oracle
from flock_amd import instances as I

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EXPECTED = json.load(open(os.path.join(GOLD, "expected.json")))


def load(name):
    from conftest import load_golden
    return load_golden(name)


def lp_opt(cost, supply, demand):
    m, n = cost.shape
    rows = np.concatenate([np.repeat(np.arange(m), n), m + np.tile(np.arange(n), m)])
    cols = np.concatenate([np.arange(m * n), np.arange(m * n)])
    A = coo_matrix((np.ones(2 * m * n), (rows, cols)), shape=(m + n, m * n)).tocsr()
    b = np.concatenate([supply, demand]).astype(float)
    r = linprog(cost.ravel().astype(float), A_eq=A, b_eq=b, bounds=(0, None), method="highs")
    assert r.status == 0
    return int(round(r.fun))


def rand_small(rng, m, n, cmax=20, smax=30, zeros=True):
    cost = rng.integers(1, cmax + 1, size=(m, n)).astype(np.int32)
    supply = rng.integers(0 if zeros else 1, smax, size=m).astype(np.int32)
    supply[0] += 1
    cuts = np.sort(rng.integers(0, int(supply.sum()) + 1, size=n - 1))
    demand = np.diff(np.concatenate([[0], cuts, [supply.sum()]])).astype(np.int32)
    return cost, supply, demand


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_golden_optimum(name):
    e = EXPECTED[name]
    inst = load(name)
    assert (inst.m, inst.n, inst.scale) == (e["m"], e["n"], e["scale"])
    s = oracle.solve(inst.cost, inst.supply, inst.demand)
    assert s.status == 0
    assert s.objective == e["optimum"]
    bad, obj = oracle.certify(inst.cost, inst.supply, inst.demand, s.rows, s.cols, s.flows, s.u, s.v)
    assert bad == 0 and obj == s.objective
    if e["gurobi"]:
        # Gurobi solved the float32-rounded costs: agree to float32 noise
        g = float(e["gurobi"]) * 10 ** e["scale"]
        assert abs(g - s.objective) <= 2e-7 * abs(g)
    if e["seq_log"]:
        # the reference's sequential path printed 6 significant digits
        assert float(f"{s.objective / 10 ** e['scale']:.6g}") == float(e["seq_log"])


def test_vogel_textbook_allocation():
    # classic 3x4 example: Vogel reaches the optimum 2850 with this exact allocation
    inst = load("test_simple1.dat")
    r, c, f = oracle.vogel(inst.cost, inst.supply, inst.demand)
    alloc = {(int(a), int(b)): int(q) for a, b, q in zip(r, c, f)}
    assert alloc == {(0, 1): 300, (1, 0): 250, (1, 2): 150, (2, 1): 50, (2, 2): 250, (2, 3): 200}
    # second classic example: Vogel gives 779, the optimum is 743
    inst = load("test_simple2.dat")
    r, c, f = oracle.vogel(inst.cost, inst.supply, inst.demand)
    assert int((f * inst.cost[r, c]).sum()) == 779
    assert oracle.solve(inst.cost, inst.supply, inst.demand).objective == 743


def _alloc(r, c, f):
    return [(int(a), int(b), int(q)) for a, b, q in zip(r, c, f)]


def test_vogel_argmax_ties_go_to_the_last_maximum():
    """IBFS_vogel.cu:652-658 scans [rows | columns] with `>=`, so among equal penalties
    the LAST line wins.  Hand derivation on c = [[2,2],[6,4],[5,8]], s = [5,7,2],
    d = [6,8] (penalty = 2nd smallest - smallest uncovered cost):
      step 0: rows 0, 2, 3; columns 3, 2.  Row 2 and column 0 tie at 3; the last is
              column 0, its first minimum row is 0 (cost 2); demand 6 > supply 5 covers
              row 0: (0,0) gets 5.
      step 1: rows are recomputed (-1, 2, 3), columns are stale (3, 2): the tie repeats,
              column 0 again; cheapest uncovered row is 2 (cost 5); demand 1 <= supply 2
              covers column 0: (2,0) gets 1.
      step 2: columns recomputed (-1, 4); rows stale (-1, 2, 3): column 1, row 1 (cost
              4); demand 8 > supply 7 covers row 1: (1,1) gets 7.
      step 3: the last cell (2,1) gets 1.
    Objective 51.  A first-maximum rule picks row 2 at step 0 and ends at 48."""
    cost = np.array([[2, 2], [6, 4], [5, 8]], np.int32)
    r, c, f = oracle.vogel(cost, np.array([5, 7, 2], np.int32), np.array([6, 8], np.int32))
    assert _alloc(r, c, f) == [(0, 0, 5), (2, 0, 1), (1, 1, 7), (2, 1, 1)]
    assert int((f * cost[r, c]).sum()) == 51


def test_vogel_penalties_of_the_other_kind_stay_stale():
    """IBFS_vogel.cu:588-642 recomputes row penalties only after a row was covered and
    column penalties only after a column was covered.  Hand derivation on
    c = [[8,1,9],[2,2,4]], s = [6,4], d = [6,1,3]:
      step 0: rows 7, 0; columns 6, 1, 5: row 0, cheapest column 1; demand 1 <= supply
              6 covers column 1: (0,1) gets 1.
      step 1: columns recomputed (6, -1, 5), rows stale (7, 0) -- row 0's true penalty
              is now 9 - 8 = 1 -- so row 0 wins again with cheapest uncovered column 0
              (cost 8); demand 6 > supply 5 covers row 0: (0,0) gets 5.
      step 2: rows recomputed (-1, 2), columns stale (6, -1, 5): column 0, row 1; demand
              1 <= supply 4 covers column 0: (1,0) gets 1.
      step 3: the last cell (1,2) gets 3.
    Objective 55.  Fresh penalties would pick column 0 at step 1 and end at 52."""
    cost = np.array([[8, 1, 9], [2, 2, 4]], np.int32)
    r, c, f = oracle.vogel(cost, np.array([6, 4], np.int32), np.array([6, 1, 3], np.int32))
    assert _alloc(r, c, f) == [(0, 1, 1), (0, 0, 5), (1, 0, 1), (1, 2, 3)]
    assert int((f * cost[r, c]).sum()) == 55


def test_nwc_textbook_allocation():
    inst = load("test_simple1.dat")
    r, c, f = oracle.nwc(inst.supply, inst.demand)
    assert list(zip(r.tolist(), c.tolist(), f.tolist())) == [
        (0, 0, 250), (0, 1, 50), (1, 1, 300), (1, 2, 100), (2, 2, 300), (2, 3, 200)]
    assert int((f * inst.cost[r, c]).sum()) == 4400
    s = oracle.solve(inst.cost, inst.supply, inst.demand, initial="nwc")
    assert s.objective == 2850


def _spanning_tree(m, n, cells):
    parent = list(range(m + n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for (i, j) in cells:
        a, b = find(i), find(m + j)
        if a == b:
            return False
        parent[a] = b
    return True


def _basic_solution(m, n, cells, supply, demand):
    # peel leaves of the tree: a leaf's arc flow is its remaining supply/demand
    rs, rd = supply.astype(np.int64).copy(), demand.astype(np.int64).copy()
    left = set(cells)
    flow = {}
    while left:
        deg = {}
        for (i, j) in left:
            deg[("r", i)] = deg.get(("r", i), 0) + 1
            deg[("c", j)] = deg.get(("c", j), 0) + 1
        for (i, j) in list(left):
            if deg[("r", i)] == 1:
                q = rs[i]
            elif deg[("c", j)] == 1:
                q = rd[j]
            else:
                continue
            flow[(i, j)] = q
            rs[i] -= q
            rd[j] -= q
            left.discard((i, j))
            break
    return flow


def brute_force(cost, supply, demand):
    m, n = cost.shape
    cells = [(i, j) for i in range(m) for j in range(n)]
    best = None
    for sub in itertools.combinations(cells, m + n - 1):
        if not _spanning_tree(m, n, sub):
            continue
        flow = _basic_solution(m, n, sub, supply, demand)
        if min(flow.values()) < 0:
            continue
        obj = sum(int(q) * int(cost[c]) for c, q in flow.items())
        best = obj if best is None else min(best, obj)
    return best


def test_brute_force_tiny():
    rng = np.random.default_rng(11)
    for _ in range(120):
        m, n = int(rng.integers(1, 4)), int(rng.integers(1, 5))
        cost, supply, demand = rand_small(rng, m, n, cmax=9, smax=6)
        s = oracle.solve(cost, supply, demand)
        assert s.status == 0
        assert s.objective == brute_force(cost, supply, demand)



def test_range_contract_edges_brute_force_and_closed_form():
    """At the edges of the int32 range contract -- costs up to the largest
    (2(m+n)+1) * max cost < 2^31 admits, supplies summing to 2^31 - 1 -- the oracle's
    int64 objective equals exhaustive search over all bases (Python integers) on tiny
    instances, and the closed form of a rank-1 cost a_i + b_j on larger ones."""
    rng = np.random.default_rng(2024)
    total = 2 ** 31 - 1
    for _ in range(40):
        m, n = int(rng.integers(1, 4)), int(rng.integers(1, 5))
        cmax = (2 ** 31 - 2) // (2 * (m + n) + 1)
        cost = rng.integers(cmax // 2, cmax + 1, size=(m, n)).astype(np.int32)
        cost[rng.random((m, n)) < 0.2] = 0
        cuts = np.sort(rng.integers(0, total + 1, size=m - 1))
        supply = np.diff(np.concatenate([[0], cuts, [total]])).astype(np.int32)
        cuts = np.sort(rng.integers(0, total + 1, size=n - 1))
        demand = np.diff(np.concatenate([[0], cuts, [total]])).astype(np.int32)
        s = oracle.solve(cost, supply, demand)
        assert s.status == 0
        assert s.objective == brute_force(cost, supply, demand)
    for _ in range(10):
        m, n = int(rng.integers(2, 60)), int(rng.integers(2, 60))
        cmax = (2 ** 31 - 2) // (2 * (m + n) + 1)
        a = rng.integers(0, cmax // 2 + 1, size=m)
        b = rng.integers(0, cmax // 2 + 1, size=n)
        cost = (a[:, None] + b[None, :]).astype(np.int32)
        cuts = np.sort(rng.integers(0, total + 1, size=m - 1))
        supply = np.diff(np.concatenate([[0], cuts, [total]])).astype(np.int32)
        cuts = np.sort(rng.integers(0, total + 1, size=n - 1))
        demand = np.diff(np.concatenate([[0], cuts, [total]])).astype(np.int32)
        expect = sum(int(x) * int(y) for x, y in zip(a, supply)) + sum(int(x) * int(y) for x, y in zip(b, demand))
        assert oracle.solve(cost, supply, demand).objective == expect

def test_closed_forms():
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(1, 30))
        cost, supply, demand = rand_small(rng, 1, n)
        s = oracle.solve(cost, supply, demand)
        assert s.objective == int((cost[0].astype(np.int64) * demand).sum())
        cost, supply, demand = rand_small(rng, n, 1)
        s = oracle.solve(cost, supply, demand)
        assert s.objective == int((cost[:, 0].astype(np.int64) * supply).sum())
    for _ in range(10):
        m, n = int(rng.integers(2, 40)), int(rng.integers(2, 40))
        _, supply, demand = rand_small(rng, m, n)
        cost = np.full((m, n), 7, np.int32)
        assert oracle.solve(cost, supply, demand).objective == 7 * int(supply.sum())
        a = rng.integers(0, 50, size=m)
        b = rng.integers(0, 50, size=n)
        cost = (a[:, None] + b[None, :]).astype(np.int32)
        expect = int((a * supply).sum() + (b * demand).sum())
        assert oracle.solve(cost, supply, demand).objective == expect


def test_random_vs_lp():
    rng = np.random.default_rng(5)
    for _ in range(80):
        m, n = int(rng.integers(1, 25)), int(rng.integers(1, 25))
        cost, supply, demand = rand_small(rng, m, n, cmax=int(rng.integers(1, 60)), smax=40)
        s = oracle.solve(cost, supply, demand)
        assert s.status == 0 and s.objective == lp_opt(cost, supply, demand)
        for init in ("nwc",):
            assert oracle.solve(cost, supply, demand, initial=init).objective == s.objective


def test_degenerate_family_small():
    # the 4000x2000 family at 1/10 size: costs from {1..8}, supplies 50, demands 100
    inst = I.degenerate(seed=2, m=400, n=200)
    s = oracle.solve(inst.cost, inst.supply, inst.demand)
    assert s.status == 0
    assert oracle.certify(inst.cost, inst.supply, inst.demand, s.rows, s.cols, s.flows)[0] == 0
    assert s.objective == lp_opt(inst.cost, inst.supply, inst.demand)


def test_certificate_rejects_planted_mistakes():
    rng = np.random.default_rng(9)
    cost, supply, demand = rand_small(rng, 6, 7, zeros=False)
    s = oracle.solve(cost, supply, demand)
    base = (s.rows.copy(), s.cols.copy(), s.flows.copy())
    assert oracle.certify(cost, supply, demand, *base)[0] == 0
    # a flow moved off its row breaks the primal sums
    f = base[2].copy(); f[0] += 1
    assert oracle.certify(cost, supply, demand, base[0], base[1], f)[0] & 6
    # a repeated cell is not a spanning tree
    r = base[0].copy(); c = base[1].copy(); r[1], c[1] = r[0], c[0]
    assert oracle.certify(cost, supply, demand, r, c, base[2])[0] & 8
    # a feasible but non-optimal basis (Vogel's start of the 743 example) fails dual feasibility
    inst = load("test_simple2.dat")
    r, c, f = oracle.vogel(inst.cost, inst.supply, inst.demand)
    assert oracle.certify(inst.cost, inst.supply, inst.demand, r, c, f)[0] == 16
    # wrong duals are caught
    u = s.u.copy(); u[1] += 1
    assert oracle.certify(cost, supply, demand, *base, u, s.v)[0] & 32


def test_status_codes():
    cost = np.ones((3, 3), np.int32)
    s = oracle.solve(cost, np.array([1, 2, 3], np.int32), np.array([1, 2, 4], np.int32))
    assert s.status == 3  # unbalanced
    inst = I.plain_small(1)
    s = oracle.solve(inst.cost, inst.supply, inst.demand, max_iterations=5)
    assert s.status == 1 and s.iterations == 5  # budget exhausted, reported
