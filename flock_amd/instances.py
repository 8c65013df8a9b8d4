"""Balanced transportation instances: seeded generators and the `.dat` format.

`.dat` layout (src/instance.cpp:80-141): ``m n``, then m supplies, n demands, then m*n
costs in row-major order ``i*n+j``, whitespace separated.  Costs in the reference are
``float``; here they are parsed as exact decimals and scaled by the smallest power of
ten that makes every cost integral, so the solver works on int32 costs and an int64
objective (objective in the file's units = objective / 10**scale).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from decimal import Decimal

import numpy as np


@dataclass
class Instance:
    cost: np.ndarray      # int32 (m, n), row-major
    supply: np.ndarray    # int32 (m,)
    demand: np.ndarray    # int32 (n,)
    scale: int = 0        # costs were multiplied by 10**scale
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def m(self) -> int:
        return int(self.cost.shape[0])

    @property
    def n(self) -> int:
        return int(self.cost.shape[1])


# ---------------------------------------------------------------------------
# generators for the benchmark workloads
# ---------------------------------------------------------------------------

def balance_demands(rng: np.random.Generator, supply: np.ndarray, lo: int, hi: int, n: int) -> np.ndarray:
    """Draw demands uniform in [lo, hi], rescale them to the supply total and fix the
    last entry so the sums match exactly; redraw while the last entry would be < 1."""
    total = int(supply.sum())
    for _ in range(1000):
        d = rng.integers(lo, hi + 1, size=n).astype(np.int64)
        scaled = np.maximum((d * total) // int(d.sum()), 1)
        last = total - int(scaled[:-1].sum())
        if last >= 1:
            scaled[-1] = last
            return scaled.astype(np.int32)
    raise RuntimeError("could not balance demands")


def random_instance(rng: np.random.Generator, m: int, n: int, cmax: int, smax: int,
                    cmin: int = 1, smin: int = 1, name: str = "") -> Instance:
    cost = rng.integers(cmin, cmax + 1, size=(m, n), dtype=np.int32)
    supply = rng.integers(smin, smax + 1, size=m).astype(np.int32)
    demand = balance_demands(rng, supply, smin, smax, n)
    return Instance(cost, supply, demand, 0, name or f"rand_{m}x{n}")


def plain_small(seed: int = 1) -> Instance:
    """100 x 100, costs U{1..100}, supplies/demands U{1..1000}, balanced."""
    return random_instance(np.random.default_rng(seed), 100, 100, 100, 1000, name=f"plain100_s{seed}")


def plain_large(seed: int = 1) -> Instance:
    """1000 x 1000, costs U{1..1000}, supplies/demands U{1..10000}, balanced."""
    return random_instance(np.random.default_rng(seed), 1000, 1000, 1000, 10000, name=f"plain1000_s{seed}")


def degenerate(seed: int = 1, m: int = 4000, n: int = 2000) -> Instance:
    """m x n with costs from {1..8}; every supply 50, every demand 100 (m=2n balances)."""
    rng = np.random.default_rng(seed)
    cost = rng.integers(1, 9, size=(m, n), dtype=np.int32)
    supply = np.full(m, 50, np.int32)
    demand = np.full(n, 50 * m // n, np.int32)
    assert int(supply.sum()) == int(demand.sum())
    return Instance(cost, supply, demand, 0, f"degenerate_{m}x{n}_s{seed}")


def log_uniform_sizes(rng: np.random.Generator, count: int, lo: int, hi: int) -> np.ndarray:
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(np.int64)


def batch(seed: int = 1, count: int = 4096, lo: int = 16, hi: int = 512,
          cmax: int = 100, smax: int = 1000) -> list[Instance]:
    """`count` instances, m and n log-uniform in [lo, hi] independently."""
    rng = np.random.default_rng(seed)
    ms = log_uniform_sizes(rng, count, lo, hi)
    ns = log_uniform_sizes(rng, count, lo, hi)
    return [random_instance(rng, int(m), int(n), cmax, smax, name=f"batch{k}")
            for k, (m, n) in enumerate(zip(ms, ns))]


def stream_plan(seed: int = 1, count: int = 20000, rate: float = 1000.0, lo: int = 32, hi: int = 2000):
    """Arrival times (Poisson at `rate` per second, in seconds from 0) and sizes of a
    request stream, m and n log-uniform in [lo, hi] independently."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    ms = log_uniform_sizes(rng, count, lo, hi)
    ns = log_uniform_sizes(rng, count, lo, hi)
    return t, ms, ns


def stream_request(seed: int, k: int, m: int, n: int, cmax: int = 1000, smax: int = 10000) -> Instance:
    """Request k of the stream with seed `seed`: its own generator (seed, k), so requests
    can be built in any order or in parallel and rebuilt later."""
    return random_instance(np.random.default_rng([seed, k]), int(m), int(n), cmax, smax, name=f"req{k}")


def stream(seed: int = 1, count: int = 20000, rate: float = 1000.0, lo: int = 32, hi: int = 2000,
           cmax: int = 1000, smax: int = 10000):
    """Poisson arrivals at `rate` per second; yields (arrival_time_s, Instance)."""
    t, ms, ns = stream_plan(seed, count, rate, lo, hi)
    for k in range(count):
        yield float(t[k]), stream_request(seed, k, ms[k], ns[k], cmax, smax)


# ---------------------------------------------------------------------------
# .dat files
# ---------------------------------------------------------------------------

def _scale_decimals(tokens: list[str], max_scale: int = 6) -> tuple[np.ndarray, int]:
    decs = [Decimal(t) for t in tokens]
    for k in range(max_scale + 1):
        f = Decimal(10) ** k
        scaled = [d * f for d in decs]
        if all(s == s.to_integral_value() for s in scaled):
            arr = np.array([int(s) for s in scaled], dtype=np.int64)
            if arr.size and (arr.max() > np.iinfo(np.int32).max or arr.min() < np.iinfo(np.int32).min):
                raise ValueError("scaled costs exceed int32")
            return arr.astype(np.int32), k
    raise ValueError(f"costs are not integral at any scale 10^0..10^{max_scale}")


def scale_float32_costs(costs: np.ndarray, max_scale: int = 6) -> tuple[np.ndarray, int]:
    """Float costs (the reference's `float* costs`, structs.h:35) to exact int32: the
    smallest k such that every cost equals float32(N / 10**k) for an integer N.  Same
    rule as flock_scale_float_costs in the C ABI; never rounds silently."""
    c = np.asarray(costs, dtype=np.float32).ravel()
    for k in range(max_scale + 1):
        f = 10.0 ** k
        N = np.rint(c.astype(np.float64) * f)
        if np.all(np.abs(N) < 2 ** 31 - 1) and np.array_equal((N / f).astype(np.float32), c):
            return N.astype(np.int32), k
    raise ValueError(f"float costs have no exact decimal scale <= 10^{max_scale}")


def read_dat(path: str, max_scale: int = 6, allow_float32: bool = True) -> Instance:
    """Read a `.dat` file.  Costs are scaled exactly as decimals; if that needs more than
    `max_scale` digits and `allow_float32`, they are read the way the reference reads
    them (float, instance.cpp:127-135) and scaled with the float32 rule."""
    import gzip
    import os
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as fh:
        tok = fh.read().split()
    m, n = int(tok[0]), int(tok[1])
    p = 2
    supply = np.array([int(Decimal(t)) for t in tok[p:p + m]], np.int32); p += m
    demand = np.array([int(Decimal(t)) for t in tok[p:p + n]], np.int32); p += n
    meta = {"cost_parse": "decimal"}
    try:
        cost, k = _scale_decimals(tok[p:p + m * n], max_scale)
    except ValueError:
        if not allow_float32:
            raise
        cost, k = scale_float32_costs(np.array([float(t) for t in tok[p:p + m * n]], np.float32), max_scale)
        meta["cost_parse"] = "float32"
    name = os.path.basename(path)
    return Instance(cost.reshape(m, n), supply, demand, k, name[:-3] if name.endswith(".gz") else name, meta)


def write_dat(inst: Instance, path: str) -> None:
    with open(path, "w") as fh:
        fh.write(f"{inst.m} {inst.n}\n")
        fh.write("\t".join(str(int(x)) for x in inst.supply) + "\n")
        fh.write("\t".join(str(int(x)) for x in inst.demand) + "\n")
        div = 10 ** inst.scale
        for row in inst.cost:
            if inst.scale:
                fh.write("\n".join(str(Decimal(int(c)) / div) for c in row) + "\n")
            else:
                fh.write("\n".join(str(int(c)) for c in row) + "\n")
