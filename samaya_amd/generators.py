"""The benchmark instance families, by name, with their fixed seeds and sizes.

  c1  sparse LP, m 2,000 x n 5,000, 10 nnz per column at uniform rows, N(0,1)      seed 1001
  c2  sparse LP, m 500,000 x n 1,000,000, power-law column counts on [2, 5,000]
      with mean 12, values N(0,1) x 10^U[-2,2] (nnz ~ 12M)                          seed 1002
Both: half the rows equalities, half <=, feasible around x0 ~ U[0,5]; c ~ U[-1,1]; 0 <= x <= 10.
  c5  knapsack LP, m 20,000 x n 50,000, 8 weights U{1..100} per column, capacities 0.4 x row
      sums, 0 <= x <= 1, profits U{1..100} (maximize)                                 seed 5000
      with 1,024 bound vectors fixing 1..5% of the columns at 0 or 1                seeds 6000 + k
  c4  256 multi-period blending LPs (periods U{4..48}, 20 units x 30 streams tanks, qualities
      U{1..30}; 2.6k..60k rows each)                                                  seeds 4000 + i
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .lp import Model, _f64, _p


def sparse_uniform(seed: int, m: int, n: int, per_col: int) -> Model:
    return Model.generate(N.SMY_GEN_SPARSE_UNIFORM, seed, m, n, per_col)


def sparse_powerlaw(seed: int, m: int, n: int, kmax: int, mean: float) -> Model:
    return Model.generate(N.SMY_GEN_SPARSE_POWERLAW, seed, m, n, kmax, mean)


def c1() -> Model:
    return sparse_uniform(1001, 2000, 5000, 10)


def c2(scale: float = 1.0) -> Model:
    """The 1M-column LP; scale < 1 shrinks m and n (same structure) for quick runs."""
    m, n = int(500_000 * scale), int(1_000_000 * scale)
    return sparse_powerlaw(1002, m, n, min(5000, m), 12.0)


def blending(seed: int, periods: int = 0, units: int = 20, streams: int = 30,
             qualities: int = 0) -> Model:
    """Multi-period blending LP (staircase mass balances, capacity and quality rows);
    periods / qualities 0 draw them from U{4..48} / U{1..30}."""
    return Model.generate(N.SMY_GEN_BLENDING, seed, periods, units, streams, float(qualities))


def c4(count: int = 256) -> list[Model]:
    """The batched workload: `count` blending LPs with seeds 4000, 4001, ..."""
    return [blending(4000 + i) for i in range(count)]


def knapsack(seed: int, m: int, n: int, per_col: int) -> Model:
    """Multi-dimensional knapsack LP (maximize): weights and profits U{1..100}, capacities 0.4 x
    row sums, 0 <= x <= 1."""
    return Model.generate(N.SMY_GEN_KNAPSACK, seed, m, n, per_col)


def fixings(seed: int, lower, upper, frac_min: float = 0.01, frac_max: float = 0.05):
    """A bound vector fixing a fraction U[frac_min, frac_max] of the columns at a bound."""
    lo, up = _f64(lower), _f64(upper)
    out_lo, out_up = np.empty_like(lo), np.empty_like(up)
    N.check(N.load().smy_generate_fixings(seed, len(lo), frac_min, frac_max, _p(lo), _p(up),
                                          _p(out_lo), _p(out_up)), "generate_fixings")
    return out_lo, out_up


def c5(count: int = 1024, scale: float = 1.0):
    """The node-relaxation workload: the knapsack LP (seed 5000; scale < 1 shrinks m and n) and
    `count` bound vectors (seeds 6000 + k)."""
    m, n = int(20_000 * scale), int(50_000 * scale)
    md = knapsack(5000, m, n, 8)
    bounds = [fixings(6000 + k, md.col_lower, md.col_upper) for k in range(count)]
    return md, [b[0] for b in bounds], [b[1] for b in bounds]
