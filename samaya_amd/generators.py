"""The benchmark instance families, by name, with their fixed seeds and sizes.

  c1  sparse LP, m 2,000 x n 5,000, 10 nnz per column at uniform rows, N(0,1)      seed 1001
  c2  sparse LP, m 500,000 x n 1,000,000, power-law column counts on [2, 5,000]
      with mean 12, values N(0,1) x 10^U[-2,2] (nnz ~ 12M)                          seed 1002
Both: half the rows equalities, half <=, feasible around x0 ~ U[0,5]; c ~ U[-1,1]; 0 <= x <= 10.
  c4  256 multi-period blending LPs (periods U{4..48}, 20 units x 30 streams tanks, qualities
      U{1..30}; 2.6k..60k rows each)                                                  seeds 4000 + i
"""
from __future__ import annotations

from . import _native as N
from .lp import Model


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
