"""The benchmark instance families, by name, with their fixed seeds and sizes.

  c1  sparse LP, m 2,000 x n 5,000, 10 nnz per column at uniform rows, N(0,1)      seed 1001
  c2  sparse LP, m 500,000 x n 1,000,000, power-law column counts on [2, 5,000]
      with mean 12, values N(0,1) x 10^U[-2,2] (nnz ~ 12M)                          seed 1002
Both: half the rows equalities, half <=, feasible around x0 ~ U[0,5]; c ~ U[-1,1]; 0 <= x <= 10.
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
