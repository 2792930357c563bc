"""The branch and bound's root bound propagation (smy_mip_propagate) on the CPU: closed forms
(integer rounding, negative coefficients, one-sided rows given their activity range, crossing
bounds), validity against every integer point of small MIPs enumerated by brute force, and the
reference's MIP instance mrpl_crude_small.mps coming out with every column boxed."""
import itertools
from pathlib import Path

import numpy as np
import pytest

import samaya_amd as smy


def model(rows, lo, up, rlo, rup, obj=None):
    """A model from a dense row matrix."""
    A = np.asarray(rows, float)
    m, n = A.shape
    start, index, value = [0], [], []
    for j in range(n):
        for i in range(m):
            if A[i, j] != 0.0:
                index.append(i)
                value.append(A[i, j])
        start.append(len(index))
    return smy.Model(m, n, np.array(start, np.int64), np.array(index, np.int32), np.array(value),
                     np.zeros(n) if obj is None else np.asarray(obj, float), np.asarray(lo, float),
                     np.asarray(up, float), np.asarray(rlo, float), np.asarray(rup, float), False, 0.0)


def close(a, b, tol=1e-7):
    return np.all(np.abs(np.asarray(a) - np.asarray(b)) <= tol * (1 + np.abs(np.asarray(b))))


def test_closed_forms():
    inf = np.inf
    # x1 + x2 <= 4, x >= 0: each x_j <= 4; the row's missing lower bound is its least activity 0.
    lo, up, rlo, rup, crossed = smy.propagate_bounds(model([[1, 1]], [0, 0], [inf, inf], [-inf], [4]))
    assert not crossed and close(up, [4, 4]) and close(lo, [0, 0]) and close(rlo, [0]) and close(rup, [4])
    assert np.all(up >= 4) and rlo[0] <= 0  # relaxed outward, never inward
    # 3 x <= 10, x >= 0 integer: x <= floor(10 / 3) = 3 exactly.
    lo, up, _, _, crossed = smy.propagate_bounds(model([[3]], [0], [inf], [-inf], [10]), [True])
    assert not crossed and up[0] == 3.0
    # -2 x + y >= -4, x >= 0, 0 <= y <= 1: -2 x >= -5, so x <= 2.5 (continuous), 2 (integer).
    lo, up, _, _, _ = smy.propagate_bounds(model([[-2, 1]], [0, 0], [inf, 1], [-4], [inf]))
    assert close(up[0], 2.5) and up[0] >= 2.5
    lo, up, _, _, _ = smy.propagate_bounds(model([[-2, 1]], [0, 0], [inf, 1], [-4], [inf]), [True, False])
    assert up[0] == 2.0
    # 2 x1 + 2 x2 = 1 with binary x: x_j <= 1/2 rounds to 0, then x_j >= 1/2 rounds to 1: crossed.
    _, _, _, _, crossed = smy.propagate_bounds(model([[2, 2]], [0, 0], [1, 1], [1], [1]), [True, True])
    assert crossed
    # The same row over continuous columns has points (x = 1/4 each): not crossed.
    _, _, _, _, crossed = smy.propagate_bounds(model([[2, 2]], [0, 0], [1, 1], [1], [1]))
    assert not crossed


@pytest.mark.parametrize("seed", range(6))
def test_every_integer_point_stays_inside(seed):
    # Random 3-row MIPs over 4 integer columns in [-2, 3] (two with an open side): every integer
    # point that satisfies the rows lies within the propagated column bounds, its activities within
    # the new row bounds; when the bounds cross, there is no such point.
    rng = np.random.default_rng(seed)
    A = rng.integers(-3, 4, size=(3, 4)).astype(float)
    lo = np.array([-2.0, -2.0, -np.inf, -2.0])
    up = np.array([3.0, np.inf, 3.0, 3.0])
    act = A @ rng.integers(-2, 4, size=4)
    rlo = act - rng.integers(0, 3, size=3)
    rup = act + rng.integers(0, 3, size=3)
    rlo[0] = -np.inf
    md = model(A, lo, up, rlo, rup)
    nlo, nup, nrlo, nrup, crossed = smy.propagate_bounds(md, [True] * 4)
    box = [range(-8, 9)] * 4  # wide enough to hold every point the rows admit here
    points = [np.array(p, float) for p in itertools.product(*box)
              if np.all(np.array(p) >= lo) and np.all(np.array(p) <= up)
              and np.all(A @ np.array(p, float) >= rlo) and np.all(A @ np.array(p, float) <= rup)]
    if crossed:
        assert not points
        return
    assert np.all(nlo >= lo) and np.all(nup <= up)  # only ever tighter
    for p in points:
        assert np.all(p >= nlo) and np.all(p <= nup), (p, nlo, nup)
        a = A @ p
        assert np.all(a >= nrlo) and np.all(a <= nrup)


def test_reference_mip_instance_is_boxed():
    md = smy.read_mps(Path(__file__).parent / "golden" / "mps" / "mrpl_crude_small.mps")
    unboxed = ~(np.isfinite(md.col_lower) & np.isfinite(md.col_upper))
    one_sided = ~(np.isfinite(md.row_lower) & np.isfinite(md.row_upper))
    assert unboxed.sum() == 84 and one_sided.sum() == 72
    lo, up, rlo, rup, crossed = smy.propagate_bounds(md)  # the file's integrality markers
    assert not crossed
    assert np.all(np.isfinite(lo)) and np.all(np.isfinite(up))
    assert np.all(np.isfinite(rlo)) and np.all(np.isfinite(rup))
    assert np.all(lo >= md.col_lower) and np.all(up <= md.col_upper)
