"""Product tiles and their tuner: every tile (1, 2 or 4 row blocks per workgroup, blocks in
launch order or one contiguous range per XCD) gives every sum bit for bit, so the tuner's timing
choice never changes an answer; the tuner picks one of the tiles at the first run of a precision
(timing them on a matrix of at least 2,048 row blocks, the plain tile below that), and one per size
bin of a batch's instances, launched apart when that is faster; and the solver answers equal with
and without tuning."""
import itertools

import numpy as np
import pytest

import samaya_amd as smy
import samaya_amd.device as D
import samaya_amd.generators as G
from test_gpu_batch import _device_lp
from test_gpu_pdhg import transpose_csc

pytestmark = pytest.mark.gpu

TILES = list(itertools.product((1, 2, 4), (False, True)))


def models():
    # Uneven sizes; columns of up to 1,400 nonzeros (rows of A' longer than a block: blocks of
    # their own) and rows of A of ~1,300; an LP with a single row; enough blocks for every tile
    # to leave partial workgroups at the end.
    return [G.sparse_powerlaw(121, 1500, 3000, 1400, 8.0), G.sparse_uniform(122, 1, 5, 1),
            G.blending(123, periods=3, units=3, streams=4, qualities=2), G.sparse_uniform(124, 12, 2000, 8),
            G.sparse_uniform(125, 900, 2100, 7)]


def make_batch(inputs, precision):
    b = D.PdhgBatch.create([(eq.m, eq.n, eq.start, eq.index, eq.value) for _, eq, _, _ in inputs],
                           [transpose_csc(eq) for _, eq, _, _ in inputs],
                           [eq.cost for _, eq, _, _ in inputs], [eq.lower for _, eq, _, _ in inputs],
                           [eq.upper for _, eq, _, _ in inputs], precision)
    for i, (lp, eq, rs, cs) in enumerate(inputs):
        b.set_original(i, cs, rs, lp.cost, lp.lower, lp.upper)
    return b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_tile_gives_the_same_bits(gpu, precision):
    inputs = [_device_lp(md, 130 + i) for i, md in enumerate(models())]
    K = len(inputs)
    results = []
    for per_wg, xcd in TILES:
        b = make_batch(inputs, precision)
        for matrix in (0, 1):
            b.set_tile(matrix, precision, per_wg, xcd)
        assert b.tile(0, precision) == (per_wg, xcd)
        b.run([0.2] * K, [0.3] * K, 1.0, [0] * K, 64)  # a captured run
        b.advance()
        b.run([0.2] * K, [0.3] * K, 1.0, [64] * K, 3)  # single launches
        st = D.PdhgBatch.adaptive_state([0.1] * K, 1.0)
        b.advance()
        b.run_adaptive(st, 7)
        sums = b.evaluate()
        got = [b.download(i) for i in range(K)]
        results.append((sums, got, [(s.eta, s.rejected) for s in st]))
    ref = results[0]
    for sums, got, st in results[1:]:
        np.testing.assert_array_equal(sums, ref[0])
        assert st == ref[2]
        for a, r in zip(got, ref[1]):
            for key in a:
                np.testing.assert_array_equal(a[key], r[key])


def test_tuner_picks_a_tile_at_the_first_run(gpu):
    inputs = [_device_lp(md, 140 + i) for i, md in enumerate(models())]
    K = len(inputs)
    b = make_batch(inputs, "fp64")
    assert b.tile(0, "fp64") is None and b.tile(1, "fp64") is None
    b.run([0.2] * K, [0.3] * K, 1.0, [0] * K, 64)
    for matrix in (0, 1):
        assert b.tile(matrix, "fp64") in TILES
    assert b.tile(0, "fp32") is None  # no instance steps in FP32


def test_tuner_times_the_tiles_of_a_large_matrix(gpu):
    # 2.4M nonzeros: over 2,048 row blocks per matrix, so the tuner times every tile; a matrix of
    # fewer blocks keeps the plain tile untimed (its launches last a few microseconds).
    b = make_batch([_device_lp(G.sparse_uniform(150, 100000, 200000, 12), 151)], "fp32")
    b.run([0.2], [0.3], 1.0, [0], 64)
    for matrix in (0, 1):
        assert b.tile(matrix, "fp32") in TILES
    small = make_batch([_device_lp(G.sparse_uniform(152, 300, 600, 4), 153)], "fp32")
    small.run([0.2], [0.3], 1.0, [0], 64)
    for matrix in (0, 1):
        assert small.tile(matrix, "fp32") == (1, False)


def binned_models():
    # Instances in all three size bins of both matrices: < 64 row blocks, < 512 (120k nonzeros),
    # and more (2.4M nonzeros, so that the batch has enough blocks for the tuner to time tiles).
    return models()[:3] + [G.sparse_uniform(165, 20000, 40000, 3), G.sparse_uniform(166, 100000, 200000, 12)]


def test_size_bins_launched_apart_give_the_same_bits(gpu):
    inputs = [_device_lp(md, 170 + i) for i, md in enumerate(binned_models())]
    K = len(inputs)
    results = []
    for apart in (False, True):
        b = make_batch(inputs, "fp32")
        for matrix in (0, 1):
            b.set_tile(matrix, "fp32", 1, False)
            if apart:
                for size_bin, (pw, xcd) in enumerate([(4, True), (2, False), (1, True)]):
                    b.set_bin_tile(matrix, "fp32", size_bin, pw, xcd)
                assert b.bin_tile(matrix, "fp32", 0) == (4, True, True)
                assert b.bin_tile(matrix, "fp32", 2) == (1, True, True)
            else:
                assert b.bin_tile(matrix, "fp32", 1)[2] is False
        b.run([0.2] * K, [0.3] * K, 1.0, [0] * K, 64)  # a captured run
        b.advance()
        b.run([0.2] * K, [0.3] * K, 1.0, [64] * K, 3)  # single launches
        st = D.PdhgBatch.adaptive_state([0.1] * K, 1.0)
        b.advance()
        b.run_adaptive(st, 7)
        results.append((b.evaluate(), [b.download(i) for i in range(K)], [(s.eta, s.rejected) for s in st]))
    (s0, g0, a0), (s1, g1, a1) = results
    np.testing.assert_array_equal(s1, s0)
    assert a1 == a0
    for x, y in zip(g1, g0):
        for key in x:
            np.testing.assert_array_equal(x[key], y[key], err_msg=key)


def test_tuner_picks_a_tile_per_size_bin(gpu):
    inputs = [_device_lp(md, 180 + i) for i, md in enumerate(binned_models())]
    K = len(inputs)
    b = make_batch(inputs, "fp32")
    b.run([0.2] * K, [0.3] * K, 1.0, [0] * K, 64)
    for matrix in (0, 1):
        assert b.tile(matrix, "fp32") in TILES
        for size_bin in range(3):
            pw, xcd, split = b.bin_tile(matrix, "fp32", size_bin)
            assert (pw, xcd) in TILES


def test_solver_answers_do_not_depend_on_the_tile(gpu):
    # The tuner's pick varies with timing; results must not. Same solve twice on fresh handles.
    md = G.sparse_powerlaw(150, 2000, 5000, 400, 10.0)
    opts = smy.PdlpOptions(tol=1e-6, precision="mixed", algorithm="enhanced")
    a = smy.solve_lp(md, opts)
    b = smy.solve_lp(md, opts)
    assert a.status == b.status == smy.OPTIMAL and a.iterations == b.iterations
    np.testing.assert_array_equal(a.col_value, b.col_value)
