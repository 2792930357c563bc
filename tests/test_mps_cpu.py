"""The MPS reader and the MPS / solution writers (include/samaya_io.h) on the CPU.

Parity: tests/golden/mps_reference.json holds the reference's own read_mps_from_string
(src/io/mps_reader.cpp, built from its sources by oracle/Makefile) on the reference's instance
files (tests/instances, copied to tests/golden/mps), on texts covering tests/test_mps_reader.cpp
and further edge cases, and on two generated models as this library writes them (array hashes,
malformed lines planted deep in the section that is parsed on several threads). Every array, name,
integrality mark, error message and line number must match exactly. Quadratic sections (the
haverly files' QCMATRIX, QUADOBJ) are outside the LP path: there the reference returns a QP and
this reader raises.
"""
import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

import samaya_amd as smy

GOLDEN = Path(__file__).resolve().parent / "golden"
FX = json.loads((GOLDEN / "mps_reference.json").read_text())
CASES = {c["name"]: c for c in FX["cases"]}


def _f(a):
    return np.array([float(t) for t in a], np.float64)


def _data(case) -> bytes:
    return (GOLDEN / case["file"]).read_bytes() if "file" in case else case["text"].encode()


def _assert_same(md, case):
    rm = case["model"]
    assert (md.m, md.n, md.nnz) == (rm["m"], rm["n"], rm["start"][-1])
    assert int(md.maximize) == rm["maximize"]
    assert md.offset == rm["offset"]
    assert np.array_equal(md.start, rm["start"])
    assert np.array_equal(md.index, rm["index"])
    assert np.array_equal(md.value, _f(rm["value"]))
    for k in ("obj", "col_lower", "col_upper", "row_lower", "row_upper"):
        assert np.array_equal(getattr(md, k), _f(rm[k])), k
    assert md.name == case["model_name"]
    assert (md.row_names or []) == case["row_names"]
    assert (md.col_names or []) == case["col_names"]
    assert md.col_integer.tolist() == case["col_integer"]


@pytest.mark.parametrize("name", [c["name"] for c in FX["cases"]])
def test_reader_matches_the_reference(name):
    case = CASES[name]
    if case["ok"] and case["quadratic"]:
        with pytest.raises(smy.MpsParseError, match="not supported by the LP path"):
            smy.read_mps_text(_data(case))
        return
    if not case["ok"]:
        with pytest.raises(smy.MpsParseError) as err:
            smy.read_mps_text(_data(case))
        assert str(err.value) == case["message"]
        assert err.value.line == case["line"]
        return
    _assert_same(smy.read_mps_text(_data(case)), case)
    if "file" in case:  # the file entry point gives the same model
        _assert_same(smy.read_mps(GOLDEN / case["file"]), case)


def test_tiny_lp_known_values():
    # tests/test_mps_reader.cpp:17-33 and tests/instances/tiny_lp.mps.
    md = smy.read_mps(GOLDEN / "mps" / "tiny_lp.mps")
    assert md.name == "TINY_LP" and md.maximize
    assert (md.m, md.n, md.nnz) == (2, 2, 4)
    assert md.obj.tolist() == [3.0, 2.0]
    assert md.row_lower[0] == -math.inf and md.row_upper.tolist() == [4.0, 6.0]
    assert md.col_upper.tolist() == [3.0, math.inf]


def test_conventions_known_values():
    # The rules of mps_reader.cpp:12-19 on small texts, checked against hand-derived answers.
    md = smy.read_mps_text(CASES["row_types_ranges_offset"]["text"])
    assert md.offset == 2.5
    assert md.row_lower.tolist() == [10.0, 6.0, 7.0, 10.0]
    assert md.row_upper.tolist() == [14.0, 10.0, 10.0, 13.0]
    assert md.nnz == 4  # the entry on the second N row is dropped
    b = smy.read_mps_text(CASES["bound_types"]["text"])
    inf = math.inf
    assert b.col_lower.tolist() == [0, -2, 3.5, -inf, -inf, -inf, 0, 2, 1, -1, -inf]
    assert b.col_upper.tolist() == [5, inf, 3.5, inf, inf, -1, 1, 9, inf, -0.5, inf]
    assert b.col_integer.tolist() == [0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0]
    d = smy.read_mps_text(CASES["duplicates_and_repeats"]["text"])
    assert d.col_names == ["x", "y", "z"]  # x listed again after y stays column 0
    assert d.obj.tolist() == [1.25, 3.0, 0.0]
    assert d.value.tolist() == [1.625, 1e-300]  # x: 1 + 0.5 + 0.125; y: 2 - 2 dropped


def test_missing_file():
    with pytest.raises(OSError, match="cannot open"):
        smy.read_mps("/nonexistent/file.mps")


def test_utf8_path_with_spaces(tmp_path):
    # tests/test_mps_reader.cpp:39-55
    path = tmp_path / "samaya मॉडल test.mps"
    path.write_text("NAME UNI\nROWS\n N obj\n L c1\nCOLUMNS\n x obj 1 c1 1\n y obj 2 c1 1\n"
                    "RHS\n rhs c1 4\nENDATA\n", encoding="utf-8")
    md = smy.read_mps(path)
    assert (md.m, md.n) == (1, 2)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def powerlaw_text():
    g = FX["generated"]["powerlaw_20k"]
    md = smy.Model.generate(2, 77, 10_000, 20_000, 400, 8.0)
    text = smy.mps_text(md).encode()
    # The writer is deterministic: the text the reference parsed.
    assert len(text) == g["text_bytes"] and hashlib.sha256(text).hexdigest() == g["text_sha"]
    return md, text, g


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_generated_model_matches_the_reference_on_any_thread_count(powerlaw_text, threads):
    md0, text, g = powerlaw_text
    md = smy.read_mps_text(text, threads=threads)
    assert (md.m, md.n, md.nnz) == (g["m"], g["n"], g["nnz"])
    for k, h in g["sha"].items():
        assert _sha(getattr(md, k)) == h, k
    # ... and it is the model that was written (no names: R0.. / C0..).
    for k in ("start", "index", "value", "obj", "col_lower", "col_upper", "row_lower", "row_upper"):
        assert np.array_equal(getattr(md, k), getattr(md0, k)), k


@pytest.mark.parametrize("threads", [1, 8])
def test_errors_deep_in_a_parallel_section_match_the_reference(powerlaw_text, threads):
    _, text, g = powerlaw_text
    lines = text.split(b"\n")
    for inj in g["injections"]:
        k = inj["index"]
        bad = b"\n".join(lines[:k] + [inj["line"].encode()] + lines[k + 1:])
        if inj["ok"]:
            md = smy.read_mps_text(bad, threads=threads)
            assert _sha(md.value) == inj["sha_value"]
            assert int(md.col_integer.sum()) == inj["col_integer_sum"]
        else:
            with pytest.raises(smy.MpsParseError) as err:
                smy.read_mps_text(bad, threads=threads)
            assert (str(err.value), err.value.line) == (inj["message"], inj["error_line"])


def _roundtrip_equal(a, b):
    assert (a.m, a.n, a.nnz, a.maximize, a.offset) == (b.m, b.n, b.nnz, b.maximize, b.offset)
    for k in ("start", "index", "value", "obj", "col_lower", "col_upper", "row_lower", "row_upper"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


@pytest.mark.parametrize("name", ["tiny_lp", "mrpl_crude_small", "neos-2657525-crna",
                                  "row_types_ranges_offset", "bound_types", "integer_markers",
                                  "objname_and_sets", "empty_column_and_row", "crossed_bounds"])
def test_write_then_read_gives_the_same_model(name):
    case = CASES[name]
    md = smy.read_mps_text(_data(case))
    back = smy.read_mps_text(smy.mps_text(md))
    _roundtrip_equal(back, md)
    assert back.row_names == md.row_names and back.col_names == md.col_names
    assert np.array_equal(back.col_integer, md.col_integer)


def test_writer_on_generated_models_and_awkward_rows(tmp_path):
    for md in (smy.generators.c1(), smy.generators.blending(4003), smy.generators.c5(4)[0]):
        path = tmp_path / "m.mps"
        smy.write_mps(md, path)
        _roundtrip_equal(smy.read_mps(path), md)
    # Rows of every kind, ranged rows whose ends do not reproduce exactly from one side, empty
    # columns, columns with upper < 0 and a zero lower bound, an objective offset.
    inf = math.inf
    rl = np.array([1.0, -inf, 2.0, 0.1, -inf, 5.0, 1e-3])
    ru = np.array([1.0, 3.0, inf, 0.7, inf, 5.0 + 2 ** -40, 1e22])
    A = np.array([[1, 0, 2, 0], [0, 1, 0, 0], [3, 0, 0, 0], [1, 1, 0, 0], [0, 0, 1, 0],
                  [0, 1, 0, 0], [1, 0, 0, 0]], float)
    md = smy.Model.from_dense(A, [1, 0, 0, -2.5], [0, -inf, 0, -3], [inf, inf, -1, inf], rl, ru,
                              maximize=True, offset=0.3)
    back = smy.read_mps_text(smy.mps_text(md))
    for k in ("start", "index", "value", "obj", "col_lower", "col_upper"):
        assert np.array_equal(getattr(back, k), getattr(md, k)), k
    # A ranged row's far end is rhs +- |range|: exact where either end reproduces the other, so
    # here everywhere but the last row (1e-3, 1e22), whose lower end moves by less than an ulp of
    # the range.
    assert np.array_equal(back.row_upper, md.row_upper)
    assert np.array_equal(back.row_lower[:-1], md.row_lower[:-1])
    assert abs(back.row_lower[-1] - md.row_lower[-1]) <= np.spacing(1e22)
    assert back.offset == md.offset and back.maximize


def test_names_survive_the_native_model():
    md = smy.read_mps_text(CASES["fixed_format_blank_names"]["text"])
    assert md.col_names == ["X ONE", "X TWO"] and md.row_names == ["LIM 1", "MY ROW"]
    nm = md.native()  # a copy made from the arrays keeps the names and marks
    text = smy.mps_text(nm)
    # Names with blanks cannot be written in free MPS: generated ones replace them.
    assert "C0" in text and "X ONE" not in text and "'INTORG'" in text


def _result(status, objective=float("nan"), verified=True, x=None, d=None, act=None, y=None,
            cert=None):
    z = np.zeros(0)
    return smy.LpResult(status=status, verified=verified, objective=objective,
                        col_value=x if x is not None else z, row_dual=y if y is not None else z,
                        col_dual=d if d is not None else z,
                        row_activity=act if act is not None else z, certificate=cert,
                        max_primal_violation=0.0, max_dual_violation=0.0, iterations=0,
                        fp32_iterations=0, rounds=1, primal_residual=0.0, dual_residual=0.0,
                        gap=0.0, setup_seconds=0.0, solve_seconds=0.0, kernel_ms=0.0, message="")


def test_solution_file_format(tmp_path):
    # apps/cli/main.cpp:137-198: header lines, then name value [reduced_cost] per column and name
    # activity [dual] per row, numbers with 17 significant digits.
    md = smy.read_mps(GOLDEN / "mps" / "tiny_lp.mps")
    r = _result(smy.OPTIMAL, 11.0, x=np.array([3.0, 1.0]), d=np.array([1.0 / 3, 0.0]),
                act=np.array([4.0, 6.0]), y=np.array([2.5, 0.1]))
    text = smy.solution_text(md, r)
    assert text == ("status optimal\nverified yes\nobjective 11\n\ncolumns 2\n"
                    "# name value reduced_cost\nx 3 0.33333333333333331\ny 1 0\n\nrows 2\n"
                    "# name activity dual\nc1 4 2.5\nc2 6 0.10000000000000001\n")
    path = tmp_path / "sol.txt"
    smy.write_solution(md, r, path)
    assert path.read_text() == text
    # Without names (c0.. / r0..), not verified, the activity from A x when not given.
    plain = smy.Model(md.m, md.n, md.start, md.index, md.value, md.obj, md.col_lower,
                      md.col_upper, md.row_lower, md.row_upper, md.maximize, md.offset)
    r2 = _result(smy.OPTIMAL, 11.0, verified=False, x=np.array([3.0, 1.0]))
    r2.row_activity = None
    assert smy.solution_text(plain, r2) == ("status optimal\nverified no\nobjective 11\n\n"
                                            "columns 2\n# name value\nc0 3\nc1 1\n\nrows 2\n"
                                            "# name activity\nr0 4\nr1 6\n")


def test_solution_file_certificates():
    md = smy.read_mps(GOLDEN / "mps" / "tiny_lp.mps")
    inf = _result(smy.INFEASIBLE, x=np.array([0.5, 0.0]), act=np.array([0.5, 0.5]),
                  y=np.array([9.0, 9.0]), d=np.array([9.0, 9.0]), cert=np.array([1.0, -2e-7]))
    assert smy.solution_text(md, inf) == (
        "status infeasible\nverified yes\n\ncolumns 2\n# name value\nx 0.5\ny 0\n\nrows 2\n"
        "# name activity\nc1 0.5\nc2 0.5\n\ninfeasibility_certificate 2\nc1 1\nc2 -1.9999999999999999e-07\n")
    unb = _result(smy.UNBOUNDED, cert=np.array([1.0, 2.0]))
    assert smy.solution_text(md, unb, with_values=False) == (
        "status unbounded\nverified yes\n\nunbounded_ray 2\nx 1\ny 2\n")
    lim = _result(smy.ITERATION_LIMIT, 5.0, verified=False)
    assert smy.solution_text(md, lim, with_values=False) == "status iteration_limit\nverified no\n"
    # No answer, no point: values passed along are not written.
    err = _result(smy.NUMERICAL_ERROR, verified=False, x=np.array([7.0, 7.0]),
                  act=np.array([7.0, 7.0]))
    assert smy.solution_text(md, err) == "status numerical_error\nverified no\n"
