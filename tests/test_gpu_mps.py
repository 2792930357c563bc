"""MPS file in, LP answer by PDLP on the GPU, --solution file out: the reference CLI's path
(apps/cli/main.cpp: read_mps, solve, write_solution) through this library's C ABI.

Known answers: tests/instances/tiny_lp.mps has objective 11 at x = 3, y = 1
(tests/instances/tiny_lp.mps:1-7, tests/CMakeLists.txt:48-51); the LP relaxations of
mrpl_crude_small.mps and neos-2657525-crna.mps against HiGHS (tests/golden/mps_reference.json).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import samaya_amd as smy

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
RELAX = json.loads((GOLDEN / "mps_reference.json").read_text())["relaxations"]


def test_tiny_lp_file_to_solution_file(gpu, tmp_path):
    print("MPS: tiny_lp.mps, FP64 PDLP", flush=True)
    md = smy.read_mps(GOLDEN / "mps" / "tiny_lp.mps")
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-8))
    assert r.status == smy.OPTIMAL and r.verified
    assert abs(r.objective - 11.0) <= 1e-6
    assert np.allclose(r.col_value, [3.0, 1.0], atol=1e-6)
    path = tmp_path / "tiny.sol"
    smy.write_solution(md, r, path)
    lines = path.read_text().splitlines()
    assert lines[0] == "status optimal" and lines[1] == "verified yes"
    assert abs(float(lines[2].split()[1]) - 11.0) <= 1e-6
    assert lines[4:6] == ["columns 2", "# name value reduced_cost"]
    x = {ln.split()[0]: float(ln.split()[1]) for ln in lines[6:8]}
    assert abs(x["x"] - 3.0) <= 1e-6 and abs(x["y"] - 1.0) <= 1e-6
    assert lines[9:11] == ["rows 2", "# name activity dual"]


@pytest.mark.parametrize("name", ["mrpl_crude_small", "neos-2657525-crna"])
def test_instance_relaxations_match_highs(gpu, name):
    print(f"MPS: {name}.mps LP relaxation, FP64 PDLP", flush=True)
    md = smy.read_mps(GOLDEN / "mps" / f"{name}.mps")
    ref = RELAX[name]
    assert int(md.col_integer.sum()) == ref["integers"]
    r = smy.solve_lp(md, smy.PdlpOptions(tol=1e-8))
    assert r.status == smy.OPTIMAL
    assert abs(r.objective - ref["highs_objective"]) <= 1e-6 * (1 + abs(ref["highs_objective"]))
    # The answer is a valid bound on the relaxation (weak duality from its duals).
    if md.maximize:
        assert r.dual_bound >= ref["highs_objective"] - 1e-6 * (1 + abs(ref["highs_objective"]))
    else:
        assert r.dual_bound <= ref["highs_objective"] + 1e-6 * (1 + abs(ref["highs_objective"]))
