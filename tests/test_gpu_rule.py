"""The restart rule on the device (smy_pdhg_rule_init / smy_pdhg_rule_decide) against the same
rule taken on the host from downloaded sums (the reference's way, src/lp/pdlp.cpp:1075-1132).

Both read the same device totals and use the same formulas in the same order; only exp and log
of the primal-weight update may round differently on the device, so the two solves take the same
decisions (restart iterations and candidates) on nearly every model and reach the same answers."""
import numpy as np
import pytest

import samaya_amd as smy
import samaya_amd.generators as G
from conftest import load_golden
from oracle_lib import M

pytestmark = pytest.mark.gpu


def solve_both(lp, **kw):
    a = smy.solve_pdlp(lp, smy.PdlpOptions(device_restart=True, **kw))
    b = smy.solve_pdlp(lp, smy.PdlpOptions(device_restart=False, **kw))
    return a, b


def test_device_rule_takes_the_host_decisions(gpu):
    same = total = 0
    for c in load_golden("pdlp_random_feasible")["cases"]:
        lp = smy.make_problem(M.from_json(c["model"]).product_model())
        a, b = solve_both(lp, tol=1e-8, max_iterations=100000)
        assert a.status == b.status, c["k"]
        total += 1
        ra = [(it, av) for it, av, _, _ in a.restarts]
        rb = [(it, av) for it, av, _, _ in b.restarts]
        if ra == rb and a.iterations == b.iterations:
            same += 1
            # Same schedule: the candidates' errors and primal weights agree to the rounding the
            # two trajectories picked up (an ulp of a weight update early on grows to ~1e-9
            # relative in errors near 1e-8).
            for (_, _, ka, oa), (_, _, kb, ob) in zip(a.restarts, b.restarts):
                assert abs(ka - kb) <= 1e-6 * kb and abs(oa - ob) <= 1e-6 * ob
        if a.status == smy.OPTIMAL:
            ya, yb = a.x[:lp.n], b.x[:lp.n]
            assert np.max(np.abs(ya - yb)) <= 1e-4 * (1 + np.max(np.abs(yb)))
    assert same >= 0.9 * total, (same, total)


@pytest.mark.parametrize("precision,adaptive", [("fp64", False), ("mixed", False), ("mixed", True)])
def test_device_rule_solves_agree_with_host_rule(gpu, precision, adaptive):
    md = G.sparse_powerlaw(161, 1500, 4000, 300, 9.0)
    lp = smy.make_problem(md)
    a, b = solve_both(lp, tol=1e-6, precision=precision, algorithm="enhanced", adaptive_steps=adaptive)
    assert a.status == b.status == smy.OPTIMAL
    assert len(a.restarts) > 0
    obj_a = float(lp.cost @ a.x)
    obj_b = float(lp.cost @ b.x)
    assert abs(obj_a - obj_b) <= 1e-5 * (1 + abs(obj_b))
    # Restart decisions keep to the rule's own invariants.
    for (it, _, kkt, omega) in a.restarts:
        assert it % 64 == 63 and kkt > 0 and omega > 0


def test_device_rule_in_a_batch_equals_single_solves(gpu):
    models = [G.sparse_uniform(171, 80, 200, 4), G.blending(172, periods=2, units=2, streams=3, qualities=2),
              G.sparse_powerlaw(173, 300, 600, 30, 4.0)]
    opts = smy.PdlpOptions(tol=1e-6, max_iterations=50000, precision="mixed", algorithm="enhanced")
    singles = [smy.solve_lp(md, opts) for md in models]
    batch, _ = smy.solve_lp_batch(models, opts)
    for s, b in zip(singles, batch):
        assert s.status == b.status and s.iterations == b.iterations
        np.testing.assert_array_equal(s.col_value, b.col_value)
    # The third LP stops at the iteration cap in both: a harder draw, equal all the same.
    assert sum(s.status == smy.OPTIMAL for s in singles) >= 2
