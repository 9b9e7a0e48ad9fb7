"""Pins the oracle's branch-and-bound loop (1-Ackley/ackley50.py:1033-1380).

End-to-end checks against closed forms and independent witnesses: the
certified enclosure [lb, ub] must contain f* (0 for the separable and
Ackley-type families at x = 0; for the d = 10 chain a local minimum found by
scipy from many starts and evaluated in mpmath is a witness upper bound),
and must be no wider than the tolerance.  The reference's x-width output
rule must return boxes that contain the minimiser.
"""
import numpy as np
import pytest
from scipy.optimize import minimize as sp_minimize

from gpugo_amd import problems as pb
from tests.test_oracle_objectives import f_true

FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


def solve(oracle, spec, **kw):
    r, out = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, **kw)
    return r, out


def test_config0_certified_contains_zero(oracle):
    spec = pb.sepcos2()
    for grid, k in [((64, 2), 1), ((2, 2), 16)]:
        r, _ = solve(oracle, spec, tol_f=1e-6, wid=grid[0], cycle=grid[1], k_select=k, max_iters=20000)
        assert r["status"] == 1
        assert r["lb"] <= 0.0 <= r["ub"]
        assert r["ub"] - r["lb"] <= 1e-6


def test_config1_certified_contains_witness(oracle):
    spec = pb.chain10()
    r, _ = solve(oracle, spec, tol_f=1e-6, wid=2, cycle=5, k_select=256, max_iters=5000)
    assert r["status"] == 1 and r["ub"] - r["lb"] <= 1e-6
    # witness: best of many local minimisations, value in mpmath
    rng = np.random.default_rng(0)
    f = lambda x: float(f_true(spec.family, spec.params, x))
    best = None
    for _ in range(40):
        x0 = rng.uniform(spec.lo, spec.hi)
        res = sp_minimize(f, x0, method="L-BFGS-B", bounds=list(zip(spec.lo, spec.hi)))
        if best is None or res.fun < best.fun:
            best = res
    fw = f_true(spec.family, spec.params, best.x)
    assert r["lb"] <= float(fw) + 1e-12          # lb is a lower bound of every f value
    assert r["ub"] <= float(fw) + 1e-6           # the enclosure reaches the best witness


@pytest.mark.parametrize("spec", [pb.rastrigin(2), pb.rosenbrock(3), pb.ackley(3), pb.ackley50(d=4)],
                         ids=lambda s: s.name)
def test_reference_objectives_f_star_zero(oracle, spec):
    r, _ = solve(oracle, spec, tol_f=1e-6, wid=2, cycle=spec.d, k_select=64, max_iters=20000)
    assert r["status"] == 1
    assert r["lb"] <= 0.0 <= r["ub"] and r["ub"] - r["lb"] <= 1e-6


def test_k_select_changes_path_not_result(oracle):
    spec = pb.sepcos2()
    a, _ = solve(oracle, spec, tol_f=1e-6, wid=4, cycle=2, k_select=1, max_iters=20000)
    b, _ = solve(oracle, spec, tol_f=1e-6, wid=4, cycle=2, k_select=32, max_iters=20000)
    assert a["status"] == b["status"] == 1
    assert max(a["lb"], b["lb"]) <= min(a["ub"], b["ub"])


def test_x_width_rule_outputs_contain_minimiser(oracle):
    """Reference rule (1-Ackley/ackley50.py:1306-1343): boxes narrower than
    tol_x go to the output list; the run ends when the list is empty."""
    spec = pb.rastrigin(2)
    r, (olo, ohi, olb) = solve(oracle, spec, tol_f=0.0, tol_x=1e-3, wid=4, cycle=2, k_select=8,
                               max_iters=100000)
    assert r["status"] == 2 and r["n_out"] >= 1
    hit = [(np.all(olo[i] <= 0) and np.all(ohi[i] >= 0)) for i in range(len(olb))]
    assert any(hit)
    assert np.all(ohi - olo < 1e-3)
    assert r["lb"] <= 0.0 <= r["ub"]


def test_deterministic(oracle):
    spec = pb.chain10()
    a, _ = solve(oracle, spec, tol_f=1e-6, wid=2, cycle=5, k_select=64, max_iters=40, threads=8)
    b, _ = solve(oracle, spec, tol_f=1e-6, wid=2, cycle=5, k_select=64, max_iters=40, threads=3)
    for k in ("lb", "ub", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt", "queue_size"):
        assert a[k] == b[k], k


def test_levy_solve_certifies(oracle):
    """The oracle's loop on its Levy restatement (d = 3, the reference box
    and initial upper bound, 10 diagonal samples) certifies f* = 0."""
    d = 3
    o, _ = oracle.solve(3, np.zeros(0), np.full(d, -10.0), np.full(d, 10.0), tol_f=1e-6, wid=3, cycle=2,
                        k_select=16, n_samples=10, init_ub=1000.0, max_iters=20000)
    assert o["status"] == 1 and o["lb"] <= 0.0 <= o["ub"] and o["ub"] - o["lb"] <= 1e-6
