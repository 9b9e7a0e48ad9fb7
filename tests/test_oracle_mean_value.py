"""Pins the oracle's mean-value mode (ORACLE_MEAN_VALUE, family_mv_lower).

The mean-value bound is f(c) + sum_j G_j (X_j - c_j) with c the box midpoint
(rounded to nearest) and G the gradient enclosure over the box; by the mean
value theorem its lower end bounds f from below on the box.  The reference
has no such bound (it prunes with the natural extension only,
1-Ackley/ackley50.py:116-220), so it is pinned by mathematics:

  - closed form: for f = sum a_j x_j^2 (the separable family with b = 0) the
    gradient enclosure is 2 a_j X_j, so the bound is f(c) plus
    sum_j min of the four endpoint products of [2 a_j lo_j, 2 a_j hi_j] and
    [lo_j - c_j, hi_j - c_j]; computed here in exact rational arithmetic, the
    oracle's outward-rounded bound must lie at or below it and within
    rounding slack of it (a dropped term, a wrong index or a midpoint gradient
    moves it far outside the slack);
  - rigor: on random boxes of every family the bound is at most f evaluated
    in 256-bit mpmath at the midpoint, the corners and random points;
  - end to end: mean-value solves contain the closed-form f* = 0 (separable,
    Rastrigin, Ackley, Rosenbrock) and, for the d = 10 chain, intersect the
    plain solve whose enclosure is pinned by an L-BFGS witness in
    test_oracle_bnb.
"""
from fractions import Fraction as F
import itertools

import mpmath as mp
import numpy as np
import pytest
from scipy.optimize import minimize as sp_minimize

from gpugo_amd import problems as pb
from tests.test_oracle_objectives import f_true, random_boxes

mp.mp.prec = 256
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


def midpoint(lo, hi):
    """The oracle's midpoint: RN((RN(hi - lo) / 2) + lo)."""
    return (hi - lo) / 2.0 + lo


def mv_exact_quadratic(a, lo, hi):
    """Exact mean-value lower bound of sum a_j x_j^2 over [lo, hi]."""
    total = F(0)
    for aj, l, h in zip(a, lo, hi):
        c = midpoint(l, h)
        total += F(aj) * F(c) ** 2
        g = (2 * F(aj) * F(l), 2 * F(aj) * F(h))
        dx = (F(l) - F(c), F(h) - F(c))
        total += min(gi * di for gi in g for di in dx)
    return total


@pytest.mark.parametrize("d", [1, 2, 5, 12])
def test_mv_lower_closed_form_quadratic(oracle, d):
    rng = np.random.default_rng(100 + d)
    a = rng.uniform(0.5, 2.0, d)
    params = np.concatenate([a, np.zeros(d), rng.uniform(1.0, 6.0, d)])  # b = 0: f = sum a x^2
    lo = rng.uniform(-6, 4, (40, d))
    hi = lo + rng.uniform(0, 3, (40, d)) * rng.uniform(0, 1, (40, 1)) ** 2
    got = oracle.mv_lower(0, params, lo, hi)
    for i in range(lo.shape[0]):
        ex = mv_exact_quadratic(a, lo[i], hi[i])
        scale = sum(F(float(aj)) * (abs(F(l)) + abs(F(h))) ** 2 for aj, l, h in zip(a, lo[i], hi[i]))
        assert F(got[i]) <= ex, i                              # outward: never above the exact bound
        assert ex - F(got[i]) <= F(1e-13) * (1 + scale), i     # and within rounding slack of it


def test_mv_lower_closed_form_catches_dropped_term(oracle):
    """Teeth: leaving out one coordinate's gradient term moves the closed
    form by far more than the slack the test above allows."""
    d = 3
    a = np.array([1.0, 1.5, 0.75])
    lo, hi = np.array([[-1.0, 0.5, -2.0]]), np.array([[0.5, 2.0, -1.0]])
    ex = mv_exact_quadratic(a, lo[0], hi[0])
    wrong = ex - min(F(2 * a[1] * g) * dx for g in (lo[0, 1], hi[0, 1])
                     for dx in (F(lo[0, 1]) - F(midpoint(lo[0, 1], hi[0, 1])),
                                F(hi[0, 1]) - F(midpoint(lo[0, 1], hi[0, 1]))))
    assert abs(wrong - ex) > F(1e-3)
    got = oracle.mv_lower(0, np.concatenate([a, np.zeros(d), np.ones(d)]), lo, hi)
    assert abs(F(got[0]) - ex) < F(1e-12) and abs(F(got[0]) - wrong) > F(1e-4)


SPECS = [pb.sepcos2(), pb.chain10(), pb.ackley50(d=6), pb.rastrigin(3), pb.rosenbrock(4), pb.ackley(5)]


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: s.name)
def test_mv_lower_below_truth(oracle, spec):
    rng = np.random.default_rng(7)
    lo, hi = random_boxes(rng, spec, 40, max_frac=0.1)
    mv = oracle.mv_lower(FAM[spec.family], spec.params, lo, hi)
    for i in range(lo.shape[0]):
        pts = [midpoint(lo[i], hi[i])]
        corners = list(itertools.product(*zip(lo[i], hi[i])))
        pts += [np.array(c) for c in corners[:8]]
        pts += [lo[i] + rng.uniform(0, 1, spec.d) * (hi[i] - lo[i]) for _ in range(3)]
        for x in pts:
            assert mp.mpf(mv[i]) <= f_true(spec.family, spec.params, x), i


@pytest.mark.parametrize("spec,grid,k", [
    (pb.sepcos2(), (2, 2), 16),
    (pb.rastrigin(3), (2, 3), 64),
    (pb.ackley(4), (2, 4), 64),
    (pb.rosenbrock(4), (2, 4), 64),
], ids=lambda v: getattr(v, "name", str(v)))
def test_mean_value_solves_contain_f_star(oracle, spec, grid, k):
    r, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=grid[0],
                        cycle=grid[1], k_select=k, max_iters=20000, mean_value=True)
    assert r["status"] == 1
    assert r["lb"] <= 0.0 <= r["ub"] and r["ub"] - r["lb"] <= 1e-6


def test_chain10_mean_value_intersects_witness_pinned_plain(oracle):
    spec = pb.chain10()
    kw = dict(tol_f=1e-6, wid=2, cycle=5, k_select=256, max_iters=5000)
    mv, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, mean_value=True, **kw)
    plain, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, **kw)
    assert mv["status"] == 1 and plain["status"] == 1
    assert max(mv["lb"], plain["lb"]) <= min(mv["ub"], plain["ub"])
    # the witness of test_oracle_bnb: no lower bound above a point value
    rng = np.random.default_rng(0)
    f = lambda x: float(f_true(spec.family, spec.params, x))
    best = None
    for _ in range(20):
        res = sp_minimize(f, rng.uniform(spec.lo, spec.hi), method="L-BFGS-B", bounds=list(zip(spec.lo, spec.hi)))
        if best is None or res.fun < best.fun:
            best = res
    fw = f_true(spec.family, spec.params, best.x)
    assert mv["lb"] <= float(fw) + 1e-12 and mv["ub"] <= float(fw) + 1e-6
