"""Pins the oracle's objective code lists (value, gradient, monotonicity
test, sample bound) for the three coefficient families and Levy.

Reference code lists: 1-Ackley/ackley50.py:116-479,
7-Rastrigin/rastrigin50.py:110-372, 11-Rosenbrock/rosenbrock50.py:115-451.
Checks: the value enclosure contains f evaluated in 256-bit mpmath at
points inside the box (f written here from the mathematical definitions of
the benchmark configurations and of the reference's Rastrigin, Rosenbrock
and Ackley, not from the oracle's code); the gradient enclosure contains
mpmath's numerical derivative; closed-form minima (f(0) = 0 for the
separable and Ackley-type families); the monotonicity test never removes a
box containing the minimiser and does remove boxes where the gradient has
constant sign away from the initial box boundary.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import problems as pb

mp.mp.prec = 256
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


def f_true(family, params, x):
    """The objective in mpmath, from its mathematical definition."""
    d = len(x)
    x = [v if isinstance(v, mp.mpf) else mp.mpf(float(v)) for v in x]
    P = [mp.mpf(float(v)) for v in params]
    if family in ("sepcos", "sepcos_chain"):
        a, b, c = P[:d], P[d:2 * d], P[2 * d:3 * d]
        f = mp.fsum(a[i] * x[i] ** 2 + b[i] * (1 - mp.cos(c[i] * x[i])) for i in range(d))
        if family == "sepcos_chain":
            w = P[3 * d:]
            f += mp.fsum(w[i] * (x[i + 1] - x[i] ** 2) ** 2 + (1 - x[i]) ** 2 for i in range(d - 1))
        return f
    A, B, C = P[0], P[1], P[2]
    q = P[3:]
    m1 = mp.fsum(v ** 2 for v in x) / d
    m2 = mp.fsum(mp.cos(C * v) for v in x) / d
    return (-A * mp.exp(-B * mp.sqrt(m1)) - mp.exp(m2) + A + mp.e
            + mp.fsum(q[i] * x[i] ** 4 for i in range(d)) / 50)


def reference_fn(name, x):
    """Rastrigin / Rosenbrock / Ackley as the reference scripts define them."""
    x = [mp.mpf(float(v)) for v in x]
    n = len(x)
    if name == "rastrigin":   # 7-Rastrigin/rastrigin50.py:167-188
        return 10 * n + mp.fsum(v ** 2 - 10 * mp.cos(2 * mp.pi * v) for v in x)
    if name == "rosenbrock":  # 11-Rosenbrock/rosenbrock50.py:170-198
        return mp.fsum(100 * (x[i + 1] - x[i] ** 2) ** 2 + (x[i] - 1) ** 2 for i in range(n - 1))
    # 1-Ackley/ackley50.py:171-212
    return (-20 * mp.exp(-mp.mpf("0.02") * mp.sqrt(mp.fsum(v ** 2 for v in x) / n))
            - mp.exp(mp.fsum(mp.cos(2 * mp.pi * v) for v in x) / n) + 20 + mp.e)


def random_boxes(rng, spec, n, max_frac=0.3):
    d = spec.d
    c = rng.uniform(spec.lo, spec.hi, (n, d))
    w = (spec.hi - spec.lo) * max_frac * rng.uniform(0, 1, (n, d)) * rng.uniform(0, 1, (n, 1)) ** 4
    return np.maximum(c - w, spec.lo), np.minimum(c + w, spec.hi)


SPECS = [pb.sepcos2(), pb.chain10(), pb.ackley50(d=6), pb.rastrigin(3), pb.rosenbrock(4), pb.ackley(5)]


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: s.name)
def test_value_and_gradient_contain_truth(oracle, spec):
    rng = np.random.default_rng(11)
    lo, hi = random_boxes(rng, spec, 60)
    r = oracle.eval_boxes(FAM[spec.family], spec.params, lo, hi, spec.lo, spec.hi, 1)
    for i in range(lo.shape[0]):
        for t in range(3):
            x = lo[i] + rng.uniform(0, 1, spec.d) * (hi[i] - lo[i])
            fx = f_true(spec.family, spec.params, x)
            assert mp.mpf(r["flo"][i]) <= fx <= mp.mpf(r["fhi"][i]), (i, t)
            if t == 0:
                for j in range(spec.d):
                    def fj(s, j=j):
                        y = list(x)
                        y[j] = s
                        return f_true(spec.family, spec.params, y)
                    dj = mp.diff(fj, mp.mpf(float(x[j])))
                    assert mp.mpf(r["glo"][i, j]) - mp.mpf("1e-40") <= dj <= mp.mpf(r["ghi"][i, j]) + mp.mpf("1e-40")
        # sample bound: >= f at the midpoint
        mid = lo[i] + (hi[i] - lo[i]) / 2
        assert mp.mpf(r["ub"][i]) >= f_true(spec.family, spec.params, mid)


@pytest.mark.parametrize("name,spec", [("rastrigin", pb.rastrigin(3)), ("rosenbrock", pb.rosenbrock(4)),
                                       ("ackley", pb.ackley(5))])
def test_reference_objectives_are_family_members(oracle, name, spec):
    """The families reproduce the reference's Rastrigin, Rosenbrock and
    Ackley code lists as functions (same values at random points)."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        x = rng.uniform(spec.lo, spec.hi)
        r = oracle.eval_boxes(FAM[spec.family], spec.params, x[None], x[None], spec.lo, spec.hi, 1)
        fr = reference_fn(name, x)
        # the reference uses the double nearest to 2*pi as the coefficient;
        # the difference to exact 2*pi is < 1e-15 relative
        tol = mp.mpf("1e-12") * (1 + abs(fr))
        assert mp.mpf(r["flo"][0]) - tol <= fr <= mp.mpf(r["fhi"][0]) + tol


@pytest.mark.parametrize("spec", [pb.sepcos2(), pb.ackley50(d=50), pb.rastrigin(4), pb.ackley(8)],
                         ids=lambda s: s.name)
def test_closed_form_minimum_zero(oracle, spec):
    z = np.zeros((1, spec.d))
    r = oracle.eval_boxes(FAM[spec.family], spec.params, z, z, spec.lo, spec.hi, 1)
    assert r["flo"][0] <= 0.0 <= r["fhi"][0]
    assert r["fhi"][0] - r["flo"][0] < 1e-12
    # a box around the minimiser is never pruned by the monotonicity test
    lo, hi = np.full((1, spec.d), -1e-3), np.full((1, spec.d), 2e-3)
    r = oracle.eval_boxes(FAM[spec.family], spec.params, lo, hi, spec.lo, spec.hi, 1)
    assert r["kkt"][0] == 0
    assert r["flo"][0] <= 0.0


def test_kkt_prunes_monotone_boxes(oracle):
    spec = pb.sepcos2()
    # far from 0 the quadratic dominates: gradient has constant sign; a box
    # touching the initial boundary on the low side must not be pruned when
    # f increases away from that boundary
    lo, hi = np.array([[3.0, 3.0]]), np.array([[3.1, 3.1]])
    r = oracle.eval_boxes(0, spec.params, lo, hi, spec.lo, spec.hi, 1)
    g_sign = np.sign(r["glo"][0]) == np.sign(r["ghi"][0])
    assert g_sign.any() and r["kkt"][0] == 1
    lo2 = np.array([[-5.12, 3.0]])
    hi2 = np.array([[-5.0, 3.1]])
    r2 = oracle.eval_boxes(0, spec.params, lo2, hi2, spec.lo, spec.hi, 1)
    # coordinate 0 decreases toward the interior (gradient < 0): prunable
    assert r2["ghi"][0, 0] < 0 and r2["kkt"][0] == 1
    lo3, hi3 = np.array([[-5.12, -5.12]]), np.array([[-5.12, -5.12]])
    # the corner point itself: f decreasing into the box on both axes but the
    # box touches the boundary only on the low side -> gradient < 0 prunes
    r3 = oracle.eval_boxes(0, spec.params, lo3, hi3, spec.lo, spec.hi, 1)
    assert r3["kkt"][0] == 1


@pytest.mark.parametrize("d", [2, 3, 5])
def test_levy_restatement_contains_truth(oracle, d):
    """Family 3 (Levy, 6-Levy/levy50.py:166-224) against the independent
    256-bit definition: value and gradient enclosures on random boxes of the
    reference box [-10, 10]^d contain f and mpmath's derivative at random
    points, the sample bound is at least f at the midpoint, and the closed-
    form minimum f(1, ..., 1) = 0 is enclosed with a gradient containing 0."""
    from tests.test_families import f_ref
    rng = np.random.default_rng([12, d])
    n = 40
    c = rng.uniform(-10, 10, (n, d))
    w = 20 * 0.3 * rng.uniform(0, 1, (n, d)) * rng.uniform(0, 1, (n, 1)) ** 4
    lo, hi = np.maximum(c - w, -10.0), np.minimum(c + w, 10.0)
    r = oracle.eval_boxes(3, np.zeros(0), lo, hi, np.full(d, -10.0), np.full(d, 10.0), 1)
    mp.mp.prec = 256
    for i in range(n):
        for t in range(3):
            x = lo[i] + rng.uniform(0, 1, d) * (hi[i] - lo[i])
            assert mp.mpf(r["flo"][i]) <= f_ref("levy", x) <= mp.mpf(r["fhi"][i]), (i, t)
            if t == 0:
                for j in range(d):
                    def fj(s, j=j):
                        y = [mp.mpf(float(v)) for v in x]
                        y[j] = s
                        return f_ref("levy", y)
                    dj = mp.diff(fj, mp.mpf(float(x[j])))
                    assert r["glo"][i, j] - 1e-30 <= dj <= r["ghi"][i, j] + 1e-30, (i, j)
        assert mp.mpf(r["ub"][i]) >= f_ref("levy", lo[i] + (hi[i] - lo[i]) / 2)
    one = np.ones((1, d))
    z = oracle.eval_boxes(3, np.zeros(0), one, one, np.full(d, -10.0), np.full(d, 10.0), 1)
    assert z["flo"][0] <= 0.0 <= z["fhi"][0] and z["fhi"][0] < 1e-25
    assert np.all(z["glo"][0] <= 0.0) and np.all(z["ghi"][0] >= 0.0)
