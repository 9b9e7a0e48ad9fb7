"""The reference's eleven benchmark objectives in the expression IR.

CPU: the IR of each objective equals its formula written independently
here in 256-bit mpmath (from the reference code lists cited in
gpugo_amd/families.py), at random points of its box; the closed-form
minimum is where the table says it is (f(x*) = f*, and no random point
is below f*).
GPU (marked): interval enclosures on random boxes contain the 256-bit
values at random points, gradient enclosures contain mpmath derivatives,
and certified solves contain f*.
"""
import math

import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import families as F
from tests.test_host import ir_eval

mp.mp.prec = 256


def f_ref(name, x):
    """Independent 256-bit definitions."""
    x = [v if isinstance(v, mp.mpf) else mp.mpf(float(v)) for v in x]
    n = len(x)
    pi = mp.pi
    if name == "ackley":
        return (-20 * mp.exp(mp.mpf(-0.02) * mp.sqrt(sum(v * v for v in x) / n))
                - mp.exp(sum(mp.cos(2 * pi * v) for v in x) / n) + 20 + mp.e)
    if name == "belegundu":
        s = sum((v - 5) ** 2 for v in x)
        return mp.mpf(0.1) * s - mp.cos(5 * mp.sqrt(s))
    if name == "breiman":
        return mp.mpf(-0.1) * sum(mp.cos(5 * pi * v) for v in x) + sum(v * v for v in x)
    if name == "fu":
        return 1 + sum(8 * mp.sin(7 * (v - mp.mpf(0.9)) ** 2) ** 2 + 6 * mp.sin(14 * (v - mp.mpf(0.9)) ** 2) ** 2
                       + (v - mp.mpf(0.9)) ** 2 for v in x)
    if name == "griewank":
        p = mp.mpf(1)
        for i, v in enumerate(x):
            p *= mp.cos(v / mp.sqrt(i + 1))
        return 1 + sum(v * v for v in x) / 4000 - p
    if name == "levy":
        w = lambda v: mp.sin(pi / 4 * v + 3 * pi / 4) ** 2
        s = 10 * w(x[0]) + sum(((x[i] - 1) / 4) ** 2 * (1 + 10 * w(x[i + 1])) for i in range(n - 1))
        return pi / n * (s + ((x[-1] - 1) / 4) ** 2)
    if name == "rastrigin":
        return 10 * n + sum(v * v - 10 * mp.cos(2 * pi * v) for v in x)
    if name == "salomon":
        r = mp.sqrt(sum(v * v for v in x))
        return 1 - mp.cos(2 * pi * r) + mp.mpf(0.1) * r
    if name == "styblinski":
        p = mp.mpf(1)
        for v in x:
            p *= mp.cos(v)
        return sum(v * v for v in x) / (2 * n) - 4 * n * p
    if name == "zabinsky":
        p1, p2 = mp.mpf(1), mp.mpf(1)
        for v in x:
            p1 *= mp.sin(v - pi / 6)
            p2 *= mp.sin(5 * (v - pi / 6))
        return -(mp.mpf(2.5) * p1 + p2)
    if name == "rosenbrock":
        return sum(100 * (x[i + 1] - x[i] ** 2) ** 2 + (x[i] - 1) ** 2 for i in range(n - 1))
    raise KeyError(name)


@pytest.mark.parametrize("name", F.names())
@pytest.mark.parametrize("d", [2, 5])
def test_ir_matches_formula(name, d):
    obj = F.objective(name, d)
    obj.check()
    fam = F.REFERENCE[name]
    rng = np.random.default_rng(d)
    for _ in range(8):
        x = rng.uniform(fam.lo, fam.hi, d)
        got = ir_eval(obj.value, [mp.mpf(float(v)) for v in x], [])
        want = f_ref(name, x)
        assert abs(got - want) <= mp.mpf("1e-60") * (1 + abs(want)), (name, x)


@pytest.mark.parametrize("name", F.names())
def test_closed_form_minimum(name):
    d = 3
    fam = F.REFERENCE[name]
    xs = fam.x_star(d)
    fx = f_ref(name, xs)
    # x* given in double: f there equals f* up to the rounding of x*
    assert abs(fx - fam.f_star(d)) < 1e-12, (name, fx)
    rng = np.random.default_rng(0)
    vals = [f_ref(name, rng.uniform(fam.lo, fam.hi, d)) for _ in range(300)]
    assert min(vals) >= fam.f_star(d) - 1e-12


@pytest.mark.parametrize("name", sorted(F.TEMPLATES))
@pytest.mark.parametrize("d", [2, 3, 6])
def test_loop_templates_match_formula(name, d):
    """The loop templates (term-parallel and windowed code lists) are the
    same function as the definition, and their gradient component j is
    mpmath's derivative of the definition (256-bit)."""
    tpl = F.TEMPLATES[name](d)
    fam = F.REFERENCE[name]
    rng = np.random.default_rng(10 + d)
    for _ in range(4):
        # away from the kinks where a sqrt of a sum of squares vanishes
        xf = rng.uniform(fam.lo, fam.hi, d)
        x = [mp.mpf(float(v)) for v in xf]
        want = f_ref(name, x)
        got = ir_eval(tpl["value"], x, [])
        assert abs(got - want) <= mp.mpf("1e-60") * (1 + abs(want)), (name, xf)
        for j in range(d):
            def fj(s, j=j):
                y = list(x)
                y[j] = s
                return f_ref(name, y)
            dw = mp.diff(fj, x[j])
            dg = ir_eval(tpl["grad"], x, [], {}, j)
            assert abs(dg - dw) <= mp.mpf("1e-40") * (1 + abs(dw)), (name, j, xf)
