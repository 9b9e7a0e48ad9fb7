"""Pins the oracle's interval operations.

Reference: "Updated Interval Arithmetic Operations/IA.py":57-570.  The
reference has no tests or golden vectors for this path; the oracle is
pinned here by exact rational arithmetic (+ - * / sqrt must equal the
true round-down / round-up results bit for bit), by 256-bit mpmath ranges
for exp/sin/cos (containment plus a width bound), and by closed forms for
the extended-division and even-power special cases.
"""
import math

import numpy as np
import pytest

from tests._exact import (F, exact_div, exact_mul, random_intervals, rd, ru, sqrt_rd, sqrt_ru,
                          trig_range, ulps_between, mp)

OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "neg": 4, "scale": 5, "addc": 6, "rsubc": 7,
       "divc": 8, "sqr": 9, "pown": 10, "sqrt": 11, "exp": 12, "sin": 13, "cos": 14}


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(1234)
    alo, ahi = random_intervals(rng, 1500)
    blo, bhi = random_intervals(rng, 1500)
    blo, bhi = blo[::-1].copy(), bhi[::-1].copy()
    return alo, ahi, blo, bhi


def test_rounding_selftest(oracle):
    assert oracle.rounding_selftest()


def test_add_sub_exact(oracle, data):
    alo, ahi, blo, bhi = data
    lo, hi = oracle.ia_eval(OPS["add"], alo, ahi, blo, bhi)
    for i in range(alo.size):
        assert lo[i] == rd(F(alo[i]) + F(blo[i])) and hi[i] == ru(F(ahi[i]) + F(bhi[i])), i
    lo, hi = oracle.ia_eval(OPS["sub"], alo, ahi, blo, bhi)
    for i in range(alo.size):
        assert lo[i] == rd(F(alo[i]) - F(bhi[i])) and hi[i] == ru(F(ahi[i]) - F(blo[i])), i


def test_mul_exact(oracle, data):
    alo, ahi, blo, bhi = data
    lo, hi = oracle.ia_eval(OPS["mul"], alo, ahi, blo, bhi)
    for i in range(alo.size):
        assert (lo[i], hi[i]) == exact_mul(alo[i], ahi[i], blo[i], bhi[i]), i


def test_div_exact_and_extended(oracle, data):
    alo, ahi, blo, bhi = data
    lo, hi = oracle.ia_eval(OPS["div"], alo, ahi, blo, bhi)
    n = 0
    for i in range(alo.size):
        if blo[i] > 0 or bhi[i] < 0:
            assert (lo[i], hi[i]) == exact_div(alo[i], ahi[i], blo[i], bhi[i]), i
            n += 1
    assert n > 500
    # textbook extended division (IA.py:154-207 by the mathematics)
    cases = [((1.0, 2.0), (0.0, 4.0), (0.25, math.inf)),
             ((-2.0, -1.0), (0.0, 4.0), (-math.inf, -0.25)),
             ((1.0, 2.0), (-4.0, 0.0), (-math.inf, -0.25)),
             ((-2.0, -1.0), (-4.0, 0.0), (0.25, math.inf)),
             ((-1.0, 2.0), (0.0, 4.0), (-math.inf, math.inf)),
             ((1.0, 2.0), (-1.0, 1.0), (-math.inf, math.inf)),
             ((0.0, 2.0), (0.0, 4.0), (0.0, math.inf))]
    for (a, b, want) in cases:
        got = oracle.ia_eval(OPS["div"], [a[0]], [a[1]], [b[0]], [b[1]])
        assert (got[0][0], got[1][0]) == want, (a, b)


def test_scalar_ops_exact(oracle, data):
    alo, ahi, _, _ = data
    for s in (2.5, -2.5, 3.0, -0.1, 1e-3):
        lo, hi = oracle.ia_eval(OPS["scale"], alo, ahi, s=s)
        for i in range(alo.size):
            ps = [F(alo[i]) * F(s), F(ahi[i]) * F(s)]
            assert (lo[i], hi[i]) == (rd(min(ps)), ru(max(ps))), (i, s)
        lo, hi = oracle.ia_eval(OPS["addc"], alo, ahi, s=s)
        for i in range(alo.size):
            assert (lo[i], hi[i]) == (rd(F(alo[i]) + F(s)), ru(F(ahi[i]) + F(s)))
        lo, hi = oracle.ia_eval(OPS["rsubc"], alo, ahi, s=s)
        for i in range(alo.size):
            assert (lo[i], hi[i]) == (rd(F(s) - F(ahi[i])), ru(F(s) - F(alo[i])))
        lo, hi = oracle.ia_eval(OPS["divc"], alo, ahi, s=s)
        for i in range(alo.size):
            qs = [F(alo[i]) / F(s), F(ahi[i]) / F(s)]
            assert (lo[i], hi[i]) == (rd(min(qs)), ru(max(qs)))
    lo, hi = oracle.ia_eval(OPS["scale"], alo, ahi, s=0.0)
    assert np.all(lo == 0) and np.all(hi == 0)


def test_sqr_exact(oracle, data):
    alo, ahi, _, _ = data
    lo, hi = oracle.ia_eval(OPS["sqr"], alo, ahi)
    for i in range(alo.size):
        a, b = F(alo[i]), F(ahi[i])
        sq = [a * a, b * b]
        want_lo = 0.0 if alo[i] < 0 < ahi[i] else rd(min(sq))
        assert lo[i] == want_lo and hi[i] == ru(max(sq)), i
    # IA.py:267-281 returns -2.5e-323 as the lower bound of [0,h]^2; the
    # mathematics gives exactly 0
    lo, hi = oracle.ia_eval(OPS["sqr"], [0.0], [3.0])
    assert lo[0] == 0.0 and hi[0] == 9.0


def test_pown_contains_and_tight(oracle, data):
    alo, ahi, _, _ = data
    keep = (np.abs(alo) < 1e60) & (np.abs(ahi) < 1e60) & ((np.abs(alo) > 1e-60) | (alo == 0))
    alo, ahi = alo[keep], ahi[keep]
    for n in (0, 1, 3, 4, 5, 8):
        lo, hi = oracle.ia_eval(OPS["pown"], alo, ahi, s=n)
        for i in range(alo.size):
            a, b = F(alo[i]), F(ahi[i])
            vals = [a ** n, b ** n] + ([F(0)] if (alo[i] < 0 < ahi[i] and n % 2 == 0 and n > 0) else [])
            tlo, thi = min(vals), max(vals)
            assert F(lo[i]) <= tlo and F(hi[i]) >= thi, (n, i)
            # one rounding per multiplication: at most ~n ulps outside
            if tlo != 0:
                assert abs(ulps_between(lo[i], rd(tlo))) <= 2 * n
            if thi != 0:
                assert abs(ulps_between(ru(thi), hi[i])) <= 2 * n


def test_sqrt_exact(oracle):
    rng = np.random.default_rng(7)
    a = np.sort(np.abs(rng.standard_normal((800, 2))) * 10.0 ** rng.uniform(-30, 30, (800, 1)), axis=1)
    a[:4] = [[0.0, 0.0], [0.0, 4.0], [1.0, 2.0], [2.0 ** -1070, 2.0 ** -1060]]
    lo, hi = oracle.ia_eval(OPS["sqrt"], a[:, 0], a[:, 1])
    for i in range(a.shape[0]):
        assert lo[i] == (0.0 if a[i, 0] == 0 else sqrt_rd(a[i, 0])), i
        assert hi[i] == sqrt_ru(a[i, 1]), i
    # partly negative argument: domain restricted to [0, h]
    lo, hi = oracle.ia_eval(OPS["sqrt"], [-1.0], [4.0])
    assert (lo[0], hi[0]) == (0.0, 2.0)


@pytest.mark.parametrize("fn", ["exp", "sin", "cos"])
def test_transcendental_contains_true_range(oracle, fn):
    rng = np.random.default_rng({"exp": 1, "sin": 2, "cos": 3}[fn])
    n = 600
    c = rng.uniform(-40, 40, n) if fn != "exp" else rng.uniform(-700, 700, n)
    w = 10.0 ** rng.uniform(-14, 0.8, n)
    lo_in, hi_in = c - w, c + w
    # intervals straddling extrema k*pi/2 and degenerate points
    k = np.arange(-20, 20)
    lo_in[:40] = k * math.pi / 2 - 1e-9
    hi_in[:40] = k * math.pi / 2 + 1e-9
    lo_in[40:60] = hi_in[40:60] = c[40:60]
    lo, hi = oracle.ia_eval(OPS[fn], lo_in, hi_in)
    for i in range(n):
        if fn == "exp":
            tlo, thi = mp.exp(mp.mpf(lo_in[i])), mp.exp(mp.mpf(hi_in[i]))
        else:
            tlo, thi = trig_range(lo_in[i], hi_in[i], fn)
        assert mp.mpf(lo[i]) <= tlo and mp.mpf(hi[i]) >= thi, (fn, i, lo_in[i], hi_in[i])
        # width bound: at most 4 ulps beyond the true range (2-ulp widening
        # of a <= 1-ulp nearest value, plus the conservative extremum test);
        # sin/cos of arguments narrower than 2^-40 use one evaluation widened
        # by the argument width (Lipschitz bound), which may add that width
        w = hi_in[i] - lo_in[i]
        extra = w if (fn != "exp" and w <= 2.0 ** -40) else 0.0
        if tlo != -1 and tlo != 0:
            assert ulps_between(lo[i], float(tlo)) <= 4 or float(tlo) - lo[i] <= 2 * extra, (fn, i)
        if thi != 1 and math.isfinite(hi[i]):
            assert ulps_between(float(thi), hi[i]) <= 4 or hi[i] - float(thi) <= 2 * extra, (fn, i)


def test_checker_catches_nearest_rounding():
    """A planted mistake (round-to-nearest instead of directed) must fail the
    exact check: guards against a vacuous comparison."""
    rng = np.random.default_rng(5)
    a = rng.standard_normal(200)
    b = rng.standard_normal(200)
    wrong = a + b  # nearest, not round-down
    mismatches = sum(wrong[i] != rd(F(a[i]) + F(b[i])) for i in range(200))
    assert mismatches > 50


def test_digits15_restatement():
    """The checker's restatement of outRound (the IA behind the paper's
    results, 1-Ackley/IAAckley.py:106-141) on closed-form cases: decimals of
    15 significant digits as doubles, tiny magnitudes widened to 0 or
    -/+1e-292; on random values the bounds lie within 4 ulps of a decimal of
    15 significant digits (the reference scales by inexact powers of ten) and
    within one ulp of the outside of x (the scaling may land an ulp inside)."""
    from oracle.oracle import digits15 as g
    assert g(-2.5, False) == -2.5 and g(-2.5, True) == -2.5 and g(0.0, False) == 0.0
    assert g(math.pi, False) == 3.14159265358979 and g(math.pi, True) == 3.1415926535898
    assert g(-math.pi, False) == -3.1415926535898 and g(-math.pi, True) == -3.14159265358979
    assert g(1 / 3, True) == 0.333333333333334 and g(1 / 3, False) == 0.333333333333333
    assert g(2 / 3, True) == 0.666666666666667 and g(2 / 3, False) == 0.666666666666666
    assert g(123456789012345678.0, False) == 123456789012345000.0
    assert g(123456789012345678.0, True) == 123456789012346000.0
    assert g(1e-300, False) == 0.0 and g(1e-300, True) == 1e-292 and g(-1e-300, False) == -1e-292
    from decimal import Decimal
    rng = np.random.default_rng(3)
    for x in np.concatenate([rng.standard_normal(200) * 10.0 ** rng.integers(-200, 200, 200), [1e23, 9.999e22]]):
        x = float(x)
        lo, hi = g(x, False), g(x, True)
        assert lo <= math.nextafter(x, math.inf) and hi >= math.nextafter(x, -math.inf)
        for b in (lo, hi):   # a 15-digit decimal, up to the scaling's rounding
            d = Decimal(b)
            near = float(d.quantize(Decimal(1).scaleb(d.adjusted() - 14)))
            assert abs(near - b) <= 4 * math.ulp(b), (x, b, near)
