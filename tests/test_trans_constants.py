"""Pins the constants of the device library's exp/sin/cos enclosures
(gpugo_amd/csrc/ia.hpp, tr64): the Cody-Waite splits of pi/2 and ln 2 must
bracket the true constants exactly (the reduced argument is then an
enclosure for any k), their leading parts must carry at most 33
significant bits (so k * c1 is exact for |k| < 2^20), the Taylor coefficients
must be the nearest doubles of (-1)^i / n!, and the a-priori error bounds
must cover the rounding-error analysis they are derived from.  Checked in
exact rational arithmetic and 300-bit mpmath; no GPU needed."""
import math
import os
import re
from fractions import Fraction as F

import mpmath as mp
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "gpugo_amd", "csrc", "ia.hpp")).read()
mp.mp.prec = 300


def const(name):
    m = re.search(rf"\b{name} = ([0-9a-fx.p+\-]+)[;,]", SRC)
    assert m, name
    v = m.group(1)
    return float.fromhex(v) if v.startswith(("0x", "-0x")) else float(v)


def sig_bits(x):
    m, e = math.frexp(x)
    q = F(m)
    bits = 0
    while q.denominator != 1:
        q *= 2
        bits += 1
    return bits


def test_pi_over_2_split_brackets():
    p1, p2 = const("PIO2_1"), const("PIO2_2")
    lo, hi = const("PIO2_3_LO"), const("PIO2_3_HI")
    assert sig_bits(p1) <= 33 and sig_bits(p2) <= 33
    assert lo < hi and math.nextafter(lo, math.inf) == hi
    pio2 = mp.pi / 2
    assert mp.mpf(p1) + mp.mpf(p2) + mp.mpf(lo) < pio2 < mp.mpf(p1) + mp.mpf(p2) + mp.mpf(hi)
    # teeth: the nearest double of pi/2 alone does not bracket the remainder
    assert not (mp.mpf(float(pio2)) == pio2)


def test_ln2_split_brackets():
    l1 = const("LN2_1")
    lo, hi = const("LN2_2_LO"), const("LN2_2_HI")
    assert sig_bits(l1) <= 33
    assert lo < hi and math.nextafter(lo, math.inf) == hi
    ln2 = mp.log(2)
    assert mp.mpf(l1) + mp.mpf(lo) < ln2 < mp.mpf(l1) + mp.mpf(hi)


def test_reduction_guesses():
    # only used to choose k: any value works, these are the nearest doubles
    assert const("TWO_OVER_PI") == float(2 / mp.pi)
    assert const("LOG2E") == float(1 / mp.log(2))


def coeffs(fn):
    body = SRC[SRC.index(f"double {fn}(double r)"):]
    body = body[:body.index("}")]
    return [float(v) for v in re.findall(r"[-+]?\d+\.\d+e[-+]\d+|[-+]?\d+\.\d+", body)]


@pytest.mark.parametrize("fn,expect", [
    ("sin_poly", [F((-1) ** i, math.factorial(2 * i + 1)) for i in range(10, 0, -1)]),
    ("cos_poly", [F((-1) ** i, math.factorial(2 * i)) for i in range(11, 0, -1)]),
    ("exp_poly", [F(1, math.factorial(i)) for i in range(13, 0, -1)]),
])
def test_taylor_coefficients_are_nearest(fn, expect):
    got = coeffs(fn)[:len(expect)]
    assert len(got) == len(expect)
    for g, e in zip(got, expect):
        assert g == float(e), (fn, g, e)   # float(Fraction) rounds to nearest


def test_error_bounds_cover_the_analysis():
    u = 2.0 ** -53
    gamma = lambda n: n * u / (1 - n * u)
    r_sin = math.pi / 4 + 2.0 ** -11
    z = r_sin * r_sin
    # sin: |z Q(z)| and the bound of section 5.1 (fma Horner) + roundings
    zq = sum(z ** i / math.factorial(2 * i + 1) for i in range(1, 11))
    sin_bound = (zq * (gamma(9) + 4 * u) + u)
    assert 3 * sin_bound <= const("EPS_SIN")
    cq = sum(z ** i / math.factorial(2 * i) for i in range(1, 12))
    cos_bound = cq * (gamma(10) + 4 * u) + u
    assert 3 * cos_bound <= const("EPS_COS")
    r_exp = math.log(2) / 2 + 2.0 ** -40
    sa = sum(r_exp ** i / math.factorial(i) for i in range(14))
    exp_bound = sa * (gamma(13) + 2 * u)
    assert 3 * exp_bound <= const("EPS_EXP")
    # truncation terms are far below the bounds
    assert r_sin ** 23 / math.factorial(23) < 1e-24 and r_sin ** 24 / math.factorial(24) < 1e-25
    assert math.exp(r_exp) * r_exp ** 14 / math.factorial(14) < 1e-17


@pytest.mark.parametrize("fn,expect", [
    ("sin_poly_lp", [F((-1) ** i, math.factorial(2 * i + 1)) for i in range(4, 0, -1)]),
    ("cos_poly_lp", [F((-1) ** i, math.factorial(2 * i)) for i in range(5, 0, -1)]),
    ("exp_poly_lp", [F(1, math.factorial(i)) for i in range(7, 0, -1)]),
])
def test_short_kernel_coefficients_are_nearest(fn, expect):
    """The kernels of the FP32 interval functions (endpoints rounded to FP32
    afterwards): the same Taylor series, cut shorter."""
    got = coeffs(fn)[:len(expect)]
    assert len(got) == len(expect)
    for g, e in zip(got, expect):
        assert g == float(e), (fn, g, e)


def test_short_kernel_bounds_cover_the_analysis():
    u = 2.0 ** -53
    gamma = lambda n: n * u / (1 - n * u)
    r_sin = math.pi / 4 + 2.0 ** -11
    z = r_sin * r_sin
    # sin to r^9: truncation |r|^11/11! relative to |r|, plus rounding
    trunc = r_sin ** 10 / math.factorial(11)
    zq = sum(z ** i / math.factorial(2 * i + 1) for i in range(1, 5))
    assert 3 * (trunc + zq * (gamma(3) + 4 * u) + u) <= const("EPS_SIN_LP")
    # cos to r^10
    trunc = r_sin ** 12 / math.factorial(12)
    cq = sum(z ** i / math.factorial(2 * i) for i in range(1, 6))
    assert 3 * (trunc + cq * (gamma(4) + 4 * u) + u) <= const("EPS_COS_LP")
    # exp to r^7 (bound on e^r, which lies in [0.70, 1.42])
    r_exp = math.log(2) / 2 + 2.0 ** -40
    trunc = math.exp(r_exp) * r_exp ** 8 / math.factorial(8)
    sa = sum(r_exp ** i / math.factorial(i) for i in range(8))
    assert 3 * (trunc + sa * (gamma(7) + 2 * u)) <= const("EPS_EXP_LP")
    # far below an FP32 ulp (2^-24 relative), so FP32 endpoints lose at most one ulp
    assert const("EPS_SIN_LP") < 2.0 ** -24 / 8 and const("EPS_EXP_LP") < 0.7 * 2.0 ** -24
    # teeth: one term fewer would not be covered
    assert 3 * r_sin ** 8 / math.factorial(9) > const("EPS_SIN_LP")
