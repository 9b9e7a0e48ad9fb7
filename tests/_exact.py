"""Independent references for the interval tests.

Exact rational arithmetic (fractions.Fraction) decides round-down/round-up
results of + - * / sqrt; mpmath at 256 bits gives the exact ranges of
exp/sin/cos (endpoint values plus interior extrema located with a
high-precision pi).  Neither shares code with the oracle or the product.
"""
from __future__ import annotations

import math
from fractions import Fraction as F

import mpmath as mp
import numpy as np

mp.mp.prec = 256
INF = math.inf


_MAXF = F(1.7976931348623157e308)


def rd(q: F) -> float:
    """IEEE round toward -inf of an exact rational."""
    if q > _MAXF:
        return 1.7976931348623157e308
    if q < -_MAXF:
        return -INF
    x = float(q)                       # correctly rounded to nearest
    if F(x) > q:
        x = math.nextafter(x, -INF)
    return x


def ru(q: F) -> float:
    """IEEE round toward +inf of an exact rational."""
    if q > _MAXF:
        return INF
    if q < -_MAXF:
        return -1.7976931348623157e308
    x = float(q)
    if F(x) < q:
        x = math.nextafter(x, INF)
    return x


def sqrt_rd(a: float) -> float:
    """Largest double r with r*r <= a (a >= 0), decided exactly."""
    q = F(a)
    r = math.sqrt(a)
    while F(r) * F(r) > q:
        r = math.nextafter(r, -INF)
    while F(math.nextafter(r, INF)) ** 2 <= q:
        r = math.nextafter(r, INF)
    return r


def sqrt_ru(a: float) -> float:
    q = F(a)
    r = math.sqrt(a)
    while F(r) * F(r) < q:
        r = math.nextafter(r, INF)
    while r > 0 and F(math.nextafter(r, -INF)) ** 2 >= q:
        r = math.nextafter(r, -INF)
    return r


def exact_mul(alo, ahi, blo, bhi):
    ps = [F(x) * F(y) for x in (alo, ahi) for y in (blo, bhi)]
    return rd(min(ps)), ru(max(ps))


def exact_div(alo, ahi, blo, bhi):
    qs = [F(x) / F(y) for x in (alo, ahi) for y in (blo, bhi)]
    return rd(min(qs)), ru(max(qs))


def trig_range(lo: float, hi: float, fn: str):
    """[min, max] of sin or cos over [lo, hi] as mpmath numbers (256 bits)."""
    f = mp.cos if fn == "cos" else mp.sin
    a, b = mp.mpf(lo), mp.mpf(hi)
    vals = [f(a), f(b)]
    # extrema: cos at k*pi, sin at pi/2 + k*pi
    shift = 0 if fn == "cos" else mp.pi / 2
    k0 = int(mp.ceil((a - shift) / mp.pi))
    k1 = int(mp.floor((b - shift) / mp.pi))
    for k in range(k0, min(k1, k0 + 3) + 1):
        vals.append(mp.mpf(1) if k % 2 == 0 else mp.mpf(-1))
    if k1 - k0 >= 2:
        vals += [mp.mpf(1), mp.mpf(-1)]
    return min(vals), max(vals)


def ulps_between(x: float, y: float) -> int:
    """Signed distance in ulps from x to y (finite doubles)."""
    def key(v):
        b = np.frombuffer(np.float64(v).tobytes(), dtype=np.int64)[0]
        return int(b) if b >= 0 else -(int(b) & 0x7FFFFFFFFFFFFFFF)
    return key(y) - key(x)


def random_intervals(rng, n, scale_exp=(-8, 8), with_edges=True):
    """Sorted pairs with magnitudes spread over 10^scale_exp, plus edge cases."""
    mag = 10.0 ** rng.uniform(*scale_exp, (n, 1))
    a = rng.standard_normal((n, 2)) * mag
    a.sort(axis=1)
    lo, hi = a[:, 0].copy(), a[:, 1].copy()
    if with_edges:
        edges = [(0.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (-0.0, 2.5), (1.0, 1.0), (-3.0, -3.0),
                 (0.1, 0.3), (-0.7, 0.2), (2.0 ** -1000, 2.0 ** -999), (-(2.0 ** 900), 2.0 ** 900),
                 (1.0, 1.0 + 2.0 ** -52), (-(1.0 + 2.0 ** -52), -1.0), (5e-324, 1e-323), (-1e-300, 1e-300)]
        k = min(len(edges), n)
        for i in range(k):
            lo[i], hi[i] = edges[i]
    return lo, hi
