"""GPU interval ops (gpugo_ia_eval) against the oracle and exact references.

Bar: EXACT mode equals the oracle's true directed rounding bit for bit for
+ - * / sqrt and the scalar ops (the documented exception, nonzero
products below 2^-968 whose residual may underflow, may be one ulp wider);
FAST mode contains the oracle's result and is at most one ulp wider per
endpoint; exp/sin/cos contain the 256-bit range in both modes and stay
within 4 ulps of it; FP32 results contain the exact result.
"""
import math

import numpy as np
import pytest

from tests._exact import mp, random_intervals, trig_range, ulps_between

pytestmark = pytest.mark.gpu

RATIONAL = ["add", "sub", "mul", "div", "neg", "scale", "addc", "rsubc", "divc", "sqr", "pown", "sqrt"]
OPC = {"add": 0, "sub": 1, "mul": 2, "div": 3, "neg": 4, "scale": 5, "addc": 6, "rsubc": 7,
       "divc": 8, "sqr": 9, "pown": 10, "sqrt": 11, "exp": 12, "sin": 13, "cos": 14}
TINY = 2.0 ** -968


def inputs(op, rng):
    alo, ahi = random_intervals(rng, 3000)
    blo, bhi = random_intervals(rng, 3000)
    blo, bhi = blo[::-1].copy(), bhi[::-1].copy()
    if op == "sqrt":
        alo, ahi = np.sort(np.abs(np.stack([alo, ahi])), axis=0)
    s = {"scale": -2.7, "addc": 1.3, "rsubc": 1.0, "divc": -3.0, "pown": 5}.get(op, 0.0)
    return alo, ahi, blo, bhi, s


def tiny_case(lo, hi):
    return ((np.abs(lo) < TINY) & (lo != 0)) | ((np.abs(hi) < TINY) & (hi != 0))


@pytest.mark.parametrize("op", RATIONAL)
def test_exact_mode_bit_identical(gpu, oracle, op):
    rng = np.random.default_rng(OPC[op])
    alo, ahi, blo, bhi, s = inputs(op, rng)
    want = oracle.ia_eval(OPC[op], alo, ahi, blo, bhi, s)
    got = gpu.ia_eval(op, alo, ahi, blo, bhi, s, rounding="exact")
    same = (got[0] == want[0]) & (got[1] == want[1])
    # documented exception: one ulp wider where a product underflows
    exc = tiny_case(want[0], want[1]) | tiny_case(got[0], got[1])
    assert np.all(same | exc), np.flatnonzero(~(same | exc))[:10]
    assert np.all(got[0][exc] <= want[0][exc]) and np.all(got[1][exc] >= want[1][exc])
    assert same.mean() > 0.97


@pytest.mark.parametrize("op", RATIONAL)
def test_fast_mode_contains_and_one_ulp(gpu, oracle, op):
    rng = np.random.default_rng(100 + OPC[op])
    alo, ahi, blo, bhi, s = inputs(op, rng)
    want = oracle.ia_eval(OPC[op], alo, ahi, blo, bhi, s)
    got = gpu.ia_eval(op, alo, ahi, blo, bhi, s, rounding="fast")
    assert np.all(got[0] <= want[0]) and np.all(got[1] >= want[1])
    fin = np.isfinite(want[0]) & np.isfinite(want[1]) & np.isfinite(got[0]) & np.isfinite(got[1])
    fin &= ~tiny_case(want[0], want[1])
    per_op = 2 * 5 if op == "pown" else 2   # pown: one step per multiplication
    for i in np.flatnonzero(fin):
        assert ulps_between(got[0][i], want[0][i]) <= per_op
        assert ulps_between(want[1][i], got[1][i]) <= per_op


@pytest.mark.parametrize("fn", ["exp", "sin", "cos"])
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_transcendentals_contain_truth(gpu, oracle, fn, rounding):
    rng = np.random.default_rng(7)
    n = 4000
    c = rng.uniform(-700, 700, n) if fn == "exp" else rng.uniform(-300, 300, n)
    w = 10.0 ** rng.uniform(-15, 0.8, n)
    w[: n // 2] = 0.0                       # points: checks the nearest-value error bound
    lo_in, hi_in = c - w, c + w
    k = np.arange(-40, 40)
    if fn != "exp":
        lo_in[n // 2: n // 2 + 80] = k * math.pi / 2 - 1e-11
        hi_in[n // 2: n // 2 + 80] = k * math.pi / 2 + 1e-11
    got = gpu.ia_eval(fn, lo_in, hi_in, rounding=rounding)
    orc = oracle.ia_eval(OPC[fn], lo_in, hi_in)
    for i in range(n):
        if fn == "exp":
            tlo, thi = mp.exp(mp.mpf(lo_in[i])), mp.exp(mp.mpf(hi_in[i]))
        else:
            tlo, thi = trig_range(lo_in[i], hi_in[i], fn)
        assert mp.mpf(got[0][i]) <= tlo and mp.mpf(got[1][i]) >= thi, (i, lo_in[i], hi_in[i])
        # narrow sin/cos arguments (width <= 2^-40 max(1, |lo|)) may add their
        # width (one evaluation widened by it: |sin'|, |cos'| <= 1)
        w = hi_in[i] - lo_in[i]
        extra = 2 * w if (fn != "exp" and w <= 2.0 ** -40 * max(1.0, abs(lo_in[i]))) else 0.0
        # width: the a-priori bounds of the FP64 evaluation (sin: 1e-15 |r|,
        # cos: 2e-15, exp: 8e-15 before the 2^k scaling) plus outward steps
        slack = (TRANS_ABS[fn] * max(abs(float(tlo)), abs(float(thi)), 1e-300 if fn == "exp" else 1.0)
                 + extra)
        if math.isfinite(got[0][i]):
            assert float(tlo) - got[0][i] <= slack or ulps_between(got[0][i], float(tlo)) <= 4, (i, fn)
        if math.isfinite(got[1][i]):
            assert got[1][i] - float(thi) <= slack or ulps_between(float(thi), got[1][i]) <= 4, (i, fn)
        # the oracle's glibc-based enclosure also contains the range: the two intersect
        assert max(got[0][i], orc[0][i]) <= min(got[1][i], orc[1][i])


# absolute error bound of a transcendental endpoint relative to the result's
# magnitude (exp: relative), with margin for the final outward steps
TRANS_ABS = {"exp": 1.2e-14, "sin": 5e-15, "cos": 5e-15}


def mp_trig(fn, x):
    return mp.sin(mp.mpf(x)) if fn == "sin" else mp.cos(mp.mpf(x))


@pytest.mark.parametrize("rounding", ["exact", "fast"])
@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_trig_adversarial(gpu, fn, rounding):
    """Arguments next to k pi/2 (where the reduced argument cancels) for k up
    to 2^39, where k pi/2 is exact-reducible (|k| < 2^20) and beyond; tiny,
    subnormal and signed-zero arguments.  Point intervals and intervals of
    width 1e-12 around them: each enclosure contains the 300-bit range."""
    mp.mp.prec = 300
    ks = [1, 2, 3, 4, 7, 100, 1023, 1025, 2 ** 19 - 1, 2 ** 19 + 1, 2 ** 20 - 3, 2 ** 20 + 5, 2 ** 25 + 1,
          2 ** 30 - 7, 2 ** 35 + 3, 2 ** 39 - 1]
    xs = []
    for k in ks:
        c = float(mp.mpf(k) * mp.pi / 2)
        for s in (1, -1):
            x = s * c
            xs += [x, math.nextafter(x, math.inf), math.nextafter(x, -math.inf)]
    xs += [0.0, -0.0, 5e-324, -5e-324, 1e-310, 1e-300, -1e-200, 1e-8, 0.785398, -0.7854, 2.0 ** -27]
    xs = np.array(xs)
    for width in (0.0, 1e-12):
        lo, hi = xs - width, xs + width
        got = gpu.ia_eval(fn, lo, hi, rounding=rounding)
        for i in range(len(xs)):
            if width == 0.0:
                t = mp_trig(fn, lo[i])
                tlo, thi = t, t
            else:
                tlo, thi = trig_range(lo[i], hi[i], fn)
            assert mp.mpf(got[0][i]) <= tlo and thi <= mp.mpf(got[1][i]), (fn, i, lo[i], hi[i])
            if width == 0.0 and abs(lo[i]) < 2.0 ** 20:
                # exact reduction: the a-priori bound holds
                assert got[1][i] - got[0][i] <= 2 * TRANS_ABS[fn] * max(abs(float(t)), 1e-300) + 1e-300 + \
                    4 * abs(math.ulp(float(t))), (fn, i, lo[i], got[0][i], got[1][i])


@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_exp_adversarial(gpu, rounding):
    """exp near overflow (ln MAXV = 709.78...), near and below the subnormal
    range (-708 ... -745.2), at tiny and zero arguments, and at k ln 2 where
    the reduced argument cancels: enclosures contain the 300-bit value."""
    mp.mp.prec = 300
    xs = [709.78, 709.782712893383, 709.782712893384, 709.79, 709.8, 710.0, 800.0, -700.0, -708.0, -708.4,
          -708.5, -720.0, -740.0, -744.0, -745.0, -745.13, -745.1332191019411, -745.2, -746.0, -1000.0,
          0.0, -0.0, 5e-324, -5e-324, 1e-300, -1e-300, 1e-17, -1e-17, 0.3465, -0.3466]
    for k in (1, 2, 10, 100, 512, 1000, 1023, -1, -10, -1000, -1074):
        c = float(mp.mpf(k) * mp.log(2))
        xs += [c, math.nextafter(c, math.inf), math.nextafter(c, -math.inf)]
    xs = np.array(xs)
    got = gpu.ia_eval("exp", xs, xs, rounding=rounding)
    for i, x in enumerate(xs):
        t = mp.exp(mp.mpf(x))
        assert got[0][i] >= 0 and mp.mpf(got[0][i]) <= t, (i, x, got[0][i])
        assert t <= mp.mpf(got[1][i]), (i, x, got[1][i])
        if -700 < x < 709:
            assert got[1][i] - got[0][i] <= 2 * TRANS_ABS["exp"] * float(t) + 4 * math.ulp(float(t)), (i, x)


@pytest.mark.parametrize("op", ["add", "mul", "div", "sqr", "sqrt", "scale", "exp", "cos", "sin"])
def test_fp32_contains(gpu, oracle, op):
    rng = np.random.default_rng(9)
    alo, ahi, blo, bhi, s = inputs(op, rng)
    if op in ("exp", "cos", "sin"):
        c = rng.uniform(-20, 20, alo.size)
        alo, ahi = c - 0.01, c + 0.02
    ok = (np.abs(alo) < 1e30) & (np.abs(ahi) < 1e30) & (np.abs(alo) > 1e-30) & (np.abs(blo) > 1e-30) \
        & (np.abs(bhi) < 1e30) & (np.abs(bhi) > 1e-30)
    alo, ahi, blo, bhi = alo[ok], ahi[ok], blo[ok], bhi[ok]
    for rounding in ("exact", "fast"):
        got = gpu.ia_eval(op, alo, ahi, blo, bhi, s, prec="fp32", rounding=rounding)
        if op in ("exp", "cos", "sin"):
            for i in range(alo.size):
                tlo, thi = ((mp.exp(mp.mpf(alo[i])), mp.exp(mp.mpf(ahi[i]))) if op == "exp"
                            else trig_range(alo[i], ahi[i], op))
                assert mp.mpf(got[0][i]) <= tlo and mp.mpf(got[1][i]) >= thi
        else:
            want = oracle.ia_eval(OPC[op], alo, ahi, blo, bhi, s)
            fin = np.isfinite(want[0]) & np.isfinite(want[1])
            assert np.all(got[0][fin] <= want[0][fin]) and np.all(got[1][fin] >= want[1][fin])


@pytest.mark.parametrize("fn", ["sin", "cos", "exp"])
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_fp32_transcendentals_adversarial(gpu, fn, rounding):
    """FP32 intervals (their endpoints come from the short FP64 kernels,
    rounded outward): FP32 arguments next to k pi/2 for k up to 2^18 (the
    FP32 trig range is |x| <= 2^20), next to k ln 2, near the FP32 exp
    overflow (88.72) and underflow (-87.3 normal, -103.97 subnormal) edges,
    tiny and zero arguments; points and intervals a few ulps wide (the narrow
    path: one evaluation widened by the width) and 1e-3 wide.  Each
    enclosure contains the 300-bit range, and a point's is at most six FP32
    ulps of the value wide (plus 8 * 2^-126: FAST steps a zero bound by about the
    smallest normal): the rounding of the FP64 enclosure to FP32 and, in FAST
    rounding, two outward steps per end."""
    mp.mp.prec = 300
    f32 = lambda v: float(np.float32(v))
    xs = []
    if fn == "exp":
        for c in [88.72, 88.7228, 88.73, 89.0, -87.3, -87.34, -100.0, -103.9, -103.98, -104.0, -110.0, 0.0,
                  1e-30, -1e-30, 0.34657, -0.34657, 20.5, -33.25]:
            xs.append(f32(c))
        for k in (1, 3, 10, 50, 100, 127, -1, -10, -100, -126, -149):
            c = f32(mp.mpf(k) * mp.log(2))
            xs += [c, float(np.nextafter(np.float32(c), np.float32(np.inf))),
                   float(np.nextafter(np.float32(c), np.float32(-np.inf)))]
    else:
        for k in (1, 2, 3, 7, 100, 1023, 2 ** 12 + 1, 2 ** 15 - 1, 2 ** 18 + 3):
            for sgn in (1, -1):
                c = np.float32(sgn * float(mp.mpf(k) * mp.pi / 2))
                xs += [float(c), float(np.nextafter(c, np.float32(np.inf))), float(np.nextafter(c, np.float32(-np.inf)))]
        xs += [0.0, -0.0, 1e-40, -1e-38, 1e-8, f32(0.785398), f32(-0.7854), 1000.25, -65536.5]
    xs = np.array(xs)
    tf = {"sin": mp.sin, "cos": mp.cos, "exp": mp.exp}[fn]
    for width in (0.0, 4.0, 1e-3):
        if width == 4.0:   # a few FP32 ulps of the argument
            lo = np.array([float(np.nextafter(np.nextafter(np.float32(x), np.float32(-np.inf)), np.float32(-np.inf)))
                           for x in xs])
            hi = np.array([float(np.nextafter(np.nextafter(np.float32(x), np.float32(np.inf)), np.float32(np.inf)))
                           for x in xs])
        else:
            lo, hi = np.array([f32(x - width) for x in xs]), np.array([f32(x + width) for x in xs])
            lo, hi = np.minimum(lo, xs), np.maximum(hi, xs)
        got = gpu.ia_eval(fn, lo, hi, prec="fp32", rounding=rounding)
        for i in range(len(xs)):
            if fn == "exp":
                tlo, thi = mp.exp(mp.mpf(lo[i])), mp.exp(mp.mpf(hi[i]))
            elif lo[i] == hi[i]:
                tlo = thi = tf(mp.mpf(lo[i]))
            else:
                tlo, thi = trig_range(lo[i], hi[i], fn)
            assert mp.mpf(got[0][i]) <= tlo and thi <= mp.mpf(got[1][i]), (fn, width, i, lo[i], hi[i], got[0][i], got[1][i])
            if width == 0.0 and not (fn == "exp" and (lo[i] > 88.7 or lo[i] < -87.3)):
                t = float(tlo)
                ulp = float(np.spacing(np.float32(abs(t)))) if t != 0 else 0.0
                assert got[1][i] - got[0][i] <= 6 * ulp + 8 * 2.0 ** -126, (fn, i, lo[i], got[0][i], got[1][i], t)


def test_domain_errors(gpu):
    from gpugo_amd import GpugoError
    with pytest.raises(GpugoError, match="DOMAIN"):
        gpu.ia_eval("exp", [1.0], [0.0])
    with pytest.raises(GpugoError, match="DOMAIN"):
        gpu.ia_eval("divc", [1.0], [2.0], s=0.0)
    lo, hi = gpu.ia_eval("add", [], [], [], [])
    assert lo.size == 0


@pytest.mark.parametrize("op", RATIONAL)
def test_digits15_mode(gpu, oracle, op):
    """GPUGO_ROUND_DIGITS15 (the IA of the paper's results): each bound
    contains the true directed-rounding bound w (the oracle's EXACT result)
    and is the 15-significant-digit rounding of the nearest result, as
    outRound(lower1 + lower2, ...) rounds it (1-Ackley/IAAckley.py:106-141),
    kept no narrower than w.  The nearest result is w or the double one ulp
    inside it, so the checker's restatement of outRound (oracle.digits15) of
    either, held no narrower than w, is the expected bound: within one
    15-digit step everywhere, and equal for at least 90% of the bounds where
    the reference's power of ten 10^k (k = 14 - floor(log10 |w|)) is an exact
    integer (0 <= k <= 18; beyond, the reference's int64 power overflows and
    the device's and the host's pow may round differently)."""
    from oracle.oracle import digits15
    rng = np.random.default_rng(200 + OPC[op])
    alo, ahi, blo, bhi, s = inputs(op, rng)
    want = oracle.ia_eval(OPC[op], alo, ahi, blo, bhi, s)
    got = gpu.ia_eval(op, alo, ahi, blo, bhi, s, rounding="digits15")
    assert np.all(got[0] <= want[0]) and np.all(got[1] >= want[1])
    if op == "neg":
        return  # exact: nothing is rounded
    if op == "pown":
        # repeated multiplication: each step rounds to 15 digits
        fin = np.isfinite(want[0]) & np.isfinite(want[1])
        # (plus the documented widening of magnitudes below 1e-292)
        tol = 8 * 1e-14 * np.maximum(np.abs(want[0]), np.abs(want[1])) + 2e-292
        with np.errstate(invalid="ignore"):  # inf - inf outside `fin`
            dlo, dhi = want[0] - got[0], got[1] - want[1]
        assert np.all(dlo[fin] <= tol[fin]) and np.all(dhi[fin] <= tol[fin])
        return
    same = total = 0
    for i in range(len(alo)):
        for g, w, up in ((got[0][i], want[0][i], False), (got[1][i], want[1][i], True)):
            if not np.isfinite(w):
                continue
            w = float(w)
            inside = math.nextafter(w, -math.inf if up else math.inf)
            refs = [max(digits15(v, up), w) if up else min(digits15(v, up), w) for v in (w, inside)]
            assert min(abs(g - r) for r in refs) <= 1e-14 * abs(w) + 2 * math.ulp(w), (op, i, w, g, refs)
            if w != 0.0 and 0 <= 14 - math.floor(math.log10(abs(w))) <= 18:
                total += 1
                same += g in refs
    assert total > 0.2 * 2 * len(alo) and same >= 0.9 * total, (same, total)


@pytest.mark.parametrize("fn", ["exp", "sin", "cos"])
def test_digits15_transcendentals_contain(gpu, fn):
    rng = np.random.default_rng(300)
    lo = rng.uniform(-20, 20, 400)
    hi = lo + rng.uniform(0, 2, 400)
    g = gpu.ia_eval(fn, lo, hi, rounding="digits15")
    f = gpu.ia_eval(fn, lo, hi, rounding="exact")
    assert np.all(g[0] <= f[0]) and np.all(g[1] >= f[1])


def test_digits15_solve_certifies(gpu, oracle):
    """A solve in the paper's IA: certified enclosure of f* = 0 (sepcos d=2)
    that intersects the oracle's (true directed rounding)."""
    from gpugo_amd import problems as pb
    spec = pb.sepcos2()
    obj = pb.objective(spec.family, spec.d)
    m = gpu.get_module(obj, roundings=("digits15",))
    r = gpu.minimize(obj, spec.lo, spec.hi, tol=1e-6, grid=(2, 2), k_select=16, params=spec.params,
                     rounding="digits15", module=m, max_iters=10000)
    o, _ = oracle.solve(0, spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=2, cycle=2, k_select=16, max_iters=10000)
    assert r["status_name"] == "certified" and r["lb"] <= 0.0 <= r["ub"] and r["ub"] - r["lb"] <= 1e-6
    assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])
