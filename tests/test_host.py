"""Host-side product checks that need no GPU.

- the objective families (gpugo_amd.problems) are the functions their
  definitions say: a generic interpreter of the expression IR, run in
  256-bit mpmath, is compared with the definitions written independently in
  tests/test_oracle_objectives.py;
- the explicit gradient lists equal mpmath's derivative of the value IR;
- the emitter's symbolic differentiation equals mpmath's derivative;
- libgpugo.so loads and exports every function include/gpugo.h declares;
- argument and domain errors are reported before any GPU work.
"""
import ctypes
import math
import os
import re

import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import _lib, emit, expr as E, problems as pb
from tests.test_oracle_objectives import f_true

mp.mp.prec = 256
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ir_eval(e, x, P, memo=None, j=None):
    """Point value of an IR expression in mpmath (generic interpreter); the
    dimension of guarded sums and loop templates is len(x); j is the loop
    variable inside a template sum."""
    if memo is None:
        memo = {}
    k = id(e)
    if k in memo and j is None:
        return memo[k]
    op = e.op
    idx = (lambda v: v[1] * j + v[2]) if j is not None else None
    if op == "var":
        r = x[idx(e.val)] if isinstance(e.val, tuple) else x[e.val]
    elif op == "param":
        r = mp.mpf(float(P[idx(e.val) if isinstance(e.val, tuple) else e.val]))
    elif op == "tsum":
        r = mp.mpf(0)
        for jj in range(len(x) + e.val):
            r += ir_eval(e.args[0], x, P, {}, jj)
    elif op == "tprod":
        r = mp.mpf(1)
        for jj in range(len(x) + e.val):
            r *= ir_eval(e.args[0], x, P, {}, jj)
    elif op == "tprodx":
        r = mp.mpf(1)
        for jj in range(len(x) + e.val):
            if jj != j:
                r *= ir_eval(e.args[0], x, P, {}, jj)
    elif op == "jnum":
        r = mp.mpf(j + e.val)
    elif op == "twhen":
        ok = {"j+1<d": j + 1 < len(x), "j>0": j > 0, "j==0": j == 0, "j+1==d": j + 1 == len(x)}[e.val]
        r = ir_eval(e.args[0], x, P, {}, j) if ok else mp.mpf(0)
    elif op == "const":
        r = mp.mpf(e.val)
    elif op == "dim":
        r = mp.mpf(len(x))
    elif op == "iconst":
        lo, hi = e.val
        r = mp.pi if abs(lo - math.pi) < 1e-15 else mp.e
    elif op == "guard":
        r = ir_eval(e.args[0], x, P, memo, j) if e.val < len(x) else mp.mpf(0)
    elif op == "gsum":
        r = mp.mpf(0)
        for t in e.args:
            r += ir_eval(t, x, P, memo, j)
    else:
        a = [ir_eval(t, x, P, memo if j is None else {}, j) for t in e.args]
        r = {"add": lambda: a[0] + a[1], "sub": lambda: a[0] - a[1], "mul": lambda: a[0] * a[1],
             "div": lambda: a[0] / a[1], "neg": lambda: -a[0], "sqr": lambda: a[0] ** 2,
             "pown": lambda: a[0] ** e.val, "sqrt": lambda: mp.sqrt(a[0]), "exp": lambda: mp.exp(a[0]),
             "sin": lambda: mp.sin(a[0]), "cos": lambda: mp.cos(a[0])}[op]()
    if j is None:
        memo[k] = r
    return r


SPECS = [pb.sepcos2(), pb.chain10(), pb.ackley50(d=7), pb.rastrigin(3), pb.rosenbrock(4), pb.ackley(4)]


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: s.name)
def test_family_ir_matches_definition(spec):
    obj = pb.objective(spec.family, spec.d)
    obj.check()
    rng = np.random.default_rng(0)
    for _ in range(10):
        xf = rng.uniform(spec.lo, spec.hi)
        x = [mp.mpf(float(v)) for v in xf]
        v = ir_eval(obj.value, x, spec.params)
        t = f_true(spec.family, spec.params, xf)
        assert abs(v - t) <= mp.mpf("1e-60") * (1 + abs(t))


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: s.name)
def test_gradient_lists_match_derivative(spec):
    obj = pb.objective(spec.family, spec.d)
    rng = np.random.default_rng(1)
    xf = rng.uniform(spec.lo, spec.hi)
    x = [mp.mpf(float(v)) for v in xf]
    for j in range(spec.d):
        def fj(s, j=j):
            y = list(x)
            y[j] = s
            return ir_eval(obj.value, y, spec.params)
        want = mp.diff(fj, x[j])
        got = ir_eval(obj.grad[j], x, spec.params)
        assert abs(got - want) <= mp.mpf("1e-40") * (1 + abs(want)), j


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: s.name)
def test_loop_templates_match_unrolled(spec):
    """The term-parallel code lists' loop form is the same function."""
    obj = pb.objective(spec.family, spec.d)
    t = obj.templates
    rng = np.random.default_rng(5)
    for _ in range(4):
        x = [mp.mpf(float(v)) for v in rng.uniform(spec.lo, spec.hi)]
        v0 = ir_eval(obj.value, x, spec.params)
        v1 = ir_eval(t["value"], x, spec.params)
        assert abs(v0 - v1) <= mp.mpf("1e-60") * (1 + abs(v0))
        for jj in range(spec.d):
            g0 = ir_eval(obj.grad[jj], x, spec.params)
            g1 = ir_eval(t["grad"], x, spec.params, {}, jj)
            assert abs(g0 - g1) <= mp.mpf("1e-60") * (1 + abs(g0)), jj


def test_split_code_lists_emit():
    src = emit.module_source(pb.objective("sepcos_chain", 10), ("fast",), split=4)
    assert "#define GPUGO_SPLIT_G 4" in src and "obj_value_g" in src and "group_sum<R, G>" in src
    assert "for (int j = lg; j < d - 1; j += G)" in src


def test_symbolic_diff():
    x0, x1, x2 = E.X(0), E.X(1), E.X(2)
    f = E.exp(E.sin(x0 * 1.5) * x1) + E.sqrt(E.sqr(x1) + 2.0) / (x2 + 3.0) - E.pown(x0 - x2, 3) * E.P(0)
    obj = E.Objective(3, 1, f, None, name="t")
    grads = emit.gradient(obj)
    P = [0.7]
    x = [mp.mpf("0.3"), mp.mpf("-1.2"), mp.mpf("0.9")]
    for j in range(3):
        def fj(s, j=j):
            y = list(x)
            y[j] = s
            return ir_eval(f, y, P)
        assert abs(ir_eval(grads[j], x, P) - mp.diff(fj, x[j])) < mp.mpf("1e-50")


def test_emitter_lowering_rules():
    spec = pb.ackley50(d=3)
    src = emit.code_lists(pb.objective(spec.family, spec.d))
    # A + e enters as an interval constant plus a scalar; 1/d as a positive divisor
    assert "mk<double>(0x1.5bf0a8b145769p+1, 0x1.5bf0a8b14576ap+1)" in src
    # coefficients are declared >= 0: -A and -B scale one-sidedly
    assert "A::divc_pos(" in src and "A::scale_neg(t" in src and "A::scale_pos(" in src
    assert "#define GPUGO_PSIGN 1, 1, 1, 1, 1, 1" in src
    assert src.count("bool obj_kkt") == 1 and "return true;" in src
    full = emit.module_source(pb.objective("sepcos", 2), ("exact", "fast"))
    assert "#define GPUGO_WITH_EXACT 1" in full and "gpugo_bound_fast" in full


def header_functions():
    txt = open(os.path.join(ROOT, "include", "gpugo.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpugo_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(_lib.EXPORTS) <= set(names)


def test_errors_before_gpu():
    lib = _lib.load()
    d = lambda *v: np.array(v, dtype=np.float64)
    a, b, o1, o2 = d(1.0), d(2.0), d(0.0), d(0.0)
    P = _lib.ptr
    assert lib.gpugo_ia_eval(99, 0, 0, 1, P(a), P(b), None, None, 0.0, P(o1), P(o2)) == _lib.E_INVALID
    assert b"unknown op" in lib.gpugo_last_error()
    assert lib.gpugo_ia_eval(0, 0, 0, 1, P(a), P(b), None, None, 0.0, P(o1), P(o2)) == _lib.E_INVALID
    nan = d(float("nan"))
    assert lib.gpugo_ia_eval(12, 0, 0, 1, P(nan), P(b), None, None, 0.0, P(o1), P(o2)) == _lib.E_DOMAIN
    assert lib.gpugo_ia_eval(12, 0, 0, 1, P(b), P(a), None, None, 0.0, P(o1), P(o2)) == _lib.E_DOMAIN
    assert b"inverted" in lib.gpugo_last_error()
    assert lib.gpugo_ia_eval(8, 0, 0, 1, P(a), P(b), None, None, 0.0, P(o1), P(o2)) == _lib.E_DOMAIN
    if lib.gpugo_device_count() == 0:
        assert lib.gpugo_init(0) == _lib.E_NODEV


def test_rtc_compile_without_gpu(tmp_path, monkeypatch):
    """hipRTC builds a module for gfx950 with no GPU present."""
    import gpugo_amd.solver as S
    monkeypatch.setattr(S, "_CACHE_DIR", str(tmp_path))
    path = S.rtc_compile_to_cache(pb.objective("sepcos", 2), roundings=("fast",))
    data = open(path, "rb").read()
    assert len(data) > 1000 and data[:4] == b"\x7fELF"


def test_window_ranges_exact():
    """The window rule of the windowed (large-d) modules, run from the same
    source on the host: the ranges hold exactly the terms j in [0, nt) whose
    coordinates j + k (k in [klo, khi]) meet the window {(win + t) mod d},
    brute force over every window of small d (wrapping ones included)."""
    from gpugo_amd.solver import window_ranges
    rows, want = [], []
    for d in range(1, 13):
        for c in range(1, d + 1):
            for win in range(d):
                W = {(win + t) % d for t in range(c)}
                for klo, khi in ((0, 0), (0, 1), (-1, 1), (-1, 0), (-2, 3)):
                    for nt in {d, max(d - 1, 0), d + 1}:
                        rows.append((win, c, d, klo, khi, nt))
                        want.append({j for j in range(nt) if any(j + k in W for k in range(klo, khi + 1))})
    out = window_ranges(rows)
    for row, o, w in zip(rows, out, want):
        n = o[0]
        assert 0 <= n <= 2
        got = [set(range(o[1 + 2 * r], o[2 + 2 * r])) for r in range(n)]
        assert all(g for g in got), (row, o)
        if n == 2:
            assert not (got[0] & got[1]), (row, o)
        assert set().union(*got) == w, (row, o, sorted(w))


def test_window_terms_fit_table_rows():
    """Term tables keep (c + 2(w - 1)) rows per template sum of stencil
    width w (and per gradient template): the window's ranges never hold
    more terms than that, for every window of small d (wrapping ones
    included), stencils up to the widest a table covers."""
    from gpugo_amd.emit import TAB_MAX_W
    from gpugo_amd.solver import window_ranges
    rows = []
    for d in range(1, 15):
        for c in range(1, d + 1):
            for win in range(d):
                for w in range(1, TAB_MAX_W + 1):
                    for klo in range(-(w - 1), 1):
                        for nt in {d, max(d - 1, 0), d + 1}:
                            rows.append((win, c, d, klo, klo + w - 1, nt))
    out = window_ranges(rows)
    for row, o in zip(rows, out):
        c, w = row[1], row[4] - row[3] + 1
        nq = sum(o[2 + 2 * r] - o[1 + 2 * r] for r in range(o[0]))
        assert nq <= c + 2 * (w - 1), (row, o)


def test_window_ranges_rejects_bad_rows():
    from gpugo_amd.solver import window_ranges
    with pytest.raises(_lib.GpugoError):
        window_ranges([(5, 2, 5, 0, 0, 5)])  # win >= d
    with pytest.raises(_lib.GpugoError):
        window_ranges([(0, 6, 5, 0, 0, 5)])  # c > d


def test_window_module_emits():
    """Windowed module text: loop-form code lists only (no per-coordinate
    unrolling), rest sums per template sum, the window test on the gradient
    components the window changes (all of them for a global gradient)."""
    from gpugo_amd import families as F
    src = emit.window_module_source(F.objective_loops("levy", 1000), ("fast",))
    assert "#define GPUGO_WINDOWED 1" in src and "#define GPUGO_NREST 1" in src
    assert "x[999]" in src and "x[998]" not in src           # tail term only, no unrolled sum
    assert "win_ranges(win, c, d, -1, 1, d, glo, ghi)" in src  # local gradient: window components
    src = emit.window_module_source(F.objective_loops("ackley", 1000), ("fast",))
    assert "#define GPUGO_NREST 2" in src and "for (int j = 0; j < d; ++j) {" in src
    # product families: template products (rest product per parent, window
    # terms times it), the product-except-j gradient over the window only
    src = emit.window_module_source(F.objective_loops("griewank", 1000), ("fast",))
    assert "#define GPUGO_NREST 2" in src and "rest_is_prod(int s) { return (2u >> s) & 1u; }" in src
    assert "A::mul(t" in src and "if (i != j)" in src
    kkt = src[src.index("bool obj_kkt_w("):]
    kkt = kkt[:kkt.index("\n}\n")]
    assert "win_ranges(win, c, d, 0, 0, d, glo, ghi)" in kkt and "for (int j = 0; j < d; ++j) {" not in kkt
    src = emit.window_module_source(F.objective_loops("zabinsky", 100), ("fast",))
    assert "rest_is_prod(int s) { return (3u >> s) & 1u; }" in src
    # a product-except-j next to a global gradient term is refused
    from gpugo_amd.expr import J, Objective, TProd, TProdEx, TSum, X, cos, sqr
    bad = Objective(8, 0, TProd(cos(X(J))), None, name="bad",
                    templates={"value": TProd(cos(X(J))), "grad": TSum(sqr(X(J))) * TProdEx(cos(X(J)))})
    with pytest.raises(ValueError):
        emit.window_module_source(bad, ("fast",))
    # the 15-significant-digit IA: its kernels are instantiated on request only
    src = emit.window_module_source(F.objective_loops("levy", 50), ("digits15",))
    assert "#define GPUGO_WITH_DIGITS15 1" in src and "#define GPUGO_WITH_FAST 0" in src
    assert "gpugo_bound_digits15" in src and "gpugo_prep_digits15" in src
    with pytest.raises(ValueError):
        emit.window_module_source(F.objective_loops("levy", 50), ("nearest",))


def header_defines():
    txt = open(os.path.join(ROOT, "include", "gpugo.h")).read()
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(GPUGO_[A-Z0-9_]+)\s+(-?\d+)", txt)}


def test_flags_match_header():
    """The Python flag and status constants are the header's values."""
    h = header_defines()
    for name in ("FLAG_MEAN_VALUE", "FLAG_SPLIT", "FLAG_PERSIST", "FLAG_EAGER_COMPACT"):
        assert getattr(_lib, name) == h["GPUGO_" + name], name
    assert _lib.E_INVALID == h["GPUGO_E_INVALID"] and _lib.E_DOMAIN == h["GPUGO_E_DOMAIN"]
    flags = [h[k] for k in h if k.startswith("GPUGO_FLAG_")]
    assert len(set(flags)) == len(flags) and all(f & (f - 1) == 0 for f in flags)  # distinct single bits


def test_problem_create_rejects_null():
    """gpugo_problem_create checks its pointers before any device work."""
    lib = _lib.load()
    assert lib.gpugo_problem_create(None, None, None) == _lib.E_INVALID
