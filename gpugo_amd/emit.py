"""Expression -> HIP emitter for objective modules.

The emitted text is: ia.hpp + bnb_args.hpp + bnb_box.hpp + the objective's
code lists (obj_value, obj_grad, obj_kkt) + bnb_kernel.hpp.  It is compiled by hipRTC
(through libgpugo.so) into one code object per objective.  This replaces
Numba's JIT of each driver's code lists (GPUGO_1.1/1-Ackley/ackley50.py:
116-479, 602-766).

Lowering rules (the same op sequence a hand-written code list would use):
  scalar * I, I * scalar   -> scale      I + scalar -> addc     I - scalar -> addc(I, -s)
  scalar - I               -> rsubc      I / scalar -> divc     scalar / I -> div(point, I)
  I op I                   -> add/sub/mul/div
  sqr, pown, sqrt, exp, sin, cos -> the interval functions of ia.hpp
Repeated subexpressions are evaluated once per code list.
"""
from __future__ import annotations

import hashlib
import math
import os

import numpy as np
from typing import Dict, List, Optional, Tuple

from .expr import Expr, Objective, wrap, Const, has_loop_index, TEMPLATE_REDUCTIONS

_CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")


def _lit(v: float) -> str:
    """Exact C literal of a double."""
    if v != v:
        return "__builtin_nan(\"\")"
    if v == float("inf"):
        return "__builtin_huge_val()"
    if v == float("-inf"):
        return "(-__builtin_huge_val())"
    return float(v).hex()


def _f32_bounds(v: float) -> Tuple[float, float]:
    """Outward FP32 enclosure of a double."""
    f = np.float32(v)
    lo = f if float(f) <= v else np.nextafter(f, np.float32(-np.inf))
    hi = f if float(f) >= v else np.nextafter(f, np.float32(np.inf))
    return float(lo), float(hi)


def _lit32(v: float) -> str:
    """C literal of a double that is exactly a float."""
    if v == float("inf"):
        return "__builtin_huge_valf()"
    if v == float("-inf"):
        return "(-__builtin_huge_valf())"
    return float(v).hex() + "f"


def _pow2(e: Expr) -> bool:
    """A constant whose products are exact (+-2^k)."""
    if e.op != "const" or e.val == 0 or e.val != e.val or abs(e.val) == float("inf"):
        return False
    m, _ = math.frexp(abs(e.val))
    return m == 0.5


class _Func:
    """Straight-line code of one device function with CSE.  ty = "double":
    scalars (constants, coefficients, the dimension) are exact doubles;
    ty = "float": the FP32 pruning code list, where coefficients are FP32
    intervals (P[i] is an iv<float> rounded outward from the double) and
    constants not representable in FP32 become outward intervals."""

    def __init__(self, ty: str = "double"):
        self.lines: List[str] = []
        self.memo: Dict[tuple, Tuple[str, str]] = {}
        self.n = 0
        self.ty = ty

    def tmp(self, kind: str, code: str) -> str:
        name = f"t{self.n}"
        self.n += 1
        ty = f"iv<{self.ty}>" if kind == "i" else self.ty
        self.lines.append(f"  const {ty} {name} = {code};")
        return name

    def as_iv(self, kc: Tuple[str, str]) -> str:
        kind, code = kc
        return code if kind == "i" else f"point<{self.ty}>({code})"

    def lower(self, e: Expr) -> Tuple[str, str]:
        """Returns (kind, code): kind 's' = exact scalar C expression, 'i' = interval."""
        k = e.key()
        if k in self.memo:
            return self.memo[k]
        r = self._lower(e)
        self.memo[k] = r
        return r

    def _lower(self, e: Expr) -> Tuple[str, str]:
        op = e.op
        f32 = self.ty == "float"
        T = self.ty
        if op == "var":
            return ("i", f"x[{e.val}]")
        if op == "const":
            if not f32:
                return ("s", _lit(e.val))
            lo, hi = _f32_bounds(e.val)
            if lo == hi:
                return ("s", _lit32(lo))
            return ("i", self.tmp("i", f"mk<float>({_lit32(lo)}, {_lit32(hi)})"))
        if op == "param":
            return ("i" if f32 else "s", f"P[{e.val}]")
        if op == "dim":
            return ("s", f"(({T})d)")
        if op == "guard":
            raise ValueError("Guarded terms are only allowed inside gsum")
        if op == "gsum":
            # left fold with terms that exist only for j < d: a mutable
            # accumulator, each guarded term in its own scope
            acc = f"t{self.n}"
            self.n += 1
            terms = list(e.args)
            if terms[0].op == "guard":
                init = f"mk<{T}>(0, 0)"
            else:
                init = self.as_iv(self.lower(terms.pop(0)))
            self.lines.append(f"  iv<{T}> {acc} = {init};")
            for t in terms:
                if t.op == "guard":
                    self.lines.append(f"  if ({t.val} < d) {{")
                    saved = dict(self.memo)
                    code = self.as_iv(self.lower(t.args[0]))
                    self.lines.append(f"  {acc} = A::add({acc}, {code});")
                    self.memo = saved
                    self.lines.append("  }")
                else:
                    code = self.as_iv(self.lower(t))
                    self.lines.append(f"  {acc} = A::add({acc}, {code});")
            return ("i", acc)
        if op == "iconst":
            if f32:
                lo, hi = _f32_bounds(e.val[0])[0], _f32_bounds(e.val[1])[1]
                return ("i", self.tmp("i", f"mk<float>({_lit32(lo)}, {_lit32(hi)})"))
            return ("i", self.tmp("i", f"mk<double>({_lit(e.val[0])}, {_lit(e.val[1])})"))
        if e.is_scalar():
            if op == "neg":
                a = self.lower(e.args[0])
                if a[0] == "s":
                    return ("s", f"(-{a[1]})")
                return ("i", self.tmp("i", f"A::neg({a[1]})"))
            if op == "mul":
                (ka, a), (kb, b) = (self.lower(x) for x in e.args)
                # exact only when one factor is +-2^k; otherwise an interval
                if ka == "s" and kb == "s" and (_pow2(e.args[0]) or _pow2(e.args[1])):
                    return ("s", f"({a} * {b})")
                return ("i", self.tmp("i", f"A::mul({self.as_iv((ka, a))}, {self.as_iv((kb, b))})"))
        if op == "neg":
            return ("i", self.tmp("i", f"A::neg({self.as_iv(self.lower(e.args[0]))})"))
        if op in ("sqr", "sqrt", "exp", "sin", "cos"):
            a = self.as_iv(self.lower(e.args[0]))
            return ("i", self.tmp("i", f"A::{op}({a})"))
        if op == "pown":
            a = self.as_iv(self.lower(e.args[0]))
            return ("i", self.tmp("i", f"A::pown({a}, {int(e.val)})"))
        ea, eb = e.args
        a, b = self.lower(ea), self.lower(eb)
        if op == "mul":
            if a[0] == "s" and b[0] == "s":
                return ("i", self.tmp("i", f"A::scale(point<{T}>({a[1]}), {b[1]})"))
            if a[0] == "s":
                s_code, s_expr, iv_code = a[1], ea, b[1]
            elif b[0] == "s":
                s_code, s_expr, iv_code = b[1], eb, a[1]
            else:
                # an interval coefficient of known sign (FP32 code list)
                for se, sc, ic in ((ea, a[1], b[1]), (eb, b[1], a[1])):
                    if se.is_scalar() and se.scalar_sign() > 0:
                        return ("i", self.tmp("i", f"A::mulp({ic}, {sc})"))
                return ("i", self.tmp("i", f"A::mul({a[1]}, {b[1]})"))
            sign = s_expr.scalar_sign()
            fn = "scale_pos" if sign > 0 else ("scale_neg" if sign < 0 else "scale")
            return ("i", self.tmp("i", f"A::{fn}({iv_code}, {s_code})"))
        if op == "add":
            if a[0] == "s" and b[0] == "s":
                return ("i", self.tmp("i", f"A::addc(point<{T}>({a[1]}), {b[1]})"))
            if a[0] == "s":
                return ("i", self.tmp("i", f"A::addc({b[1]}, {a[1]})"))
            if b[0] == "s":
                return ("i", self.tmp("i", f"A::addc({a[1]}, {b[1]})"))
            return ("i", self.tmp("i", f"A::add({a[1]}, {b[1]})"))
        if op == "sub":
            if a[0] == "s" and b[0] == "s":
                return ("i", self.tmp("i", f"A::rsubc({a[1]}, point<{T}>({b[1]}))"))
            if a[0] == "s":
                return ("i", self.tmp("i", f"A::rsubc({a[1]}, {b[1]})"))
            if b[0] == "s":
                return ("i", self.tmp("i", f"A::addc({a[1]}, -{b[1]})"))
            return ("i", self.tmp("i", f"A::sub({a[1]}, {b[1]})"))
        if op == "div":
            if b[0] == "s":
                ia = self.as_iv(a)
                fn = "divc_pos" if eb.scalar_sign() > 0 else "divc"
                return ("i", self.tmp("i", f"A::{fn}({ia}, {b[1]})"))
            return ("i", self.tmp("i", f"A::div({self.as_iv(a)}, {b[1]})"))
        raise ValueError(f"cannot lower op {op}")


class _FuncG(_Func):
    """Term-parallel code lists (the loop templates of an objective): the
    G lanes of a group share one box; a TSum runs over j = lg, lg + G, ...
    and ends with a butterfly reduction over the group, so every lane of the
    group holds the sum; J-indexed leaves read index a*j + b.  Loop-invariant
    subexpressions are evaluated before the loop."""

    def __init__(self, ty: str = "double"):
        super().__init__(ty)
        self.in_loop = False
        self.jvar = "j"          # C name of the innermost template index
        self.loop_keys = set()   # memo keys of values that depend on it

    def lower(self, e: Expr) -> Tuple[str, str]:
        r = super().lower(e)
        if self.in_loop and has_loop_index(e):
            self.loop_keys.add(e.key())
        return r

    def _lower(self, e: Expr) -> Tuple[str, str]:
        op = e.op
        jv = self.jvar
        if op in ("var", "param") and isinstance(e.val, tuple):
            if not self.in_loop:
                raise ValueError("loop index used outside a template loop")
            _, a, b = e.val
            idx = f"({a} * {jv} + {b})" if b >= 0 else f"({a} * {jv} - {-b})"
            if op == "var":
                return ("i", f"x[{idx}]")
            # FP32 code lists take the coefficients as outward FP32 intervals
            return ("i" if self.ty == "float" else "s", f"P[{idx}]")
        if op == "jnum":
            if not self.in_loop:
                raise ValueError("loop index used outside a template loop")
            b = e.val
            return ("s", f"(({self.ty})({jv} + {b}))" if b >= 0 else f"(({self.ty})({jv} - {-b}))")
        if op in TEMPLATE_REDUCTIONS:
            return self.loop_sum(e)
        if op == "tprodx":
            if not self.in_loop or self.jvar != "j":
                raise ValueError("TProdEx belongs in a gradient template (one level deep)")
            return self.loop_prod_except(e)
        if op == "twhen":
            cond = {"j+1<d": f"{jv} + 1 < d", "j>0": f"{jv} > 0", "j==0": f"{jv} == 0",
                    "j+1==d": f"{jv} + 1 == d"}[e.val]
            name = f"t{self.n}"
            self.n += 1
            self.lines.append(f"  iv<{self.ty}> {name} = mk<{self.ty}>(0, 0);")
            self.lines.append(f"  if ({cond}) {{")
            saved = dict(self.memo)
            code = self.as_iv(self.lower(e.args[0]))
            self.lines.append(f"  {name} = {code};")
            self.memo = saved
            self.lines.append("  }")
            return ("i", name)
        return super()._lower(e)

    def hoist(self, body: Expr):
        """Evaluate the maximal loop-invariant subexpressions of `body`."""
        def walk(e):
            if not has_loop_index(e):
                if e.op not in ("var", "param", "const", "dim", "iconst") and not e.is_scalar():
                    self.lower(e)
                return
            for a in e.args:
                walk(a)
        walk(body)

    def loop(self, body: Expr, lim: str, emit_fn):
        """for (j = lg; j < lim; j += G) { emit_fn(code of body) }"""
        if self.in_loop:
            raise ValueError("nested template loops are not supported")
        self.hoist(body)
        self.lines.append(f"  for (int j = lg; j < {lim}; j += G) {{")
        saved = dict(self.memo)
        self.in_loop = True
        code = self.as_iv(self.lower(body))
        self.lines += emit_fn(code)
        self.in_loop = False
        self.memo = saved
        self.lines.append("  }")

    def loop_sum(self, e: Expr) -> Tuple[str, str]:
        acc = f"t{self.n}"
        self.n += 1
        lim = _limit(e.val)
        prod = e.op == "tprod"
        self.lines.append(f"  iv<{self.ty}> {acc} = mk<{self.ty}>({int(prod)}, {int(prod)});")
        fn = "mul" if prod else "add"
        self.loop(e.args[0], lim, lambda code: [f"    {acc} = A::{fn}({acc}, {code});"])
        self.lines.append(f"  {acc} = group_{'prod' if prod else 'sum'}<R, G>({acc});")
        return ("i", acc)

    def loop_prod_except(self, e: Expr) -> Tuple[str, str]:
        """Product over i != j of the body at J = i: a sequential loop inside
        the component loop (O(d) per component; the windowed lists use the
        parent's rest product instead)."""
        acc = f"t{self.n}"
        self.n += 1
        self.lines.append(f"  iv<{self.ty}> {acc} = mk<{self.ty}>(1, 1);")
        saved = dict(self.memo)
        self.memo = {k: v for k, v in saved.items() if k not in self.loop_keys}
        self.jvar = "i"
        self.lines.append(f"  for (int i = 0; i < {_limit(e.val)}; ++i) {{")
        self.lines.append("    if (i == j) continue;")
        code = self.as_iv(self.lower(e.args[0]))
        self.lines.append(f"    {acc} = A::mul({acc}, {code});")
        self.lines.append("  }")
        self.jvar = "j"
        self.memo = saved
        return ("i", acc)


def _limit(off: int) -> str:
    """C expression of a template's trip count d + off."""
    return "d" if off == 0 else (f"d + {off}" if off > 0 else f"d - {-off}")


# Product over a group of G lanes (template products), the butterfly of
# group_sum (bnb_box.hpp) with A::mul; emitted by the code lists that need it.
_GROUP_PROD = """#ifndef GPUGO_GROUP_PROD
#define GPUGO_GROUP_PROD 1
namespace gpugo {
template <int R, int G>
__device__ __forceinline__ iv<double> group_prod(iv<double> v) {
  typedef IA<double, R> A;
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const double lo = __shfl_xor(v.lo, o, 64), hi = __shfl_xor(v.hi, o, 64);
    v = A::mul(v, mk<double>(lo, hi));
  }
  return v;
}
}  // namespace gpugo
#endif
"""


def split_lists(obj: Objective) -> str:
    """obj_value_g / obj_grad_g / obj_kkt_g from the objective's loop
    templates: the term-parallel code lists (G lanes per box)."""
    t = obj.templates
    if not t or "value" not in t or "grad" not in t:
        raise ValueError(f"objective {obj.name} has no loop templates")
    has_prod = any(e.op == "tprod" for e in _tsums(t["value"], t["grad"]))
    out = ([_GROUP_PROD] if has_prod else []) + ["namespace gpugo {"]
    f = _FuncG()
    v = f.lower(wrap(t["value"]))
    out.append("template <int R, int G, class XS>\n__device__ __forceinline__ iv<double> obj_value_g("
               "const XS &x, const double *__restrict__ P, int d, int lg) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")
    grad = wrap(t["grad"])
    # gradient components j = lg, lg + G, ... into a sink (one lane each)
    f = _FuncG()
    f.loop(grad, "d", lambda code: [f"    g[j] = {code};"])
    out.append("template <int R, int G, class XS, class GS>\n__device__ __forceinline__ void obj_grad_g("
               "const XS &x, const double *__restrict__ P, GS &&g, int d, int lg) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("}")
    # monotonicity test: each lane tests its components, the group ORs
    f = _FuncG()
    f.lines.append("  bool prune = false;")
    f.loop(grad, "d", lambda code: [
        f"    {{ const iv<double> gj = {code};",
        "      prune |= (gj.lo > 0.0 && x[j].lo != dlo[j]) || (gj.hi < 0.0 && x[j].hi != dhi[j]); }"])
    out.append("template <int R, int G, class XS>\n__device__ __forceinline__ bool obj_kkt_g("
               "const XS &x, const double *__restrict__ P, const double *__restrict__ dlo, "
               "const double *__restrict__ dhi, int d, int lg) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("  return group_any<G>(prune);")
    out.append("}")
    out.append("}  // namespace gpugo")
    return "\n".join(out) + "\n"


# ---------------------------------------------------------------------------
# windowed code lists (large d): a child differs from its parent record only
# in the c coordinates of the parent's split window, so each template sum is
# (rest over the parent) + (terms that touch the window, over the child)
# ---------------------------------------------------------------------------
def _walk(e: Expr, fn, seen=None):
    if seen is None:
        seen = set()
    if id(e) in seen:
        return
    seen.add(id(e))
    fn(e)
    for a in e.args:
        _walk(a, fn, seen)


def _reduction_of(e: Expr) -> Expr:
    """The template reduction a node stands for: itself for a sum or a
    product, the full product for a product-except-j."""
    return Expr("tprod", e.args, e.val) if e.op == "tprodx" else e


def _tsums(*exprs) -> List[Expr]:
    """Distinct template reductions (sums and products; a product-except-j
    counts as its full product) of the expressions, in order of appearance."""
    out, keys = [], set()

    def fn(e):
        if e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx":
            r = _reduction_of(e)
            if r.key() not in keys:
                keys.add(r.key())
                out.append(r)
    for e in exprs:
        _walk(wrap(e), fn)
    return out


def _global_reductions(e: Expr) -> List[Expr]:
    """Template sums and products that are not under a product-except-j
    (a gradient that uses one needs every component re-tested)."""
    out = []

    def walk(x):
        if x.op == "tprodx":
            return
        if x.op in TEMPLATE_REDUCTIONS:
            out.append(x)
            return
        for a in x.args:
            walk(a)
    walk(wrap(e))
    return out


def _x_offsets(body: Expr, what: str) -> Tuple[int, int]:
    """[min, max] of b over the coordinates X(J + b) of a loop body; the
    windowed form needs unit stride.  No coordinate: (0, 0)."""
    offs = []

    def fn(e):
        if e.op == "tprodx":
            return
        if e.op == "var":
            if not isinstance(e.val, tuple):
                raise ValueError(f"{what}: fixed coordinate X[{e.val}] inside a loop body is not windowable")
            if e.val[1] != 1:
                raise ValueError(f"{what}: coordinate index {e.val[1]}*j{e.val[2]:+d} is not windowable")
            offs.append(e.val[2])
        for a in e.args:
            fn(a)
    fn(wrap(body))
    return (min(offs), max(offs)) if offs else (0, 0)


def _has_op(e: Expr, op: str) -> bool:
    found = []
    _walk(wrap(e), lambda x: found.append(1) if x.op == op else None)
    return bool(found)


def _check_local_body(body: Expr, what: str):
    """A loop body may not use coordinate-dependent values computed outside
    its own term (they would differ between a parent and its children)."""
    def fn(e):
        if e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx":
            raise ValueError(f"{what}: nested template reductions are not windowable")
        if not has_loop_index(e) and e.depends_on_x():
            raise ValueError(f"{what}: loop-invariant coordinate term {e!r} is not windowable")
    _walk(body, fn)


class _FuncW(_FuncG):
    """Windowed lowering: a TSum becomes rest[s] (its terms that do not
    touch the window, enclosed over the parent record once per parent) plus
    the terms j in the window's ranges, evaluated over the child."""

    def __init__(self, slots: Dict[tuple, int], ty: str = "double"):
        super().__init__(ty)
        self.slots = slots

    def loop_sum(self, e: Expr) -> Tuple[str, str]:
        s = self.slots[e.key()]
        acc = f"t{self.n}"
        self.n += 1
        if self.ty == "float":
            # the FP64 rest sum rounded outward to an FP32 interval
            self.lines.append(f"  iv<float> {acc} = mk<float>(to_dn<float>(rest[{s}].lo), to_up<float>(rest[{s}].hi));")
            term = f"wterm32_{s}"
        else:
            self.lines.append(f"  iv<double> {acc} = rest[{s}];")
            term = f"wterm_{s}"
        fn = "mul" if e.op == "tprod" else "add"
        self.lines.append(f"  {{ int rlo[2], rhi[2]; const int nr = wranges_{s}(win, c, d, rlo, rhi);")
        self.lines.append("    for (int r = 0; r < nr; ++r)")
        self.lines.append(f"      for (int j = rlo[r]; j < rhi[r]; ++j) {acc} = A::{fn}({acc}, {term}<R>(x, P, d, j)); }}")
        return ("i", acc)

    def loop_prod_except(self, e: Expr) -> Tuple[str, str]:
        """Inside the component loop j (j in the window): the parent's rest
        product (which has no window term) times the window's terms i != j
        over the child."""
        s = self.slots[_reduction_of(e).key()]
        acc = f"t{self.n}"
        self.n += 1
        self.lines.append(f"  iv<double> {acc} = rest[{s}];")
        self.lines.append(f"  {{ int rlo[2], rhi[2]; const int nr = wranges_{s}(win, c, d, rlo, rhi);")
        self.lines.append("    for (int r = 0; r < nr; ++r)")
        self.lines.append("      for (int i = rlo[r]; i < rhi[r]; ++i)")
        self.lines.append(f"        if (i != j) {acc} = A::mul({acc}, wterm_{s}<R>(x, P, d, i)); }}")
        return ("i", acc)


class _FuncWP(_FuncW):
    """Cooperative windowed lowering: the G lanes of a group share one child;
    a template reduction splits its window terms over the lanes (term q of
    the window ranges on lane q mod G), folds them with a butterfly and
    then with the parent's rest reduction."""

    def loop_sum(self, e: Expr) -> Tuple[str, str]:
        s = self.slots[e.key()]
        acc = f"t{self.n}"
        self.n += 1
        prod = e.op == "tprod"
        fn = "mul" if prod else "add"
        self.lines.append(f"  iv<double> {acc} = mk<double>({int(prod)}, {int(prod)});")
        self.lines.append(f"  {{ int rlo[2], rhi[2]; const int nr = wranges_{s}(win, c, d, rlo, rhi);")
        self.lines.append("    const int n0 = nr > 0 ? rhi[0] - rlo[0] : 0, nq = n0 + (nr > 1 ? rhi[1] - rlo[1] : 0);")
        self.lines.append("    for (int q = lg; q < nq; q += G) {")
        self.lines.append("      const int j = q < n0 ? rlo[0] + q : rlo[1] + (q - n0);")
        self.lines.append(f"      {acc} = A::{fn}({acc}, wterm_{s}<R>(x, P, d, j)); }} }}")
        self.lines.append(f"  {acc} = A::{fn}(rest[{s}], group_{'prod' if prod else 'sum'}<R, G>({acc}));")
        return ("i", acc)


class _FuncWT(_FuncW):
    """Windowed lowering over term tables: term j of template sum s over a
    child depends only on the digits of the window coordinates in its
    stencil [klo, khi], so the step's prep kernel evaluates it once per
    (term, digit combination) and a child adds table entries instead of
    evaluating terms.  Same operations on the same intervals, in the same
    order, as _FuncW: the result is bit-identical."""

    def __init__(self, slots: Dict[tuple, int], stencils: List[Tuple[int, int]], ty: str = "double"):
        super().__init__(slots, ty)
        self.stencils = stencils

    def _lower(self, e: Expr) -> Tuple[str, str]:
        if e.op == "tpart":
            # a gradient part read from the parts table (obj_kkt_wl)
            return ("i", f"gp{e.val}")
        return super()._lower(e)

    def loop_sum(self, e: Expr) -> Tuple[str, str]:
        s = self.slots[e.key()]
        klo, khi = self.stencils[s]
        acc = f"t{self.n}"
        self.n += 1
        ty = self.ty
        if ty == "float":
            self.lines.append(f"  iv<float> {acc} = mk<float>(to_dn<float>(rest[{s}].lo), to_up<float>(rest[{s}].hi));")
        else:
            self.lines.append(f"  iv<double> {acc} = rest[{s}];")
        fn = "mul" if e.op == "tprod" else "add"
        w = khi - klo + 1
        if w == 1 and klo == 0 and e.val == 0:
            # one coordinate per term and a term for every coordinate: the
            # window's q-th term is window coordinate q (the ranges start at
            # win and wrap to 0), so its entry is the q-th packed digit; same
            # entries in the same order as the general loop below
            self.lines.append(f"  {{ const iv<{ty}> *__restrict__ Ts = T + L.off[{s}]; const unsigned Es = L.E[{s}];")
            self.lines.append("    unsigned long long pk = xd.packed;")
            self.lines.append(f"#pragma unroll {TAB_LOADS}")
            self.lines.append("    for (int q = 0; q < c; ++q) {")
            self.lines.append(f"      {acc} = A::{fn}({acc}, Ts[q * Es + (unsigned)(pk & xd.mask)]);")
            self.lines.append("      pk >>= xd.bpd;")
            self.lines.append("    } }")
            return ("i", acc)
        # the window's terms as one run q = 0 .. nq-1 (term j of range 0,
        # then of range 1); table entries are loaded TAB_LOADS at a time and
        # folded in q order (the same order as term by term)
        self.lines.append(f"  {{ int rlo[2], rhi[2]; const int nr = wranges_{s}(win, c, d, rlo, rhi);")
        self.lines.append(f"    const iv<{ty}> *__restrict__ Ts = T + L.off[{s}]; const unsigned Es = L.E[{s}];")
        self.lines.append("    const int n0 = nr > 0 ? rhi[0] - rlo[0] : 0, nq = n0 + (nr > 1 ? rhi[1] - rlo[1] : 0);")
        self.lines.append("    int q = 0;")
        self.lines.append(f"    for (; q + {TAB_LOADS} <= nq; q += {TAB_LOADS}) {{")
        self.lines.append(f"      iv<{ty}> te[{TAB_LOADS}];")
        self.lines.append("#pragma unroll")
        self.lines.append(f"      for (int u = 0; u < {TAB_LOADS}; ++u) {{")
        self.lines.append("        const int qq = q + u, j = qq < n0 ? rlo[0] + qq : rlo[1] + (qq - n0);")
        self.lines.append(f"        te[u] = Ts[qq * Es + tab_idx<{w}>(xd, j + ({klo}))];")
        self.lines.append("      }")
        self.lines.append("#pragma unroll")
        self.lines.append(f"      for (int u = 0; u < {TAB_LOADS}; ++u) {acc} = A::{fn}({acc}, te[u]);")
        self.lines.append("    }")
        self.lines.append("    for (; q < nq; ++q) {")
        self.lines.append("      const int j = q < n0 ? rlo[0] + q : rlo[1] + (q - n0);")
        self.lines.append(f"      {acc} = A::{fn}({acc}, Ts[q * Es + tab_idx<{w}>(xd, j + ({klo}))]);")
        self.lines.append("    } }")
        return ("i", acc)

    def loop_prod_except(self, e: Expr) -> Tuple[str, str]:
        """Inside the component loop j: the parent's rest product times the
        window's factors i != j, read from the table of the full product."""
        s = self.slots[_reduction_of(e).key()]
        klo, khi = self.stencils[s]
        acc = f"t{self.n}"
        self.n += 1
        self.lines.append(f"  iv<double> {acc} = rest[{s}];")
        self.lines.append(f"  {{ int rlo[2], rhi[2]; const int nr = wranges_{s}(win, c, d, rlo, rhi);")
        self.lines.append(f"    const iv<double> *__restrict__ Ts = T + L.off[{s}]; const unsigned Es = L.E[{s}];")
        self.lines.append("    unsigned q = 0;")
        self.lines.append("    for (int r = 0; r < nr; ++r)")
        self.lines.append("      for (int i = rlo[r]; i < rhi[r]; ++i, ++q)")
        self.lines.append(f"        if (i != j) {acc} = A::mul({acc}, Ts[q * Es + tab_idx<{khi - klo + 1}>(xd, i + ({klo}))]); }}")
        return ("i", acc)


# widest stencil (coordinates per term) a term table covers: wid^w entries per term
TAB_MAX_W = 3
# table entries a child loads before folding them (loads in flight together)
TAB_LOADS = 4
# windowed modules keep exp/sin/cos out of line (inlined, measured with term
# tables: the d = 50 headline unchanged, its ten-sample solve 10% faster, Salomon,
# Belegundu, Griewank and Zabinsky at d = 1000 5-9% slower, compile time doubled)
WINDOW_TRANS_INLINE = False


def _grad_parts(grad: Expr) -> List[Expr]:
    """Maximal subexpressions of a gradient template that read the
    coordinates only through X(J + k): no template reduction and no
    loop-invariant coordinate value (those are the gradient's hoisted
    globals).  Over a child, part p of component j depends only on the digits
    of the window coordinates in its stencil, like a template term."""
    parts, keys = [], {}

    def local(e):
        if e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx":
            return False
        if not has_loop_index(e) and e.depends_on_x():
            return False
        return all(local(a) for a in e.args)

    def walk(e):
        if e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx":
            return  # a reduction binds its own J: hoisted, not a part
        if e.depends_on_x() and has_loop_index(e) and local(e):
            if e.key() not in keys:
                keys[e.key()] = len(parts)
                parts.append(e)
            return
        for a in e.args:
            walk(a)
    walk(wrap(grad))
    return parts


def _replace(e: Expr, mapping: Dict[tuple, Expr], memo=None) -> Expr:
    """e with the subexpressions whose keys are in `mapping` replaced
    (not inside template reductions, which bind their own J)."""
    if memo is None:
        memo = {}
    if e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx":
        return e
    k = e.key()
    if k in mapping:
        return mapping[k]
    if k in memo:
        return memo[k]
    r = Expr(e.op, tuple(_replace(a, mapping, memo) for a in e.args), e.val, e.val2) if e.args else e
    memo[k] = r
    return r


def _parts_ok(grad: Expr, global_grad: bool) -> bool:
    """A gradient that uses template reductions (so every component is
    tested) and whose coordinate-dependent local work can be tabled per
    component: the parts table of obj_kkt_wl."""
    if not global_grad or _has_op(grad, "tprodx"):
        return False
    parts = _grad_parts(grad)
    if not parts or len(parts) > 8:
        return False
    # a table read only pays for a part that does interval work beyond a
    # scalar operation on a coordinate (measured: Salomon, whose only part is
    # X(J), was ~10% slower with the table)
    if max(_count([prt]) for prt in parts) < _FLOP["mul"]:
        return False
    klo, khi = _x_offsets(grad, "gradient template")
    return khi - klo + 1 <= TAB_MAX_W


def _window_tables(value: Expr, grad: Expr, sums: List[Expr], slots: Dict[tuple, int], global_grad: bool,
                   grad_tab: bool, parts_tab: bool) -> List[str]:
    """Term tables of the windowed lists (FP64 part):
      TabLayout, tab_layout(c, wid, S)   entries per parent record: (1 + S)
          points (the box, then its S diagonal sample points) of NT entries
          (slot s: (c + 2(w_s - 1)) terms x wid^w_s digit combinations at
          off[s]), then the gradient table (box only) when the gradient
          template is local
      wterm_any<R>(s, x, P, d, j)         term j of slot s
      wgrad_t<R>(x, P, d, j)              gradient component j (local templates)
      obj_value_wt<R>(x, xd, T, L, P, d, rest, win, c)   value at a point of
          a child (x: the child or one of its diagonal points, for the
          coordinates read outside the template sums; xd: the child, for the
          digits) from that point's table
      obj_kkt_wt<R>(x, TG, L, P, dlo, dhi, d, win, c)  monotonicity test from
          the gradient table
      obj_kkt_wl<R>(x, T, L, P, dlo, dhi, d, rest, win, c)  monotonicity test
          of a gradient with template reductions (every component tested):
          the reductions from the point-0 tables, the components' local parts
          (_grad_parts) from the parts table: per part, d entries over the
          record's box, then (c + 2(w - 1)) x wid^w entries of the components
          whose stencil meets the window"""
    stencils = [_x_offsets(e.args[0], f"template sum {s}") for s, e in enumerate(sums)]
    widths = [khi - klo + 1 for klo, khi in stencils]
    if grad_tab or parts_tab:
        gklo, gkhi = _x_offsets(grad, "gradient template")
        gw = gkhi - gklo + 1
    else:
        gklo, gkhi, gw = 0, 0, 1
    parts = _grad_parts(grad) if parts_tab else []
    out = ["// term tables (see _window_tables in gpugo_amd/emit.py for the layout)",
           f"#define GPUGO_TAB_GRAD {int(grad_tab)}",
           f"#define GPUGO_TAB_PARTS {len(parts)}",
           f"#define GPUGO_TAB_GKLO {gklo}",
           f"#define GPUGO_TAB_GW {gw}",
           "struct TabLayout {",
           f"  int off[{len(sums) + 1}];  // per point: slot s at off[s]; off[NREST] = NT entries",
           f"  unsigned E[{len(sums)}];    // wid^w_s digit combinations per term",
           "  int goff;                   // gradient table: (1 + S) * NT",
           "  unsigned gE;                // wid^w_g (gradient and parts tables)",
           "  int poff, pstride;          // parts table: part k at poff + k * pstride (d + (c + 2(w_g - 1)) * gE)",
           "  int total;                  // entries per parent record",
           "};",
           "__host__ __device__ __forceinline__ TabLayout tab_layout(int c, unsigned wid, int S, int d) {",
           "  TabLayout L;",
           "  int o = 0;"]
    for s, w in enumerate(widths):
        out.append(f"  L.off[{s}] = o; L.E[{s}] = " + " * ".join(["wid"] * w) + f"; o += (c + {2 * (w - 1)}) * (int)L.E[{s}];")
    out += [f"  L.off[{len(sums)}] = o;",
            "  L.goff = (1 + S) * o;",
            "  L.gE = " + " * ".join(["wid"] * gw) + ";",
            f"  L.poff = L.goff + (GPUGO_TAB_GRAD ? (c + {2 * (gw - 1)}) * (int)L.gE : 0);",
            f"  L.pstride = d + (c + {2 * (gw - 1)}) * (int)L.gE;",
            "  L.total = L.poff + GPUGO_TAB_PARTS * L.pstride;",
            "  return L;",
            "}"]
    # term of any slot (the prep kernel's table build)
    out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> wterm_any("
               "int s, const XS &x, const double *__restrict__ P, int d, int j) {")
    out.append("  switch (s) {")
    for s in range(len(sums)):
        out.append(f"  case {s}: return wterm_{s}<R>(x, P, d, j);")
    out.append("  default: return mk<double>(0.0, 0.0);")
    out.append("  }")
    out.append("}")
    out.append("__device__ __forceinline__ int wranges_any(int s, int win, int c, int d, int *lo, int *hi) {")
    out.append("  switch (s) {")
    for s in range(len(sums)):
        out.append(f"  case {s}: return wranges_{s}(win, c, d, lo, hi);")
    out.append("  default: return 0;")
    out.append("  }")
    out.append("}")
    out.append("__device__ __forceinline__ int tab_klo(int s) {")
    out.append("  switch (s) {")
    for s, (klo, _) in enumerate(stencils):
        out.append(f"  case {s}: return {klo};")
    out.append("  default: return 0;")
    out.append("  }")
    out.append("}")
    out.append("__device__ __forceinline__ int tab_w(int s) {")
    out.append("  switch (s) {")
    for s, w in enumerate(widths):
        out.append(f"  case {s}: return {w};")
    out.append("  default: return 1;")
    out.append("  }")
    out.append("}")
    # value over a child from a point's table
    f = _FuncWT(slots, stencils)
    v = f.lower(value)
    out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> obj_value_wt("
               "const XS &x, const ChildBox &xd, const iv<double> *__restrict__ T, const TabLayout &L, "
               "const double *__restrict__ P, int d, const iv<double> *__restrict__ rest, int win, int c) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")
    if parts:
        # local parts of component j (the prep kernel's parts table)
        out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> wgpart("
                   "int k, const XS &x, const double *__restrict__ P, int d, int j) {")
        out.append("  typedef IA<double, R> A;")
        out.append("  switch (k) {")
        for k, prt in enumerate(parts):
            f = _FuncG()
            f.in_loop = True
            code = f.as_iv(f.lower(prt))
            out.append(f"  case {k}: {{")
            out += ["  " + ln for ln in f.lines]
            out.append(f"    return {code};")
            out.append("  }")
        out.append("  default: return mk<double>(0.0, 0.0);")
        out.append("  }")
        out.append("}")
        skel = _replace(wrap(grad), {prt.key(): Expr("tpart", (), k) for k, prt in enumerate(parts)})
        f = _FuncWT(slots, stencils)
        f.hoist(skel)
        f.lines.append(f"  int glo[2], ghi[2]; const int ng = win_ranges(win, c, d, {gklo}, {gkhi}, d, glo, ghi);")
        f.lines.append("  const int g0 = ng > 0 ? ghi[0] - glo[0] : 0;")
        f.lines.append("  const iv<double> *__restrict__ TP = T + L.poff;")
        # the window's components first (see _kkt_lines), then the others
        f.lines.append("  for (int r = 0; r <= ng; ++r)")
        f.lines.append("  for (int j = r < ng ? glo[r] : 0, je = r < ng ? ghi[r] : d; j < je; ++j) {")
        f.lines.append("    if (r == ng && in_ranges(j, ng, glo, ghi)) continue;")
        f.lines.append("    const int q = r == ng ? -1 : (r == 0 ? j - glo[0] : g0 + (j - glo[1]));")
        f.lines.append(f"    const iv<double> *__restrict__ tp = q < 0 ? TP + j : TP + d + q * (int)L.gE + "
                       f"(int)tab_idx<{gw}>(xd, j + ({gklo}));")
        for k in range(len(parts)):
            f.lines.append(f"    const iv<double> gp{k} = tp[{k} * L.pstride];")
        saved = dict(f.memo)
        f.in_loop = True
        code = f.as_iv(f.lower(skel))
        f.in_loop = False
        f.memo = saved
        f.lines += [f"    const iv<double> gj = {code};",
                    "    if (gj.lo > 0.0 && x[j].lo != dlo[j]) return true;",
                    "    if (gj.hi < 0.0 && x[j].hi != dhi[j]) return true;",
                    "  }"]
        out.append("template <int R>\n__device__ __forceinline__ bool obj_kkt_wl("
                   "const ChildBox &x, const iv<double> *__restrict__ T, const TabLayout &L, "
                   "const double *__restrict__ P, const double *__restrict__ dlo, const double *__restrict__ dhi, "
                   "int d, const iv<double> *__restrict__ rest, int win, int c) {")
        out.append("  typedef IA<double, R> A;")
        out.append("  const ChildBox &xd = x;")
        out += f.lines
        out.append("  return false;")
        out.append("}")
    if not grad_tab:
        # monotonicity test with the gradient's reductions from the tables
        # and each component evaluated over the child
        f = _FuncWT(slots, stencils)
        _kkt_lines(f, grad, global_grad)
        out.append("template <int R>\n__device__ __forceinline__ bool obj_kkt_wr("
                   "const ChildBox &x, const iv<double> *__restrict__ T, const TabLayout &L, "
                   "const double *__restrict__ P, const double *__restrict__ dlo, const double *__restrict__ dhi, "
                   "int d, const iv<double> *__restrict__ rest, int win, int c) {")
        out.append("  typedef IA<double, R> A;")
        out.append("  const ChildBox &xd = x;")
        out += f.lines
        out.append("  return false;")
        out.append("}")
    if grad_tab:
        f = _FuncG()
        f.in_loop = True
        code = f.as_iv(f.lower(grad))
        out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> wgrad_t("
                   "const XS &x, const double *__restrict__ P, int d, int j) {")
        out.append("  typedef IA<double, R> A;")
        out += f.lines
        out.append(f"  return {code};")
        out.append("}")
        out.append("template <int R>\n__device__ __forceinline__ bool obj_kkt_wt("
                   "const ChildBox &x, const iv<double> *__restrict__ TG, const TabLayout &L, "
                   "const double *__restrict__ dlo, const double *__restrict__ dhi, int d, int win, int c) {")
        out.append(f"  int glo[2], ghi[2]; const int ng = win_ranges(win, c, d, {gklo}, {gkhi}, d, glo, ghi);")
        out.append("  unsigned q = 0;")
        out.append("  for (int r = 0; r < ng; ++r)")
        out.append("  for (int j = glo[r]; j < ghi[r]; ++j, ++q) {")
        out.append(f"    const iv<double> gj = TG[q * L.gE + tab_idx<{gw}>(x, j + ({gklo}))];")
        out.append("    if (gj.lo > 0.0 && x[j].lo != dlo[j]) return true;")
        out.append("    if (gj.hi < 0.0 && x[j].hi != dhi[j]) return true;")
        out.append("  }")
        out.append("  return false;")
        out.append("}")
    return out


def _tab_ok(sums: List[Expr]) -> bool:
    """Every template reduction has a stencil narrow enough for a table
    (and the module info has room for the slots' widths)."""
    if len(sums) > 8:
        return False
    for s, e in enumerate(sums):
        klo, khi = _x_offsets(e.args[0], f"template sum {s}")
        if khi - klo + 1 > TAB_MAX_W:
            return False
    return True


def _grad_tab_ok(grad: Expr, global_grad: bool) -> bool:
    """The gradient template is local (no reduction, no product-except-j)
    with a narrow stencil: one table entry per (component, digits)."""
    if global_grad or _has_op(grad, "tprodx") or _tsums(grad):
        return False
    klo, khi = _x_offsets(grad, "gradient template")
    return khi - klo + 1 <= TAB_MAX_W


def window_lists_f32(obj: Objective) -> str:
    """FP32 windowed value code list of the FP32-prune form, after
    window_lists (it uses its wranges_s):
      wterm32_s<R>(x, P, d, j)   term j of template sum s in FP32 intervals
      obj_value_w32<R>(x, P, d, rest, win, c)  value over a child: each FP64
          rest sum rounded outward to FP32 plus the window's terms in FP32
    x is an FP32 accessor (coordinates rounded outward) and P the
    coefficients as outward FP32 intervals."""
    t = obj.templates
    value, grad = wrap(t["value"]), wrap(t["grad"])
    sums = _tsums(value, grad)
    slots = {e.key(): i for i, e in enumerate(sums)}
    out = ["namespace gpugo {"]
    for s, e in enumerate(sums):
        f = _FuncG("float")
        f.in_loop = True
        code = f.as_iv(f.lower(e.args[0]))
        out.append(f"template <int R, class XS>\n__device__ __forceinline__ iv<float> wterm32_{s}("
                   "const XS &x, const iv<float> *__restrict__ P, int d, int j) {")
        out.append("  typedef IA<float, R> A;")
        out += f.lines
        out.append(f"  return {code};")
        out.append("}")
    f = _FuncW(slots, "float")
    v = f.lower(value)
    out.append("template <int R, class XS>\n__device__ __forceinline__ iv<float> obj_value_w32("
               "const XS &x, const iv<float> *__restrict__ P, int d, const iv<double> *__restrict__ rest, "
               "int win, int c) {")
    out.append("  typedef IA<float, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")
    if _tab_ok(sums):
        # FP32 term tables (same layout as the FP64 ones, FP32 entries)
        stencils = [_x_offsets(e.args[0], f"template sum {s}") for s, e in enumerate(sums)]
        out.append("template <int R, class XS>\n__device__ __forceinline__ iv<float> wterm32_any("
                   "int s, const XS &x, const iv<float> *__restrict__ P, int d, int j) {")
        out.append("  switch (s) {")
        for s in range(len(sums)):
            out.append(f"  case {s}: return wterm32_{s}<R>(x, P, d, j);")
        out.append("  default: return mk<float>(0.0f, 0.0f);")
        out.append("  }")
        out.append("}")
        f = _FuncWT(slots, stencils, "float")
        v = f.lower(value)
        out.append("template <int R, class XS>\n__device__ __forceinline__ iv<float> obj_value_wt32("
                   "const XS &x, const ChildBox &xd, const iv<float> *__restrict__ T, const TabLayout &L, "
                   "const iv<float> *__restrict__ P, int d, const iv<double> *__restrict__ rest, int win, int c) {")
        out.append("  typedef IA<float, R> A;")
        out += f.lines
        out.append(f"  return {f.as_iv(v)};")
        out.append("}")
    out.append("}  // namespace gpugo")
    return "\n".join(out) + "\n"


def window_lists(obj: Objective) -> str:
    """Windowed code lists from the objective's loop templates:
      wterm_s<R>(x, P, d, j)   term j of template sum s
      wranges_s(win, c, d, lo, hi)  its terms that touch the window
      obj_rest<R>(s, x, P, d, win, c, t, nt, jend)  thread t of nt: partial
          sum of sum s's other terms j < jend over x (the parent box or one of
          its diagonal sample points), j = t, t + nt, ...
      obj_value_w<R>(x, P, d, rest, win, c)   value over a child
      obj_kkt_w<R>(x, P, dlo, dhi, d, rest, win, c)  monotonicity test
    The monotonicity test covers the gradient components the window can
    change: all of them if the gradient template uses a template sum,
    else those whose coordinates meet the window (the others equal the
    parent's, which did not prune it)."""
    t = obj.templates
    if not t or "value" not in t or "grad" not in t:
        raise ValueError(f"objective {obj.name} has no loop templates")
    value, grad = wrap(t["value"]), wrap(t["grad"])
    sums = _tsums(value, grad)
    slots = {e.key(): i for i, e in enumerate(sums)}
    S = len(sums)
    signs = param_signs(obj)
    out = [_GROUP_PROD, f"#define GPUGO_D {obj.d}", f"#define GPUGO_NPARAMS {obj.n_params}",
           "#define GPUGO_PSIGN " + (", ".join(str(v) for v in signs) if signs else "0"),
           f"#define GPUGO_NREST {max(S, 1)}", "namespace gpugo {"]
    prod_mask = sum(1 << s for s, e in enumerate(sums) if e.op == "tprod")
    out.append("// slot s of the rest reductions is a product (identity 1, combined\n"
               "// with A::mul) when bit s is set, else a sum (identity 0, A::add)")
    out.append(f"__device__ __forceinline__ bool rest_is_prod(int s) {{ return ({prod_mask}u >> s) & 1u; }}")
    for s, e in enumerate(sums):
        body = e.args[0]
        _check_local_body(body, f"template sum {s}")
        klo, khi = _x_offsets(body, f"template sum {s}")
        f = _FuncG()
        f.in_loop = True
        code = f.as_iv(f.lower(body))
        out.append(f"template <int R, class XS>\n__device__ __forceinline__ iv<double> wterm_{s}("
                   "const XS &x, const double *__restrict__ P, int d, int j) {")
        out.append("  typedef IA<double, R> A;")
        out += f.lines
        out.append(f"  return {code};")
        out.append("}")
        nt = _limit(e.val)
        out.append(f"__device__ __forceinline__ int wranges_{s}(int win, int c, int d, int *lo, int *hi) {{\n"
                   f"  return win_ranges(win, c, d, {klo}, {khi}, {nt}, lo, hi);\n}}")
    # rest sums of a parent record: slot s over an accessor (the box, or
    # one of its diagonal sample points)
    out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> obj_rest("
               "int s, const XS &x, const double *__restrict__ P, int d, int win, int c, int t, int nt, int jend) {")
    out.append("  typedef IA<double, R> A;")
    out.append("  iv<double> acc = mk<double>(0.0, 0.0);")
    out.append("  int rlo[2], rhi[2];")
    out.append("  switch (s) {")
    for s_, e in enumerate(sums):
        nt = _limit(e.val)
        prod = e.op == "tprod"
        out.append(f"  case {s_}: {{ const int nr = wranges_{s_}(win, c, d, rlo, rhi);")
        if prod:
            out.append("    acc = mk<double>(1.0, 1.0);")
        out.append(f"    for (int j = t, je = min({nt}, jend); j < je; j += nt)")
        out.append(f"      if (!in_ranges(j, nr, rlo, rhi)) acc = A::{'mul' if prod else 'add'}(acc, wterm_{s_}<R>(x, P, d, j));")
        out.append("    break; }")
    out.append("  default: break;")
    out.append("  }")
    out.append("  return acc;")
    out.append("}")
    # value over a child
    f = _FuncW(slots)
    v = f.lower(value)
    out.append("template <int R, class XS>\n__device__ __forceinline__ iv<double> obj_value_w("
               "const XS &x, const double *__restrict__ P, int d, const iv<double> *__restrict__ rest, "
               "int win, int c) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")
    # monotonicity test
    f = _FuncW(slots)
    global_grad = bool(_global_reductions(grad))
    if _has_op(grad, "tprodx"):
        # the rest product leaves out exactly the window's terms, so the
        # components tested must be the window's coordinates
        if global_grad or _x_offsets(grad, "gradient template") != (0, 0):
            raise ValueError("a product-except-j gradient is windowable only with X(J) alone outside it")
        for e in _tsums(grad):
            if e.op == "tprod" and _x_offsets(e.args[0], "template product") != (0, 0):
                raise ValueError("a windowed template product may read X(J) only")
    _kkt_lines(f, grad, global_grad)
    out.append("template <int R, class XS>\n__device__ __forceinline__ bool obj_kkt_w("
               "const XS &x, const double *__restrict__ P, const double *__restrict__ dlo, "
               "const double *__restrict__ dhi, int d, const iv<double> *__restrict__ rest, int win, int c) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("  return false;")
    out.append("}")
    out += _window_lists_coop(value, grad, slots, global_grad)
    tab = _tab_ok(sums)
    out.append(f"#define GPUGO_TAB {int(tab)}")
    if tab:
        out += _window_tables(value, grad, sums, slots, global_grad, _grad_tab_ok(grad, global_grad),
                              _parts_ok(grad, global_grad))
    out.append("}  // namespace gpugo")
    return "\n".join(out) + "\n"


def _kkt_lines(f: "_FuncW", grad: Expr, global_grad: bool):
    """The windowed monotonicity test into f's lines: the gradient's
    reductions (hoisted), then each component the window can change; true
    at the first component that rules the child out."""
    f.hoist(grad)
    if global_grad:
        # every component is tested, the window's first: they changed from
        # the parent's, which did not rule the parent out, so they are the
        # likeliest to rule the child out (the test is an OR: same result)
        gklo, gkhi = _x_offsets(grad, "gradient template")
        f.lines.append(f"  int glo[2], ghi[2]; const int ng = win_ranges(win, c, d, {gklo}, {gkhi}, d, glo, ghi);")
        f.lines.append("  for (int r = 0; r <= ng; ++r)")
        f.lines.append("  for (int j = r < ng ? glo[r] : 0, je = r < ng ? ghi[r] : d; j < je; ++j) {")
        f.lines.append("    if (r == ng && in_ranges(j, ng, glo, ghi)) continue;")
    else:
        gklo, gkhi = _x_offsets(grad, "gradient template")
        f.lines.append(f"  int glo[2], ghi[2]; const int ng = win_ranges(win, c, d, {gklo}, {gkhi}, d, glo, ghi);")
        f.lines.append("  for (int r = 0; r < ng; ++r)")
        f.lines.append("  for (int j = glo[r]; j < ghi[r]; ++j) {")
    saved = dict(f.memo)
    f.in_loop = True
    code = f.as_iv(f.lower(grad))
    f.in_loop = False
    f.memo = saved
    f.lines += [f"    const iv<double> gj = {code};",
                "    if (gj.lo > 0.0 && x[j].lo != dlo[j]) return true;",
                "    if (gj.hi < 0.0 && x[j].hi != dhi[j]) return true;",
                "  }"]


def _window_lists_coop(value: Expr, grad: Expr, slots: Dict[tuple, int], global_grad: bool) -> List[str]:
    """Cooperative value and monotonicity lists: the G lanes of a group
    (lane lg of G) evaluate one child together; every lane returns the same
    value / flag.  Used for the few children of a wavefront that reach the
    FP64 tests, whose one-lane evaluation would be the step's critical path.
      obj_value_wp<R, G>(x, P, d, rest, win, c, lg)
      obj_kkt_wp<R, G>(x, P, dlo, dhi, d, rest, win, c, lg)"""
    out = []
    f = _FuncWP(slots)
    v = f.lower(value)
    out.append("template <int R, int G, class XS>\n__device__ __forceinline__ iv<double> obj_value_wp("
               "const XS &x, const double *__restrict__ P, int d, const iv<double> *__restrict__ rest, "
               "int win, int c, int lg) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")
    f = _FuncWP(slots)
    f.hoist(grad)
    f.lines.append("  bool prune = false;")
    if global_grad:
        f.lines.append("  for (int j = lg; j < d; j += G) {")
    else:
        gklo, gkhi = _x_offsets(grad, "gradient template")
        f.lines.append(f"  int glo[2], ghi[2]; const int ng = win_ranges(win, c, d, {gklo}, {gkhi}, d, glo, ghi);")
        f.lines.append("  const int g0 = ng > 0 ? ghi[0] - glo[0] : 0, gq = g0 + (ng > 1 ? ghi[1] - glo[1] : 0);")
        f.lines.append("  for (int q = lg; q < gq; q += G) {")
        f.lines.append("    const int j = q < g0 ? glo[0] + q : glo[1] + (q - g0);")
    saved = dict(f.memo)
    f.in_loop = True
    code = f.as_iv(f.lower(grad))
    f.in_loop = False
    f.memo = saved
    f.lines += [f"    const iv<double> gj = {code};",
                "    prune |= (gj.lo > 0.0 && x[j].lo != dlo[j]) || (gj.hi < 0.0 && x[j].hi != dhi[j]);",
                "  }"]
    out.append("template <int R, int G, class XS>\n__device__ __forceinline__ bool obj_kkt_wp("
               "const XS &x, const double *__restrict__ P, const double *__restrict__ dlo, "
               "const double *__restrict__ dhi, int d, const iv<double> *__restrict__ rest, int win, int c, int lg) {")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("  return group_any<G>(prune);")
    out.append("}")
    return out


def window_module_source(obj: Objective, roundings=("fast",), f32: bool = False) -> str:
    """hipRTC source of a windowed (large-d) module: loop-form code lists
    only (no per-coordinate unrolling, so the code does not grow with d),
    the windowed split+bound kernel and the per-parent rest-sum kernel.
    `f32` adds the FP32 windowed value list of the FP32-prune form (FAST
    and EXACT roundings only)."""
    bad = set(roundings) - {"exact", "fast", "digits15"}
    if bad or not roundings:
        raise ValueError(f"roundings must be a non-empty subset of ('exact', 'fast', 'digits15'), got {roundings}")
    if f32 and "digits15" in roundings:
        raise ValueError("the 15-digit IA has no FP32 form")
    parts = [
        "// generated by gpugo_amd.emit (windowed form) for objective " + obj.name,
        f"#define GPUGO_WITH_EXACT {int('exact' in roundings)}",
        f"#define GPUGO_WITH_FAST {int('fast' in roundings)}",
        f"#define GPUGO_WITH_DIGITS15 {int('digits15' in roundings)}",
        f"#define GPUGO_WITH_F32 {int(bool(f32))}",
        "#define GPUGO_WINDOWED 1",
        # a child's coordinates are read only inside its window (c of them):
        # let the compiler keep the parent loads instead of redoing them
        "#define GPUGO_ACCESSOR_D 0",
        f"#define GPUGO_TRANS_INLINE {int(WINDOW_TRANS_INLINE)}",
        _read("ia.hpp"), _read("bnb_args.hpp"), _read("bnb_box.hpp"),
        window_lists(obj),
    ]
    if f32:
        parts.append(window_lists_f32(obj))
    parts += [split_lists(obj), "#define GPUGO_MODULE_STEP 1", _read("bnb_step.hpp"), _read("bnb_window.hpp")]
    return "\n".join(parts)


# ---------------------------------------------------------------------------
# symbolic differentiation (used when the objective has no gradient lists)
# ---------------------------------------------------------------------------
def diff(e: Expr, j: int, memo: Optional[dict] = None) -> Optional[Expr]:
    """d e / d X[j]; None means identically zero."""
    if memo is None:
        memo = {}
    k = e.key()
    if k in memo:
        return memo[k]
    r = _diff(e, j, memo)
    memo[k] = r
    return r


def _mul(a: Optional[Expr], b: Optional[Expr]) -> Optional[Expr]:
    if a is None or b is None:
        return None
    if a.op == "const" and a.val == 1.0:
        return b
    if b.op == "const" and b.val == 1.0:
        return a
    return a * b


def _add(a: Optional[Expr], b: Optional[Expr]) -> Optional[Expr]:
    if a is None:
        return b
    if b is None:
        return a
    return a + b


def _neg(a: Optional[Expr]) -> Optional[Expr]:
    return None if a is None else -a


def _diff(e: Expr, j: int, memo) -> Optional[Expr]:
    op = e.op
    if op == "var":
        return Const(1.0) if e.val == j else None
    if op in ("const", "param", "iconst"):
        return None
    if not e.depends_on_x():
        return None
    if op == "neg":
        return _neg(diff(e.args[0], j, memo))
    if op == "add":
        return _add(diff(e.args[0], j, memo), diff(e.args[1], j, memo))
    if op == "sub":
        return _add(diff(e.args[0], j, memo), _neg(diff(e.args[1], j, memo)))
    if op == "mul":
        a, b = e.args
        return _add(_mul(diff(a, j, memo), b), _mul(a, diff(b, j, memo)))
    if op == "div":
        a, b = e.args
        da, db = diff(a, j, memo), diff(b, j, memo)
        t1 = None if da is None else da / b
        t2 = None if db is None else (a * db) / (b * b)
        return _add(t1, _neg(t2))
    a = e.args[0]
    da = diff(a, j, memo)
    if da is None:
        return None
    if op == "sqr":
        return _mul(a * 2.0, da)
    if op == "pown":
        n = int(e.val)
        if n == 0:
            return None
        inner = Const(float(n)) if n == 1 else (Expr("pown", (a,), n - 1) * float(n) if n > 2 else a * 2.0)
        return _mul(inner, da)
    if op == "sqrt":
        return _mul(da, 0.5 / e)
    if op == "exp":
        return _mul(e, da)
    if op == "sin":
        return _mul(Expr("cos", (a,)), da)
    if op == "cos":
        return _neg(_mul(Expr("sin", (a,)), da))
    raise ValueError(f"cannot differentiate op {op}")


def gradient(obj: Objective) -> List[Expr]:
    if obj.grad is not None:
        return [wrap(g) for g in obj.grad]
    out = []
    for j in range(obj.d):
        g = diff(obj.value, j)
        out.append(g if g is not None else Const(0.0))
    return out


# ---------------------------------------------------------------------------
# module text
# ---------------------------------------------------------------------------
def param_signs(obj: Objective) -> List[int]:
    """Declared sign of each coefficient (0 = any), collected from the IR."""
    signs = [0] * obj.n_params
    seen = set()

    def walk(e):
        k = id(e)
        if k in seen:
            return
        seen.add(k)
        if e.op == "param" and e.val2:
            # P(a*J + b) inside a loop template: the indices of j in [0, d)
            idxs = ([e.val[1] * j + e.val[2] for j in range(obj.d)] if isinstance(e.val, tuple) else [e.val])
            for i in idxs:
                if not 0 <= i < obj.n_params:
                    continue
                if signs[i] not in (0, e.val2):
                    raise ValueError(f"P[{i}] declared with both signs")
                signs[i] = e.val2
        for a in e.args:
            walk(a)

    walk(obj.value)
    for g in obj.grad or ():
        walk(wrap(g))
    return signs


def code_lists(obj: Objective, suffix: str = "", dynamic: bool = False, header: bool = True) -> str:
    """obj_value / obj_grad / obj_kkt for one objective.  With dynamic=True
    the functions take the runtime dimension d <= obj.d (coordinates j >= d
    are absent) -- the variable-dimension form used by batches."""
    obj.check()
    grads = gradient(obj)
    d = obj.d
    dp = ", int d"
    out = []
    if header:
        signs = param_signs(obj)
        out += [f"#define GPUGO_D {d}", f"#define GPUGO_NPARAMS {obj.n_params}",
                "#define GPUGO_PSIGN " + (", ".join(str(v) for v in signs) if signs else "0")]
    out.append("namespace gpugo {")

    # value
    f = _Func()
    v = f.lower(obj.value)
    # x is an array of intervals or an accessor whose operator[] decodes
    # coordinate j on use (fewer live registers at large d)
    out.append(f"template <int R, class XS>\n__device__ __forceinline__ iv<double> obj_value{suffix}("
               f"const XS &x, const double *__restrict__ P{dp}) {{")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append(f"  return {f.as_iv(v)};")
    out.append("}")

    def component(f, j, g, body_fn, absent=None):
        """Lines for gradient component j; absent coordinates skip it (or
        run the `absent` lines instead)."""
        guarded = dynamic and j > 0
        if guarded:
            f.lines.append(f"  if ({j} < d) {{")
            saved = dict(f.memo)
        code = f.as_iv(f.lower(g))
        f.lines += body_fn(code)
        if guarded:
            f.memo = saved
            f.lines.append("  }" + (f" else {{ {absent} }}" if absent else ""))

    # gradient (parity hook, mean-value bound); absent components are zero
    # and are only written when absent, so a sink sees each present
    # component exactly once
    f = _Func()
    for j, g in enumerate(grads):
        component(f, j, g, lambda code, j=j: [f"  g[{j}] = {code};"],
                  absent=f"g[{j}] = mk<double>(0.0, 0.0);")
    # g is an array, or a sink whose g[j] = G_j folds each component into a
    # mean-value sum as it is produced
    out.append(f"template <int R, class XS, class GS>\n__device__ __forceinline__ void obj_grad{suffix}("
               f"const XS &x, const double *__restrict__ P, "
               f"GS &&g{dp}) {{")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("}")

    # monotonicity test with early exit per component
    # (GPUGO_1.1/1-Ackley/ackley50.py:458-477)
    f = _Func()
    for j, g in enumerate(grads):
        component(f, j, g, lambda code, j=j: [
            f"  {{ const iv<double> gj = {code};",
            f"    if (gj.lo > 0.0 && x[{j}].lo != dlo[{j}]) return true;",
            f"    if (gj.hi < 0.0 && x[{j}].hi != dhi[{j}]) return true; }}"])
    out.append(f"template <int R, class XS>\n__device__ __forceinline__ bool obj_kkt{suffix}("
               f"const XS &x, const double *__restrict__ P, "
               f"const double *__restrict__ dlo, const double *__restrict__ dhi{dp}) {{")
    out.append("  typedef IA<double, R> A;")
    out += f.lines
    out.append("  return false;")
    out.append("}")
    out.append("}  // namespace gpugo")
    return "\n".join(out) + "\n"


def value_list_f32(obj: Objective, suffix: str = "") -> str:
    """obj_value32: the value code list in FP32 intervals (outward rounded,
    one ulp per operation in FAST mode) for the pruning pass.  P is the
    coefficient array rounded outward to FP32 intervals."""
    f = _Func("float")
    v = f.lower(obj.value)
    out = ["namespace gpugo {",
           f"template <int R, class XS>\n__device__ __forceinline__ iv<float> obj_value32{suffix}("
           f"const XS &x, const iv<float> *__restrict__ P, int d) {{",
           "  typedef IA<float, R> A;"]
    out += f.lines
    out += [f"  return {f.as_iv(v)};", "}", "}  // namespace gpugo"]
    return "\n".join(out) + "\n"


# Algorithmic work per interval operation, in FLOP (the accounting used for
# the roofline): +,-,scalar ops 2; sqr 2; pown(k) 2*ceil(log2 k); interval
# product 8; interval quotient 10; sqrt 2; exp/sin/cos 2 endpoints x 20.
_FLOP = {"add": 2, "sub": 2, "neg": 0, "scale": 2, "addc": 2, "rsubc": 2, "divc": 2, "sqr": 2,
         "mul": 8, "div": 10, "sqrt": 2, "exp": 40, "sin": 40, "cos": 40}


def _count(exprs, tsum_leaf: bool = False, only=None) -> int:
    """FLOP of the expressions (each distinct subexpression once); with
    tsum_leaf, template sums count as leaves (their terms are counted by the
    caller); `only` (a predicate) counts only the nodes it accepts."""
    seen = set()
    total = 0

    def walk(e):
        nonlocal total
        k = e.key()
        if k in seen:
            return
        seen.add(k)
        if tsum_leaf and (e.op in TEMPLATE_REDUCTIONS or e.op == "tprodx"):
            return
        for a in e.args:
            walk(a)
        if only is not None and not only(e):
            return
        op = e.op
        if op in ("var", "const", "param", "iconst", "guard", "twhen", "dim", "jnum", "tpart") or e.is_scalar():
            return
        if op == "gsum":
            total += _FLOP["add"] * (len(e.args) - 1)
            return
        if op == "pown":
            n = max(int(e.val), 2)
            total += 2 * (n - 1).bit_length()
            return
        if op in ("add", "sub", "mul", "div"):
            sa, sb = e.args[0].is_scalar(), e.args[1].is_scalar()
            if op == "mul" and (sa or sb):
                total += _FLOP["scale"]
                return
            if op == "div" and sb:
                total += _FLOP["divc"]
                return
        total += _FLOP[op]

    for e in exprs:
        walk(wrap(e))
    return total


def flop_counts(obj: Objective) -> dict:
    """Algorithmic FLOP of the code lists of one child box: value enclosure,
    full gradient enclosure (the monotonicity test may stop earlier) and one
    sample evaluation (the value code list on a point)."""
    v = _count([obj.value])
    g = _count(gradient(obj))
    return {"value": v, "grad": g, "sample": v}


def window_flop_counts(obj: Objective, c: int, tables: bool = False, wid: int = 3, n_samples: int = 1) -> dict:
    """Algorithmic FLOP per child of the windowed code lists at a nominal
    window (c coordinates away from the ends): value = the non-sum part plus,
    per template sum, its terms that touch the window (c + stencil span) and
    their additions; grad = the monotonicity test over the components the
    window changes (c + span, or all d for a gradient that uses a template
    sum); sample = the value code list at one point.  The per-parent rest
    sums (O(d) once per parent record) are not included.
    tables: the step reads window terms from its term tables, so a child
    pays only the folds of the table entries (and, for a local gradient
    template, no arithmetic at all in the monotonicity test); `table` is the
    FLOP of filling one record's tables (wid per split coordinate,
    n_samples diagonal points), paid once per split record by the prep
    kernel."""
    t = obj.templates
    value, grad = wrap(t["value"]), wrap(t["grad"])
    sums = _tsums(value, grad)
    tab = tables and _tab_ok(sums)
    gtab = tab and _grad_tab_ok(grad, bool(_global_reductions(grad)))

    def sums_cost(e):
        total = 0
        for ts in _tsums(e):
            klo, khi = _x_offsets(ts.args[0], "template sum")
            fold = _FLOP["mul" if ts.op == "tprod" else "add"]
            total += min(c + khi - klo, obj.d + ts.val) * ((0 if tab else _count([ts.args[0]])) + fold)
        return total
    v = _count([value], tsum_leaf=True) + sums_cost(value)
    if _global_reductions(grad):
        ncomp, gsum = obj.d, sums_cost(grad)
    else:
        gklo, gkhi = _x_offsets(grad, "gradient template")
        ncomp, gsum = min(c + gkhi - gklo, obj.d), 0
    # a product-except-j: the window's other terms, c - 1 products each
    px = sum(min(c, obj.d) * ((0 if tab else _count([ts.args[0]])) + _FLOP["mul"])
             for ts in _tsums(grad) if _has_op(grad, "tprodx") and ts.op == "tprod")
    ptab = tab and not gtab and _parts_ok(grad, bool(_global_reductions(grad)))
    parts = _grad_parts(grad) if ptab else []
    # the parts come from the parts table: a component costs its skeleton
    skel = _replace(grad, {prt.key(): Expr("tpart", (), k) for k, prt in enumerate(parts)}) if ptab else grad
    # loop-invariant subexpressions (the gradient's globals) are evaluated
    # once per child, the rest once per component
    comp = _count([skel], tsum_leaf=True, only=has_loop_index)
    hoisted = _count([skel], tsum_leaf=True, only=lambda e: not has_loop_index(e))
    g = 0 if gtab else gsum + hoisted + ncomp * (comp + px)
    table = 0
    if ptab:
        gklo, gkhi = _x_offsets(grad, "gradient template")
        table += (obj.d + min(c + gkhi - gklo, obj.d) * wid ** (gkhi - gklo + 1)) * _count(parts)
    if tab:
        for ts in sums:
            klo, khi = _x_offsets(ts.args[0], "template sum")
            table += (1 + n_samples) * min(c + khi - klo, obj.d + ts.val) * wid ** (khi - klo + 1) * _count([ts.args[0]])
        if gtab:
            gklo, gkhi = _x_offsets(grad, "gradient template")
            table += min(c + gkhi - gklo, obj.d) * wid ** (gkhi - gklo + 1) * _count([grad])
    return {"value": v, "grad": g, "sample": v, "table": table}


def _read(name: str) -> str:
    with open(os.path.join(_CSRC, name)) as fh:
        return fh.read()


def module_source(obj: Objective, roundings=("fast",), f32: bool = False, split: int = 0) -> str:
    """Full hipRTC source; `roundings` selects the kernel variants built,
    `f32` adds the FP32 value code list of the FP32-prune precision form,
    `split` = G > 1 adds the term-parallel kernels (G lanes per box; needs
    the objective's loop templates)."""
    if split and (split < 2 or split > 64 or split & (split - 1)):
        raise ValueError("split must be a power of two in [2, 64]")
    bad = set(roundings) - {"exact", "fast", "digits15"}
    if bad or not roundings:
        raise ValueError(f"roundings must be a non-empty subset of ('exact', 'fast', 'digits15'), got {roundings}")
    parts = [
        "// generated by gpugo_amd.emit for objective " + obj.name,
        f"#define GPUGO_WITH_EXACT {int('exact' in roundings)}",
        f"#define GPUGO_WITH_FAST {int('fast' in roundings)}",
        f"#define GPUGO_WITH_DIGITS15 {int('digits15' in roundings)}",
        f"#define GPUGO_ACCESSOR_D {obj.d}",
        f"#define GPUGO_WITH_F32 {int(bool(f32))}",
        f"#define GPUGO_SPLIT_G {int(split)}",
        _read("ia.hpp"),
        _read("bnb_args.hpp"),
        _read("bnb_box.hpp"),
        code_lists(obj),
    ]
    if f32:
        parts.append(value_list_f32(obj))
    if split:
        parts.append(split_lists(obj))
    # select / compact / root of a step, so all of a step's launches use
    # this code object (and the persistent step loop can call them)
    parts += ["#define GPUGO_MODULE_STEP 1", _read("bnb_step.hpp")]
    parts.append(_read("bnb_kernel.hpp"))
    return "\n".join(parts)


def batch_module_source(objs, roundings=("fast",), mean_value: bool = False) -> str:
    """hipRTC source of a batch module: the families `objs` (variable-
    dimension objectives from problems.objective_dynamic, all with the same
    maximum dimension) behind the dispatchers fam_value / fam_kkt (/ fam_grad
    with mean_value), plus the persistent batch split+bound kernel."""
    bad = set(roundings) - {"exact", "fast"}
    if bad or not roundings:
        raise ValueError(f"roundings must be a non-empty subset of ('exact', 'fast'), got {roundings}")
    objs = list(objs)
    if not objs:
        raise ValueError("a batch module needs at least one family")
    dmax = objs[0].d
    if any(o.d != dmax for o in objs):
        raise ValueError("all families of a batch module share one maximum dimension")
    pmax = max(o.n_params for o in objs)
    psign = []
    for o in objs:
        s = param_signs(o)
        psign += s + [0] * (pmax - len(s))
    lines = [
        "// generated by gpugo_amd.emit: batch module " + ", ".join(o.name for o in objs),
        f"#define GPUGO_WITH_EXACT {int('exact' in roundings)}",
        f"#define GPUGO_WITH_FAST {int('fast' in roundings)}",
        f"#define GPUGO_BATCH_MV {int(bool(mean_value))}",
        f"#define GPUGO_DMAX {dmax}",
        f"#define GPUGO_NFAM {len(objs)}",
        f"#define GPUGO_PMAX {pmax}",
        "#define GPUGO_FAM_NPARAMS " + ", ".join(str(o.n_params) for o in objs),
        "#define GPUGO_FAM_PSIGN " + ", ".join(str(v) for v in psign),
        _read("ia.hpp"),
        _read("bnb_args.hpp"),
        _read("bnb_box.hpp"),
    ]
    for i, o in enumerate(objs):
        lines.append(code_lists(o, suffix=f"_f{i}", dynamic=True, header=False))
    val = "\n".join(f"    case {i}: return obj_value_f{i}<R>(x, P, d);" for i in range(len(objs)))
    kkt = "\n".join(f"    case {i}: return obj_kkt_f{i}<R>(x, P, dlo, dhi, d);" for i in range(len(objs)))
    grd = "\n".join(f"    case {i}: obj_grad_f{i}<R>(x, P, g, d); return;" for i in range(len(objs)))
    lines.append(f"""namespace gpugo {{
template <int R, class XS>
__device__ __forceinline__ iv<double> fam_value(int f, const XS &x, const double *P, int d) {{
  switch (f) {{
{val}
    default: return mk<double>(-__builtin_huge_val(), __builtin_huge_val());
  }}
}}
template <int R, class XS>
__device__ __forceinline__ bool fam_kkt(int f, const XS &x, const double *P, const double *dlo,
                                        const double *dhi, int d) {{
  switch (f) {{
{kkt}
    default: return false;
  }}
}}
template <int R, class XS, class GS>
__device__ __forceinline__ void fam_grad(int f, const XS &x, const double *P, GS &&g, int d) {{
  switch (f) {{
{grd}
    default: return;
  }}
}}
}}  // namespace gpugo""")
    lines.append(_read("bnb_batch.hpp"))
    return "\n".join(lines)


def source_key(src: str, options: str = "") -> str:
    return hashlib.sha256((options + "\0" + src).encode()).hexdigest()[:32]
