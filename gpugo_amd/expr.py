"""Expression IR for objectives.

An objective is written once as an expression of the coordinates ``X[j]``
and the coefficient vector ``P[i]``; gpugo_amd.emit lowers it to interval
code lists (value, gradient, monotonicity test) in HIP.  This replaces the
per-script hand-written device functions <fn>Value_IA / <fn>Derivative_IA /
<fn>Sample_IA of the reference (GPUGO_1.1/1-Ackley/ackley50.py:116-479).

Values are intervals except *exact scalars*: literal doubles, parameters,
their negations and their products with powers of two, which are exact in
binary floating point and are passed to the interval ops as scalars
(``scale``, ``addc``, ``rsubc``, ``divc``).  Anything else is evaluated with
outward rounding.  Irrational constants are written as ``IConst(lo, hi)``
with bracketing doubles (``PI``, ``E``).
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

_OPS_UNARY = ("neg", "sqr", "sqrt", "exp", "sin", "cos")
_OPS_BINARY = ("add", "sub", "mul", "div")


class Expr:
    __slots__ = ("op", "args", "val", "val2", "_key", "_hash")

    def __init__(self, op: str, args: tuple = (), val=None, val2=None):
        self.op = op
        self.args = args
        self.val = val
        self.val2 = val2      # param: declared sign (+1 / -1) or None
        self._key = None
        self._hash = None

    # structural identity (used for common-subexpression elimination)
    def key(self):
        if self._key is None:
            self._key = (self.op, self.val, self.val2, tuple(a.key() for a in self.args))
        return self._key

    def __hash__(self):
        if self._hash is None:
            self._hash = hash(self.key())
        return self._hash

    def __eq__(self, other):  # structural
        return isinstance(other, Expr) and self.key() == other.key()

    # arithmetic
    def __add__(self, o): return Expr("add", (self, wrap(o)))
    def __radd__(self, o): return Expr("add", (wrap(o), self))
    def __sub__(self, o): return Expr("sub", (self, wrap(o)))
    def __rsub__(self, o): return Expr("sub", (wrap(o), self))
    def __mul__(self, o): return Expr("mul", (self, wrap(o)))
    def __rmul__(self, o): return Expr("mul", (wrap(o), self))
    def __truediv__(self, o): return Expr("div", (self, wrap(o)))
    def __rtruediv__(self, o): return Expr("div", (wrap(o), self))
    def __neg__(self): return Expr("neg", (self,))

    def __pow__(self, n):
        if not isinstance(n, int) or n < 0:
            raise ValueError("only non-negative integer powers are supported")
        if n == 2:
            return sqr(self)
        return Expr("pown", (self,), n)

    def __repr__(self):
        if self.op == "var":
            return f"X[{self.val}]"
        if self.op == "param":
            return f"P[{self.val}]"
        if self.op == "const":
            return repr(self.val)
        if self.op == "iconst":
            return f"[{self.val[0]!r},{self.val[1]!r}]"
        if self.op == "pown":
            return f"pown({self.args[0]!r},{self.val})"
        return f"{self.op}({', '.join(repr(a) for a in self.args)})"

    # exact scalars -------------------------------------------------------
    def is_scalar(self) -> bool:
        """True if the value is a single double known exactly (no rounding)."""
        if self.op in ("const", "param", "dim", "jnum"):
            return True
        if self.op == "neg":
            return self.args[0].is_scalar()
        if self.op == "mul":
            a, b = self.args
            return (a.is_scalar() and b.is_scalar() and
                    (_is_pow2_const(a) or _is_pow2_const(b)))
        return False

    def depends_on_x(self) -> bool:
        if self.op == "var":
            return True
        return any(a.depends_on_x() for a in self.args)

    def scalar_sign(self):
        """+1 / -1 if the sign of an exact scalar is known statically, else 0."""
        if self.op == "const":
            return 1 if self.val > 0 else (-1 if self.val < 0 else 0)
        if self.op == "dim":
            return 1
        if self.op == "jnum":
            return 1 if self.val > 0 else 0
        if self.op == "param":
            return 0 if self.val2 is None else self.val2
        if self.op == "neg":
            return -self.args[0].scalar_sign()
        if self.op == "mul":
            return self.args[0].scalar_sign() * self.args[1].scalar_sign()
        return 0


def _is_pow2_const(e: Expr) -> bool:
    if e.op != "const" or e.val == 0 or not math.isfinite(e.val):
        return False
    m, _ = math.frexp(abs(e.val))
    return m == 0.5


def wrap(v) -> Expr:
    if isinstance(v, Expr):
        return v
    if isinstance(v, (int, float)):
        f = float(v)
        if isinstance(v, int) and float(int(f)) != v:
            raise ValueError(f"integer constant {v} is not exact in double")
        return Expr("const", (), f)
    raise TypeError(f"cannot use {type(v).__name__} in an objective expression")


class Idx:
    """Affine index a*J + b of the loop variable J of a template (TSum and
    gradient-component templates); X(J + 1), P(J + d) etc."""
    __slots__ = ("a", "b")

    def __init__(self, a: int = 1, b: int = 0):
        self.a, self.b = int(a), int(b)

    def __add__(self, k: int): return Idx(self.a, self.b + int(k))
    __radd__ = __add__
    def __sub__(self, k: int): return Idx(self.a, self.b - int(k))
    def __mul__(self, k: int): return Idx(self.a * int(k), self.b * int(k))
    __rmul__ = __mul__

    def at(self, j: int) -> int:
        return self.a * j + self.b


J = Idx()


def X(j) -> Expr:
    if isinstance(j, Idx):
        return Expr("var", (), ("J", j.a, j.b))
    return Expr("var", (), int(j))


def P(i, sign: int | None = None) -> Expr:
    """Coefficient P[i].  `sign` = +1 (or -1) declares P[i] >= 0 (<= 0); the
    emitter then uses the one-sided scalar product, and problems whose
    coefficients violate a declared sign are rejected."""
    if sign not in (None, 1, -1):
        raise ValueError("sign must be None, +1 or -1")
    if isinstance(i, Idx):
        return Expr("param", (), ("J", i.a, i.b), sign)
    return Expr("param", (), int(i), sign)


def TSum(body, off: int = 0) -> Expr:
    """Template sum: sum over j in [0, d + off) of `body` with J = j (left
    fold in increasing j).  Used by the term-parallel code lists, where the
    lanes of a group take j = lane, lane + G, ... and reduce."""
    return Expr("tsum", (wrap(body),), int(off))


def TProd(body, off: int = 0) -> Expr:
    """Template product: product over j in [0, d + off) of `body` with
    J = j (left fold in increasing j); the product form of TSum (Griewank,
    Styblinski-Tang, Zabinsky)."""
    return Expr("tprod", (wrap(body),), int(off))


def TProdEx(body, off: int = 0) -> Expr:
    """Inside a gradient template: the product over i in [0, d + off),
    i != J, of `body` with J = i (the factor a product's j-th partial
    derivative leaves).  `body` binds its own J; the excluded index is the
    enclosing template's J."""
    return Expr("tprodx", (wrap(body),), int(off))


def JNum(b: int = 0) -> Expr:
    """Inside a template: the loop index j + b as an exact scalar."""
    return Expr("jnum", (), int(b))


TEMPLATE_REDUCTIONS = ("tsum", "tprod")

TWHEN_CONDS = ("j+1<d", "j>0", "j==0", "j+1==d")


def TWhen(cond: str, body) -> Expr:
    """Inside a template: `body` where the condition on j holds, else 0.
    cond is one of TWHEN_CONDS ("j+1<d", "j>0", "j==0", "j+1==d")."""
    if cond not in TWHEN_CONDS:
        raise ValueError(f"unknown template condition {cond!r}")
    return Expr("twhen", (wrap(body),), cond)


def has_loop_index(e: Expr) -> bool:
    """True if e depends on the loop variable J of an enclosing template
    (a TSum binds J inside itself)."""
    if e.op in ("var", "param") and isinstance(e.val, tuple):
        return True
    if e.op in ("jnum", "tprodx", "tpart"):
        return True
    if e.op in TEMPLATE_REDUCTIONS:
        return False
    return any(has_loop_index(a) for a in e.args)


def Dim() -> Expr:
    """The problem dimension d as an exact scalar (runtime value in the
    variable-dimension code lists of a batch)."""
    return Expr("dim")


def Guarded(j: int, e) -> Expr:
    """A term that exists only when coordinate j exists (j < d); inside a
    sum it is skipped otherwise.  Used by variable-dimension code lists."""
    return Expr("guard", (wrap(e),), int(j))


def gsum(terms) -> Expr:
    """Left fold like fsum, allowing Guarded terms (skipped when absent)."""
    terms = [wrap(t) for t in terms]
    if not terms:
        return Const(0.0)
    if not any(t.op == "guard" for t in terms):
        return fsum(terms)
    return Expr("gsum", tuple(terms))


def Const(v: float) -> Expr:
    return Expr("const", (), float(v))


def IConst(lo: float, hi: float) -> Expr:
    if not lo <= hi:
        raise ValueError("IConst needs lo <= hi")
    return Expr("iconst", (), (float(lo), float(hi)))


PI = IConst(float.fromhex("0x1.921fb54442d18p+1"), float.fromhex("0x1.921fb54442d19p+1"))
E = IConst(float.fromhex("0x1.5bf0a8b145769p+1"), float.fromhex("0x1.5bf0a8b14576ap+1"))


def sqr(e) -> Expr: return Expr("sqr", (wrap(e),))
def sqrt(e) -> Expr: return Expr("sqrt", (wrap(e),))
def exp(e) -> Expr: return Expr("exp", (wrap(e),))
def sin(e) -> Expr: return Expr("sin", (wrap(e),))
def cos(e) -> Expr: return Expr("cos", (wrap(e),))


def pown(e, n: int) -> Expr:
    if n == 2:
        return sqr(e)
    return Expr("pown", (wrap(e),), int(n))


def fsum(terms: Iterable) -> Expr:
    """Left fold in the given order: ((t0 + t1) + t2) + ..."""
    it = iter(terms)
    try:
        acc = wrap(next(it))
    except StopIteration:
        return Const(0.0)
    for t in it:
        acc = acc + t
    return acc


class Objective:
    """An objective with its code lists.

    value     Expr of X[0..d-1] and P[0..n_params-1]
    grad      list of d Exprs (exact derivative enclosures); if None the
              emitter differentiates ``value`` symbolically
    name      used for the hipRTC program name and the module cache key
    """

    def __init__(self, d: int, n_params: int, value: Expr, grad: Sequence[Expr] | None = None,
                 name: str = "objective", templates: dict | None = None):
        self.d = int(d)
        self.n_params = int(n_params)
        self.value = wrap(value)
        self.grad = list(grad) if grad is not None else None
        self.name = name
        # optional loop form of the same function for the term-parallel code
        # lists: {"value": Expr with TSum nodes, "grad": component template
        # in J (j in [0, d))}
        self.templates = templates
        if self.grad is not None and len(self.grad) != self.d:
            raise ValueError(f"gradient has {len(self.grad)} components, d={self.d}")

    def check(self):
        def walk(e):
            if e.op == "var" and not 0 <= e.val < self.d:
                raise ValueError(f"X[{e.val}] out of range for d={self.d}")
            if e.op == "param" and not 0 <= e.val < self.n_params:
                raise ValueError(f"P[{e.val}] out of range for n_params={self.n_params}")
            for a in e.args:
                walk(a)
        walk(self.value)
        for g in self.grad or ():
            walk(wrap(g))
