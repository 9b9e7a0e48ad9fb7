"""The reference's benchmark objectives as expression-IR objectives.

Each entry follows the code list of the reference script named in its
docstring (paths relative to GPUGO_1.1/), with the box and initial upper
bound of that script.  Constants are the double values the reference uses;
gradients are derived symbolically by the emitter (the reference
hand-writes them).  Rastrigin, Rosenbrock and Ackley are special cases of
the coefficient families in problems.py; they are listed here too with the
reference's constants.

Every function has a closed-form global minimum on its box (f_star at
x_star), which the tests use as the check of certified enclosures.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, List

import numpy as np

from .expr import E, PI, Const, J, JNum, Objective, TProd, TProdEx, TSum, TWhen, X, cos, exp, fsum, sin, sqr, sqrt
from .problems import ProblemSpec


def fprod(terms) -> "Expr":
    """Left fold of interval products: ((t0 * t1) * t2) * ..."""
    it = iter(terms)
    acc = next(it)
    for t in it:
        acc = acc * t
    return acc


def _ackley(d):
    # 1-Ackley/ackley50.py:171-212
    s1 = fsum(sqr(X(i)) for i in range(d))
    s2 = fsum(cos(X(i) * (PI * 2.0)) for i in range(d))
    return ((exp(sqrt(s1 / float(d)) * -0.02) * -20.0 - exp(s2 / float(d))) + 20.0) + E


def _belegundu(d):
    # 2-Belegundu/deflected50.py:175-199
    s = fsum(sqr(X(i) - 5.0) for i in range(d))
    return s * 0.1 - cos(sqrt(s) * 5.0)


def _breiman(d):
    # 3-Breiman/cosine50.py:164-188 (the value code list; its Sample list
    # omits the quadratic term, :283-299, which is not reproduced)
    c = fsum(cos(X(i) * (PI * 5.0)) for i in range(d))
    q = fsum(sqr(X(i)) for i in range(d))
    return c * -0.1 + q


def _fu(d):
    # 4-Fu/trigonometric50.py:168-209
    terms = []
    for i in range(d):
        t = sqr(X(i) - 0.9)
        terms.append((sqr(sin(t * 7.0)) * 8.0 + sqr(sin(t * 14.0)) * 6.0) + t)
    return fsum(terms) + 1.0


def _griewank(d):
    # 5-Griewank/griewank50.py:166-198 (coordinates numbered from 1)
    s = fsum(sqr(X(i)) for i in range(d)) / 4000.0
    p = fprod(cos(X(i) / sqrt(Const(float(i + 1)))) for i in range(d))
    return (s + 1.0) - p


def _levy(d):
    # 6-Levy/levy50.py:166-224: (pi/n) [10 sin^2(pi/4 x_1 + 3pi/4)
    #   + sum_{i<n} ((x_i - 1)/4)^2 (1 + 10 sin^2(pi/4 x_{i+1} + 3pi/4))
    #   + ((x_n - 1)/4)^2]
    q = PI * 0.25
    t3 = PI * 0.75
    head = sqr(sin(X(0) * q + t3)) * 10.0
    body = [sqr((X(i) - 1.0) * 0.25) * (sqr(sin(X(i + 1) * q + t3)) * 10.0 + 1.0) for i in range(d - 1)]
    tail = sqr((X(d - 1) - 1.0) * 0.25)
    return (PI / float(d)) * fsum([head] + body + [tail])


def _rastrigin(d):
    # 7-Rastrigin/rastrigin50.py:167-188
    return fsum(sqr(X(i)) - cos(X(i) * (PI * 2.0)) * 10.0 for i in range(d)) + 10.0 * d


def _salomon(d):
    # 8-Salomon/salomon50.py:167-189
    r = sqrt(fsum(sqr(X(i)) for i in range(d)))
    return (1.0 - cos(r * (PI * 2.0))) + r * 0.1


def _styblinski(d):
    # 9-Styblinski/corrupted50.py:173-200
    s = fsum(sqr(X(i)) for i in range(d)) / (2.0 * d)
    return s - fprod(cos(X(i)) for i in range(d)) * (4.0 * d)


def _zabinsky(d):
    # 10-Zabinsky/sinusoidal50.py:169-194
    z = PI / 6.0
    p1 = fprod(sin(X(i) - z) for i in range(d))
    p2 = fprod(sin((X(i) - z) * 5.0) for i in range(d))
    return -(p1 * 2.5 + p2)


def _rosenbrock(d):
    # 11-Rosenbrock/rosenbrock50.py:170-198
    return fsum(sqr(X(i + 1) - sqr(X(i))) * 100.0 + sqr(X(i) - 1.0) for i in range(d - 1))


# Loop templates (value with template sums, one gradient component in J) of
# the families whose terms are sums: the same functions as the unrolled
# code lists above (checked in 256-bit arithmetic by the tests), used by
# the term-parallel and the windowed (large-d) code lists.  Gradients are
# derived by hand from the definitions.
def _t_ackley(d):
    s1 = TSum(sqr(X(J)))
    s2 = TSum(cos(X(J) * (PI * 2.0)))
    r = sqrt(s1 / float(d))
    t1 = exp(r * -0.02)
    t2 = exp(s2 / float(d))
    value = ((t1 * -20.0 - t2) + 20.0) + E
    k1 = (((t1 / r) * 0.02) * 20.0) / float(d)
    k2 = (t2 * (PI * 2.0)) / float(d)
    return {"value": value, "grad": k1 * X(J) + k2 * sin(X(J) * (PI * 2.0))}


def _t_belegundu(d):
    s = TSum(sqr(X(J) - 5.0))
    r = sqrt(s)
    return {"value": s * 0.1 - cos(r * 5.0),
            "grad": (X(J) - 5.0) * 0.2 + ((sin(r * 5.0) * 5.0) / r) * (X(J) - 5.0)}


def _t_breiman(d):
    return {"value": TSum(cos(X(J) * (PI * 5.0))) * -0.1 + TSum(sqr(X(J))),
            "grad": (sin(X(J) * (PI * 5.0)) * (PI * 5.0)) * 0.1 + X(J) * 2.0}


def _t_fu(d):
    t = sqr(X(J) - 0.9)
    body = (sqr(sin(t * 7.0)) * 8.0 + sqr(sin(t * 14.0)) * 6.0) + t
    return {"value": TSum(body) + 1.0,
            "grad": ((X(J) - 0.9) * 2.0) * ((sin(t * 14.0) * 56.0 + sin(t * 28.0) * 84.0) + 1.0)}


def _t_levy(d):
    q, t3 = PI * 0.25, PI * 0.75
    head = sqr(sin(X(0) * q + t3)) * 10.0
    body = sqr((X(J) - 1.0) * 0.25) * (sqr(sin(X(J + 1) * q + t3)) * 10.0 + 1.0)
    tail = sqr((X(d - 1) - 1.0) * 0.25)
    # d/dx 10 sin^2(pi/4 x + 3pi/4) = (5pi/2) sin(pi/2 x + 3pi/2)
    s2 = sin(X(J) * (PI * 0.5) + PI * 1.5) * (PI * 2.5)
    g = ((TWhen("j==0", s2)
          + TWhen("j+1<d", ((X(J) - 1.0) * 0.125) * (sqr(sin(X(J + 1) * q + t3)) * 10.0 + 1.0)))
         + TWhen("j>0", sqr((X(J - 1) - 1.0) * 0.25) * s2)) + TWhen("j+1==d", (X(J) - 1.0) * 0.125)
    return {"value": (PI / float(d)) * ((head + TSum(body, off=-1)) + tail), "grad": (PI / float(d)) * g}


def _t_rastrigin(d):
    return {"value": TSum(sqr(X(J)) - cos(X(J) * (PI * 2.0)) * 10.0) + 10.0 * d,
            "grad": X(J) * 2.0 + sin(X(J) * (PI * 2.0)) * (PI * 20.0)}


def _t_salomon(d):
    r = sqrt(TSum(sqr(X(J))))
    return {"value": (1.0 - cos(r * (PI * 2.0))) + r * 0.1,
            "grad": ((sin(r * (PI * 2.0)) * (PI * 2.0) + 0.1) / r) * X(J)}


def _t_rosenbrock(d):
    return {"value": TSum(sqr(X(J + 1) - sqr(X(J))) * 100.0 + sqr(X(J) - 1.0), off=-1),
            "grad": TWhen("j+1<d", (X(J) * (X(J + 1) - sqr(X(J)))) * -400.0 + (X(J) - 1.0) * 2.0)
            + TWhen("j>0", (X(J) - sqr(X(J - 1))) * 200.0)}


# Product families: the product is a template product; its j-th partial
# derivative keeps the product of the other factors (TProdEx), which the
# windowed lists form from the parent's rest product and the window's other
# terms.
def _t_griewank(d):
    rj = sqrt(JNum(1))          # sqrt(j + 1): coordinates numbered from 1
    return {"value": (TSum(sqr(X(J))) / 4000.0 + 1.0) - TProd(cos(X(J) / rj)),
            "grad": X(J) / 2000.0 + (sin(X(J) / rj) / rj) * TProdEx(cos(X(J) / rj))}


def _t_styblinski(d):
    return {"value": TSum(sqr(X(J))) / (2.0 * d) - TProd(cos(X(J))) * (4.0 * d),
            "grad": X(J) / float(d) + (sin(X(J)) * TProdEx(cos(X(J)))) * (4.0 * d)}


def _t_zabinsky(d):
    z = PI / 6.0
    u, v = X(J) - z, (X(J) - z) * 5.0
    return {"value": -(TProd(sin(u)) * 2.5 + TProd(sin(v))),
            "grad": -((cos(u) * TProdEx(sin(u))) * 2.5 + (cos(v) * 5.0) * TProdEx(sin(v)))}


TEMPLATES = {"ackley": _t_ackley, "belegundu": _t_belegundu, "breiman": _t_breiman, "fu": _t_fu,
             "levy": _t_levy, "rastrigin": _t_rastrigin, "salomon": _t_salomon, "rosenbrock": _t_rosenbrock,
             "griewank": _t_griewank, "styblinski": _t_styblinski, "zabinsky": _t_zabinsky}


@dataclass
class RefFamily:
    build: Callable[[int], "Expr"]
    lo: float
    hi: float
    init_ub: float
    x_star: Callable[[int], np.ndarray]
    f_star: Callable[[int], float]
    source: str


REFERENCE: Dict[str, RefFamily] = {
    "ackley": RefFamily(_ackley, -35.0, 40.0, 1000.0, lambda d: np.zeros(d), lambda d: 0.0,
                        "1-Ackley/ackley50.py:171-212, box :958, ub :994"),
    "belegundu": RefFamily(_belegundu, -10.0, 11.0, 1e6, lambda d: np.full(d, 5.0), lambda d: -1.0,
                           "2-Belegundu/deflected50.py:175-199, box :906, ub :943"),
    "breiman": RefFamily(_breiman, -1.0, 2.0, 1e6, lambda d: np.zeros(d), lambda d: -0.1 * d,
                         "3-Breiman/cosine50.py:164-188, box :857, ub :894"),
    "fu": RefFamily(_fu, -10.0, 10.0, 1e6, lambda d: np.full(d, 0.9), lambda d: 1.0,
                    "4-Fu/trigonometric50.py:168-209, box :923, ub :960"),
    "griewank": RefFamily(_griewank, -100.0, 110.0, 1e6, lambda d: np.zeros(d), lambda d: 0.0,
                          "5-Griewank/griewank50.py:166-198, box :913, ub :949"),
    "levy": RefFamily(_levy, -10.0, 10.0, 1000.0, lambda d: np.ones(d), lambda d: 0.0,
                      "6-Levy/levy50.py:166-224, box :1041, ub :1076"),
    "rastrigin": RefFamily(_rastrigin, -5.5, 6.0, 1e6, lambda d: np.zeros(d), lambda d: 0.0,
                           "7-Rastrigin/rastrigin50.py:167-188, box :852, ub :888"),
    "salomon": RefFamily(_salomon, -100.0, 110.0, 1e6, lambda d: np.zeros(d), lambda d: 0.0,
                         "8-Salomon/salomon50.py:167-189, box :882, ub :918"),
    "styblinski": RefFamily(_styblinski, -10.0, 11.0, 1e6, lambda d: np.zeros(d), lambda d: -4.0 * d,
                            "9-Styblinski/corrupted50.py:173-200, box :917, ub :954"),
    "zabinsky": RefFamily(_zabinsky, 0.0, math.pi, 1e6, lambda d: np.full(d, 2.0 * math.pi / 3.0),
                          lambda d: -3.5, "10-Zabinsky/sinusoidal50.py:169-194, box :922, ub :958"),
    "rosenbrock": RefFamily(_rosenbrock, -30.0, 30.0, 1e9, lambda d: np.ones(d), lambda d: 0.0,
                            "11-Rosenbrock/rosenbrock50.py:170-198, box :1059, ub :1096"),
}


def objective(name: str, d: int) -> Objective:
    """The reference objective `name` in d dimensions (no coefficients),
    with its loop templates where the family has them."""
    fam = REFERENCE[name]
    tpl = TEMPLATES[name](d) if name in TEMPLATES else None
    return Objective(d, 0, fam.build(d), None, name=f"ref_{name}_d{d}", templates=tpl)


def objective_loops(name: str, d: int) -> Objective:
    """The reference objective `name` in loop form only (no unrolled code
    lists; for the windowed large-d modules, whose size does not grow with d).
    The unrolled value is not built: its `value` is the template."""
    if name not in TEMPLATES:
        raise ValueError(f"{name} has no loop templates")
    tpl = TEMPLATES[name](d)
    return Objective(d, 0, tpl["value"], None, name=f"ref_{name}_loops_d{d}", templates=tpl)


def spec(name: str, d: int, tol: float = 1e-6) -> ProblemSpec:
    fam = REFERENCE[name]
    return ProblemSpec(name, d, np.zeros(0), np.full(d, fam.lo), np.full(d, fam.hi), tol, f_star=fam.f_star(d),
                       name=f"{name} d={d}", extra={"init_ub": fam.init_ub, "x_star": fam.x_star(d)})


def names() -> List[str]:
    return list(REFERENCE)
