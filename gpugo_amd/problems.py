"""Objective families and the seeded benchmark configurations.

Three families cover the benchmark configurations and, as special cases,
three of the reference's objectives:

  sepcos        f = sum_i a_i x_i^2 + b_i (1 - cos(c_i x_i))
                params [a(d), b(d), c(d)]; Rastrigin
                (GPUGO_1.1/7-Rastrigin/rastrigin50.py:167-188) is a=1, b=10,
                c=2pi (x^2 - 10 cos(2 pi x) + 10 per coordinate).
  sepcos_chain  sepcos + sum_{i<d-1} w_i (x_{i+1} - x_i^2)^2 + (1 - x_i)^2
                params [a(d), b(d), c(d), w(d-1)]; Rosenbrock
                (GPUGO_1.1/11-Rosenbrock/rosenbrock50.py:170-198) is
                a=b=0, w=100.
  ackley_q      f = -A exp(-B sqrt(mean x^2)) - exp(mean cos(C x)) + A + e
                    + sum_i q_i x_i^4 / 50
                params [A, B, C, q(d)]; Ackley
                (GPUGO_1.1/1-Ackley/ackley50.py:171-212) is A=20, B=0.02,
                C=2pi, q=0.

Gradients are given explicitly (as the reference hand-writes
<fn>Derivative_IA); sums are left folds in increasing index.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .expr import E, J, Dim, Guarded, Objective, P, TSum, TWhen, X, cos, exp, fsum, gsum, pown, sin, sqr, sqrt

FAMILIES = ("sepcos", "sepcos_chain", "ackley_q")


def n_params(family: str, d: int) -> int:
    return {"sepcos": 3 * d, "sepcos_chain": 4 * d - 1, "ackley_q": 3 + d}[family]


def _build(family: str, d: int, dynamic: bool) -> Objective:
    """Code lists of a family.  dynamic=False: fixed dimension d.  dynamic=True:
    d is the maximum dimension, the runtime dimension is Dim(), coordinates
    j >= Dim() are absent (Guarded) and coefficients use the layout of d.
    Both forms perform the same operations in the same order for the
    coordinates that exist.  All coefficients are declared >= 0."""
    fold = gsum if dynamic else fsum
    G = (lambda j, e: Guarded(j, e) if (dynamic and j > 0) else e)
    nd = Dim() if dynamic else float(d)
    p = lambda i: P(i, +1)
    if family in ("sepcos", "sepcos_chain"):
        a, b, c = (lambda i: p(i)), (lambda i: p(d + i)), (lambda i: p(2 * d + i))
        sep = [G(i, sqr(X(i)) * a(i) + (1.0 - cos(X(i) * c(i))) * b(i)) for i in range(d)]
        sep_g = lambda j: X(j) * (2.0 * a(j)) + (sin(X(j) * c(j)) * c(j)) * b(j)
        if family == "sepcos":
            value = fold(sep)
            grad = [sep_g(j) for j in range(d)]
        else:
            w = lambda i: p(3 * d + i)
            chain = [G(i + 1, sqr(X(i + 1) - sqr(X(i))) * w(i) + sqr(1.0 - X(i))) for i in range(d - 1)]
            value = fold(sep + chain)
            grad = []
            for j in range(d):
                terms = [sep_g(j)]
                if j + 1 < d:
                    terms.append(G(j + 1, (X(j) * (X(j + 1) - sqr(X(j)))) * (-4.0 * w(j))
                                   - (1.0 - X(j)) * 2.0))
                if j > 0:
                    terms.append((X(j) - sqr(X(j - 1))) * (2.0 * w(j - 1)))
                grad.append(fold(terms))
    elif family == "ackley_q":
        A, B, C = p(0), p(1), p(2)
        q = lambda i: p(3 + i)
        s1 = fold(G(i, sqr(X(i))) for i in range(d))
        s2 = fold(G(i, cos(X(i) * C)) for i in range(d))
        s3 = fold(G(i, pown(X(i), 4) * q(i)) for i in range(d))
        r = sqrt(s1 / nd)
        t1 = exp(r * (-B))
        v1 = t1 * (-A)
        t2 = exp(s2 / nd)
        value = ((v1 - t2) + (A + E)) + s3 / 50.0
        k1 = (((t1 / r) * B) * A) / nd
        k2 = (exp(s2 / nd) * C) / nd
        grad = [(k1 * X(j) + k2 * sin(X(j) * C)) + (pown(X(j), 3) * (4.0 * q(j))) / 50.0
                for j in range(d)]
    else:
        raise ValueError(f"unknown family {family!r}")
    tag = "dyn" if dynamic else "d"
    return Objective(d, n_params(family, d), value, grad, name=f"{family}_{tag}{d}",
                     templates=_templates(family, d))


def _templates(family: str, d: int) -> dict:
    """The same family as loops over j (for the term-parallel code lists):
    the value with TSum nodes and one gradient component in J.  d is the
    coefficient layout (the maximum dimension); loops run to the runtime d."""
    p = lambda i: P(i, +1)
    if family in ("sepcos", "sepcos_chain"):
        a, b, c = (lambda i: p(i)), (lambda i: p(i + d)), (lambda i: p(i + 2 * d))
        sep = sqr(X(J)) * a(J) + (1.0 - cos(X(J) * c(J))) * b(J)
        sep_g = X(J) * (2.0 * a(J)) + (sin(X(J) * c(J)) * c(J)) * b(J)
        if family == "sepcos":
            return {"value": TSum(sep), "grad": sep_g}
        w = lambda i: p(i + 3 * d)
        chain = sqr(X(J + 1) - sqr(X(J))) * w(J) + sqr(1.0 - X(J))
        fwd = (X(J) * (X(J + 1) - sqr(X(J)))) * (-4.0 * w(J)) - (1.0 - X(J)) * 2.0
        bwd = (X(J) - sqr(X(J - 1))) * (2.0 * w(J - 1))
        return {"value": TSum(sep) + TSum(chain, off=-1),
                "grad": (sep_g + TWhen("j+1<d", fwd)) + TWhen("j>0", bwd)}
    A, B, C = p(0), p(1), p(2)
    q = lambda i: p(i + 3)
    nd = Dim()
    s1 = TSum(sqr(X(J)))
    s2 = TSum(cos(X(J) * C))
    s3 = TSum(pown(X(J), 4) * q(J))
    r = sqrt(s1 / nd)
    t1 = exp(r * (-B))
    value = ((t1 * (-A) - exp(s2 / nd)) + (A + E)) + s3 / 50.0
    k1 = (((t1 / r) * B) * A) / nd
    k2 = (exp(s2 / nd) * C) / nd
    grad = (k1 * X(J) + k2 * sin(X(J) * C)) + (pown(X(J), 3) * (4.0 * q(J))) / 50.0
    return {"value": value, "grad": grad}


# Above this dimension the fixed-d code lists use the guarded form too: the
# per-coordinate branches on the runtime d (always taken) split the code
# into short blocks, so coordinates decoded on use do not stay live.
GUARDED_D = 8


def objective(family: str, d: int) -> Objective:
    return _build(family, d, d > GUARDED_D)


def objective_loops(family: str, d: int) -> Objective:
    """A family in loop form only (its templates; no unrolled code lists),
    for the windowed large-d modules."""
    tpl = _templates(family, d)
    return Objective(d, n_params(family, d), tpl["value"], None, name=f"{family}_loops_d{d}", templates=tpl)


def objective_dynamic(family: str, dmax: int) -> Objective:
    """Variable-dimension form (runtime d <= dmax) used by batches."""
    return _build(family, dmax, True)


def pad_params(family: str, d: int, params, dmax: int) -> np.ndarray:
    """Coefficients of a d-dimensional problem in the dmax layout of
    objective_dynamic (absent coordinates get 0)."""
    params = np.asarray(params, dtype=np.float64)
    out = np.zeros(n_params(family, dmax))
    if family == "ackley_q":
        out[:3] = params[:3]
        out[3:3 + d] = params[3:3 + d]
        return out
    blocks = 4 if family == "sepcos_chain" else 3
    for k in range(blocks):
        n = d - 1 if k == 3 else d
        out[k * dmax:k * dmax + n] = params[k * d:k * d + n]
    return out


# ---------------------------------------------------------------------------
# seeded problem instances
# ---------------------------------------------------------------------------
@dataclass
class ProblemSpec:
    family: str
    d: int
    params: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    tol: float
    f_star: Optional[float] = None      # closed form where known
    name: str = ""
    grid: tuple = (2, 10)               # (wid, cycle) default schedule
    k_select: int = 64
    extra: dict = field(default_factory=dict)


def draw_sepcos(rng, d):
    a = rng.uniform(0.5, 2.0, d)
    b = rng.uniform(1.0, 10.0, d)
    c = rng.uniform(1.0, 2.0 * math.pi, d)
    return a, b, c


def sepcos2(seed: int = 0) -> ProblemSpec:
    """d=2 separable, box [-5.12,5.12]^2, 64x64 grid, tol 1e-6; f* = 0 at x = 0."""
    rng = np.random.default_rng(seed)
    d = 2
    a, b, c = draw_sepcos(rng, d)
    return ProblemSpec("sepcos", d, np.concatenate([a, b, c]), np.full(d, -5.12), np.full(d, 5.12),
                       1e-6, f_star=0.0, name="sepcos d=2", grid=(64, 2), k_select=1)


def chain10(seed: int = 1) -> ProblemSpec:
    """d=10 separable + chain, box [-5,10]^10, reference grid 4^10, tol 1e-6."""
    rng = np.random.default_rng(seed)
    d = 10
    a, b, c = draw_sepcos(rng, d)
    w = rng.uniform(1.0, 100.0, d - 1)
    return ProblemSpec("sepcos_chain", d, np.concatenate([a, b, c, w]), np.full(d, -5.0),
                       np.full(d, 10.0), 1e-6, name="sepcos_chain d=10", grid=(4, 10), k_select=1)


def ackley50(seed: int = 2, d: int = 50) -> ProblemSpec:
    """d=50 Ackley-type + sum q x^4/50, box [-32,32]^50, tol 1e-6; f* = 0 at x = 0."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(10.0, 30.0)
    B = rng.uniform(0.1, 0.3)
    C = rng.uniform(math.pi, 3.0 * math.pi)
    q = rng.uniform(0.0, 0.1, d)
    return ProblemSpec("ackley_q", d, np.concatenate([[A, B, C], q]), np.full(d, -32.0),
                       np.full(d, 32.0), 1e-6, f_star=0.0, name=f"ackley_q d={d}", grid=(3, 12),
                       k_select=1)


def _family_params(rng, family: str, d: int) -> np.ndarray:
    """Coefficients of a family drawn as in the single-problem workloads."""
    if family == "sepcos":
        return np.concatenate(draw_sepcos(rng, d))
    if family == "sepcos_chain":
        a, b, c = draw_sepcos(rng, d)
        return np.concatenate([a, b, c, rng.uniform(1.0, 100.0, d - 1)])
    A = rng.uniform(10.0, 30.0)
    B = rng.uniform(0.1, 0.3)
    C = rng.uniform(math.pi, 3.0 * math.pi)
    return np.concatenate([[A, B, C], rng.uniform(0.0, 0.1, d)])


def batch_problem(k: int) -> ProblemSpec:
    """Problem k of the 4096-problem batch: family uniform over the three
    families, d = round(exp(U[ln 2, ln 32])), box [-h, h]^d with h ~ U[2, 32],
    tol in {1e-4, 1e-6}; drawn from default_rng([3, k])."""
    rng = np.random.default_rng([3, k])
    family = FAMILIES[int(rng.integers(0, 3))]
    d = int(round(math.exp(rng.uniform(math.log(2.0), math.log(32.0)))))
    h = rng.uniform(2.0, 32.0)
    tol = (1e-4, 1e-6)[int(rng.integers(0, 2))]
    params = _family_params(rng, family, d)
    f_star = None if family == "sepcos_chain" else 0.0
    return ProblemSpec(family, d, params, np.full(d, -h), np.full(d, h), tol, f_star=f_star,
                       name=f"batch#{k} {family} d={d}")


def served_problem(k: int) -> ProblemSpec:
    """Problem k of the served workload: family uniform, d log-uniform in
    [2, 100], tol log-uniform in [1e-8, 1e-4], box [-h, h]^d with
    h ~ U[2, 32]; drawn from default_rng([4, k])."""
    rng = np.random.default_rng([4, k])
    family = FAMILIES[int(rng.integers(0, 3))]
    d = int(round(math.exp(rng.uniform(math.log(2.0), math.log(100.0)))))
    tol = float(math.exp(rng.uniform(math.log(1e-8), math.log(1e-4))))
    h = rng.uniform(2.0, 32.0)
    params = _family_params(rng, family, d)
    f_star = None if family == "sepcos_chain" else 0.0
    return ProblemSpec(family, d, params, np.full(d, -h), np.full(d, h), tol, f_star=f_star,
                       name=f"served#{k} {family} d={d}")


def arrivals(lam: float, n: int, seed: int = 40) -> np.ndarray:
    """Poisson arrival times (s) of n requests at rate lam."""
    return np.cumsum(np.random.default_rng(seed).exponential(1.0 / lam, n))


def rastrigin(d: int) -> ProblemSpec:
    """Reference Rastrigin on its box [-5.5, 6] (7-Rastrigin/rastrigin50.py:852)."""
    p = np.concatenate([np.ones(d), np.full(d, 10.0), np.full(d, 2.0 * math.pi)])
    return ProblemSpec("sepcos", d, p, np.full(d, -5.5), np.full(d, 6.0), 1e-6, f_star=0.0,
                       name=f"rastrigin d={d}")


def rosenbrock(d: int) -> ProblemSpec:
    """Reference Rosenbrock on [-30, 30] (11-Rosenbrock/rosenbrock50.py:1059)."""
    p = np.concatenate([np.zeros(d), np.zeros(d), np.ones(d), np.full(d - 1, 100.0)])
    return ProblemSpec("sepcos_chain", d, p, np.full(d, -30.0), np.full(d, 30.0), 1e-6, f_star=0.0,
                       name=f"rosenbrock d={d}")


def ackley(d: int) -> ProblemSpec:
    """Reference Ackley on [-35, 40] (1-Ackley/ackley50.py:958)."""
    p = np.concatenate([[20.0, 0.02, 2.0 * math.pi], np.zeros(d)])
    return ProblemSpec("ackley_q", d, p, np.full(d, -35.0), np.full(d, 40.0), 1e-6, f_star=0.0,
                       name=f"ackley d={d}")
