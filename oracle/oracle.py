"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / host-core baseline.  The product
package (gpugo_amd) never imports it.

The C sources (gpugo_oracle.c) restate the reference algorithm; see the
header comment there for the reference file:line each part follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgpugo_oracle.so")
_lib = None


class GpugoResult(ctypes.Structure):
    """Mirror of gpugo_result (include/gpugo.h)."""
    _fields_ = [
        ("lb", ctypes.c_double), ("ub", ctypes.c_double), ("width", ctypes.c_double),
        ("status", ctypes.c_int32), ("overflow", ctypes.c_int32),
        ("iterations", ctypes.c_int64), ("boxes_bounded", ctypes.c_int64),
        ("pruned_value", ctypes.c_int64), ("pruned_kkt", ctypes.c_int64),
        ("queue_size", ctypes.c_int64), ("queue_peak", ctypes.c_int64),
        ("n_out", ctypes.c_int64), ("ms", ctypes.c_double),
        ("bound_ms", ctypes.c_double), ("bound_launches", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class OracleProblem(ctypes.Structure):
    _fields_ = [
        ("d", ctypes.c_int32), ("family", ctypes.c_int32), ("wid", ctypes.c_int32),
        ("cycle", ctypes.c_int32), ("k_select", ctypes.c_int32), ("n_samples", ctypes.c_int32),
        ("threads", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("lo", ctypes.POINTER(ctypes.c_double)), ("hi", ctypes.POINTER(ctypes.c_double)),
        ("params", ctypes.POINTER(ctypes.c_double)),
        ("tol_f", ctypes.c_double), ("tol_x", ctypes.c_double), ("init_ub", ctypes.c_double),
        ("max_iters", ctypes.c_int64),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (host clang, strict FP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.POINTER(ctypes.c_double)
        u32 = ctypes.POINTER(ctypes.c_uint32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        _lib.oracle_ia_eval.argtypes = [ctypes.c_int, ctypes.c_int64, d, d, d, d, ctypes.c_double, d, d]
        _lib.oracle_decode.argtypes = [ctypes.c_int, d, d, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int64, u32, d, d]
        _lib.oracle_eval_boxes.argtypes = [ctypes.c_int, ctypes.c_int, d, ctypes.c_int64, d, d, d, d,
                                           ctypes.c_int, d, d, d, d, i32, d]
        _lib.oracle_mv_lower.argtypes = [ctypes.c_int, ctypes.c_int, d, ctypes.c_int64, d, d, d]
        _lib.oracle_solve.argtypes = [ctypes.POINTER(OracleProblem), ctypes.POINTER(GpugoResult),
                                      d, d, d, ctypes.c_int64]
        _lib.oracle_split_bound.restype = ctypes.c_int64
        _lib.oracle_split_bound.argtypes = [ctypes.c_int, ctypes.c_int, d, ctypes.c_int64, d, d, i32,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d,
                                            ctypes.c_int64, u32, u32, d]
    return _lib


def _p(a, ctype=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def rounding_selftest() -> bool:
    return lib().oracle_rounding_selftest() == 0


def ia_eval(op: int, alo, ahi, blo=None, bhi=None, s: float = 0.0):
    alo, ahi = _f64(alo), _f64(ahi)
    n = alo.size
    if blo is not None:
        blo, bhi = _f64(blo), _f64(bhi)
    olo = np.empty(n); ohi = np.empty(n)
    rc = lib().oracle_ia_eval(op, n, _p(alo), _p(ahi), _p(blo), _p(bhi), float(s), _p(olo), _p(ohi))
    if rc:
        raise ValueError(f"oracle_ia_eval: bad op {op}")
    return olo, ohi


def decode(plo, phi, win: int, wid: int, cycle: int, pos):
    plo, phi = _f64(plo), _f64(phi)
    d = plo.size
    c = min(cycle, d)
    pos = np.ascontiguousarray(pos, dtype=np.uint32)
    clo = np.empty((pos.size, d)); chi = np.empty((pos.size, d))
    lib().oracle_decode(d, _p(plo), _p(phi), win, wid, c, pos.size, _p(pos, ctypes.c_uint32),
                        _p(clo), _p(chi))
    return clo, chi


def eval_boxes(family: int, params, lo, hi, dom_lo, dom_hi, n_samples: int = 1):
    """Value / gradient / KKT / sample-ub of the family code lists on boxes."""
    lo, hi = _f64(lo), _f64(hi)
    n, d = lo.shape
    params, dom_lo, dom_hi = _f64(params), _f64(dom_lo), _f64(dom_hi)
    flo = np.empty(n); fhi = np.empty(n)
    glo = np.empty((n, d)); ghi = np.empty((n, d))
    kkt = np.empty(n, dtype=np.int32); ub = np.empty(n)
    rc = lib().oracle_eval_boxes(family, d, _p(params), n, _p(lo), _p(hi), _p(dom_lo), _p(dom_hi),
                                 n_samples, _p(flo), _p(fhi), _p(glo), _p(ghi),
                                 _p(kkt, ctypes.c_int32), _p(ub))
    if rc:
        raise ValueError("oracle_eval_boxes: bad family/d")
    return dict(flo=flo, fhi=fhi, glo=glo, ghi=ghi, kkt=kkt, ub=ub)


def mv_lower(family: int, params, lo, hi):
    """Mean-value lower bound f(c) + sum_j G_j (X_j - c_j) of each box (c the
    midpoint, G the gradient enclosure over the box)."""
    lo, hi = _f64(lo), _f64(hi)
    n, d = lo.shape
    params = _f64(params)
    out = np.empty(n)
    rc = lib().oracle_mv_lower(family, d, _p(params), n, _p(lo), _p(hi), _p(out))
    if rc:
        raise ValueError("oracle_mv_lower: bad family/d")
    return out


def split_bound(family, params, plo, phi, pwin, wid, cycle, n_samples, dom_lo, dom_hi, ub,
                out_max=1 << 22):
    """One sample + rule-out step from explicit parents; returns (ub, parent, pos, lb)."""
    plo, phi = np.atleast_2d(_f64(plo)), np.atleast_2d(_f64(phi))
    npar, d = plo.shape
    pwin = np.ascontiguousarray(pwin, dtype=np.int32)
    params, dom_lo, dom_hi = _f64(params), _f64(dom_lo), _f64(dom_hi)
    ub_io = np.array([ub], dtype=np.float64)
    op = np.empty(out_max, dtype=np.uint32); opos = np.empty(out_max, dtype=np.uint32)
    olb = np.empty(out_max)
    ns = lib().oracle_split_bound(family, d, _p(params), npar, _p(plo), _p(phi),
                                  _p(pwin, ctypes.c_int32), wid, cycle, n_samples,
                                  _p(dom_lo), _p(dom_hi), _p(ub_io), out_max,
                                  _p(op, ctypes.c_uint32), _p(opos, ctypes.c_uint32), _p(olb))
    if ns < 0:
        raise ValueError("out_max too small")
    return float(ub_io[0]), op[:ns].copy(), opos[:ns].copy(), olb[:ns].copy()


def solve(family: int, params, lo, hi, *, tol_f=1e-6, tol_x=0.0, wid=4, cycle=10, k_select=1,
          n_samples=1, init_ub=float("inf"), max_iters=1 << 40, threads=0, out_max=4096,
          mean_value=False):
    """Run the oracle branch-and-bound; returns (result dict, output boxes)."""
    lo, hi, params = _f64(lo), _f64(hi), _f64(params)
    d = lo.size
    pr = OracleProblem(d=d, family=family, wid=wid, cycle=cycle, k_select=k_select,
                       n_samples=n_samples, threads=threads, flags=1 if mean_value else 0,
                       lo=_p(lo), hi=_p(hi), params=_p(params), tol_f=tol_f, tol_x=tol_x,
                       init_ub=init_ub, max_iters=max_iters)
    res = GpugoResult()
    olo = np.empty((out_max, d)); ohi = np.empty((out_max, d)); olb = np.empty(out_max)
    rc = lib().oracle_solve(ctypes.byref(pr), ctypes.byref(res), _p(olo), _p(ohi), _p(olb), out_max)
    if rc:
        raise RuntimeError(f"oracle_solve failed: {rc}")
    r = res.as_dict()
    n = min(res.n_out, out_max)
    return r, (olo[:n], ohi[:n], olb[:n])


def digits15(x: float, up: bool) -> float:
    """outRound of the IA behind the paper's results, n = 15
    (GPUGO_1.1/1-Ackley/IAAckley.py:106-141, ru/rd :56-104): with
    k = 14 - floor(log10 |x|), the bound away from zero is
    ceil(|x| * 10^k) / 10^k and the bound towards zero floor(|x| * 10^k) / 10^k,
    in IEEE double arithmetic (Python floats) as the reference computes it;
    10^k is a double power here (the reference's integer power overflows
    int64 for k > 18); magnitudes below 1e-292 widen to 0 (towards zero) or
    to -/+1e-292 (away from zero), the guard of 10-Zabinsky/IASinusoidal.py:
    109-175.  Like the reference it can land an ulp inside x where |x| * 10^k
    rounds onto an integer; the GPU mode keeps its bound no narrower than
    the directed-rounding bound."""
    import math
    if x != x or x == 0.0 or math.isinf(x):
        return x
    ax = abs(x)
    if ax < 1e-292:
        return (1e-292 if up else -1e-292) if (x > 0) == up else 0.0
    sc = 10.0 ** (14 - math.floor(math.log10(ax)))
    q = ax * sc
    m = (math.ceil(q) if (x > 0) == up else math.floor(q)) / sc
    return m if x > 0 else -m
