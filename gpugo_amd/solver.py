"""Host API: objective modules, problems and the GPUGO-shaped entry point.

    res = minimize(objective, lo, hi, tol=1e-6, grid=(wid, cycle),
                   n_samples=1, init_ub=inf, params=..., k_select=1)

keeps the knobs of a reference driver script (GPUGO_1.1/1-Ackley/
ackley50.py:958-1015): the box (initialRegion), the subregion grid
(wid, cycleSize), numSamples, the initial upper bound (lowestUpperVal) and
the tolerance.  ``tol`` certifies the width of the value enclosure
(ub - lb <= tol); ``tol_x`` enables the reference's x-width output rule
(:1306-1328).  The returned dict holds the enclosure [lb, ub] and counters
(iterations, boxes bounded, pruned by value / by the monotonicity test).
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import GpugoError, check, f64, ptr
from .emit import module_source, source_key, window_module_source
from .expr import Objective

_CACHE_DIR = os.environ.get("GPUGO_CACHE_DIR",
                            os.path.join(os.path.dirname(os.path.abspath(__file__)), "_rtc_cache"))
_modules = {}
_mod_lock = threading.Lock()


def _source(obj: Objective, roundings, f32: bool, split: int, windowed: bool) -> str:
    if windowed:
        if split:
            raise ValueError("windowed modules have no term-parallel kernels")
        return window_module_source(obj, tuple(roundings), f32)
    return module_source(obj, tuple(roundings), f32, split)


def rtc_compile_to_cache(obj: Objective, options: str = "", roundings=("fast",), f32: bool = False,
                         split: int = 0, windowed: bool = False) -> str:
    """Compile an objective module with hipRTC (no GPU needed) into the cache.
    windowed: the large-d form (loop templates only; a child costs the
    terms its split window touches)."""
    return compile_source_to_cache(_source(obj, roundings, f32, split, windowed), obj.name, options)


def compile_source_to_cache(src: str, name: str, options: str = "") -> str:
    """hipRTC-compile generated module text into the on-disk code-object cache."""
    lib = _lib.load()
    key = source_key(src, options)
    path = os.path.join(_CACHE_DIR, key + ".co")
    if os.path.exists(path):
        return path
    size = ctypes.c_size_t(0)
    check(lib.gpugo_rtc_compile(src.encode(), (name + ".hip").encode(), options.encode(),
                                ctypes.byref(size)))
    buf = ctypes.create_string_buffer(size.value)
    check(lib.gpugo_rtc_code_copy(buf, size.value))
    os.makedirs(_CACHE_DIR, exist_ok=True)
    tmp = path + f".{os.getpid()}.tmp"
    with open(tmp, "wb") as fh:
        fh.write(buf.raw[: size.value])
    os.replace(tmp, path)
    return path


class Module:
    """A compiled objective (one hipRTC code object loaded on the GPU)."""

    def __init__(self, obj: Objective, options: str = "", roundings=("fast",), f32: bool = False,
                 split: int = 0, windowed: bool = False):
        _lib.init()
        lib = _lib.load()
        self.obj = obj
        self.roundings = tuple(roundings)
        self.f32 = bool(f32)
        self.split = int(split)
        self.windowed = bool(windowed)
        path = rtc_compile_to_cache(obj, options, self.roundings, self.f32, self.split, self.windowed)
        with open(path, "rb") as fh:
            code = fh.read()
        h = ctypes.c_void_p()
        check(lib.gpugo_module_load(code, len(code), ctypes.byref(h)))
        self.handle = h
        self.code_path = path

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.gpugo_module_destroy(h)
            self.handle = None

    def eval_boxes(self, params, lo, hi, dom_lo=None, dom_hi=None, n_samples: int = 1,
                   rounding: str = "exact", want=("f", "g", "kkt", "ub")):
        lo, hi = np.atleast_2d(f64(lo)), np.atleast_2d(f64(hi))
        n, d = lo.shape
        params = f64(params)
        dom_lo = f64(dom_lo if dom_lo is not None else np.full(d, -np.inf))
        dom_hi = f64(dom_hi if dom_hi is not None else np.full(d, np.inf))
        out = {}
        flo = fhi = glo = ghi = kkt = ub = None
        if "f" in want:
            flo, fhi = np.empty(n), np.empty(n)
        if "g" in want:
            glo, ghi = np.empty((n, d)), np.empty((n, d))
        if "kkt" in want:
            kkt = np.empty(n, dtype=np.int32)
        if "ub" in want:
            ub = np.empty(n)
        check(_lib.load().gpugo_eval_boxes(self.handle, d, ptr(params), params.size, n, ptr(lo), ptr(hi),
                                           ptr(dom_lo), ptr(dom_hi), n_samples, _lib.ROUND[rounding],
                                           ptr(flo), ptr(fhi), ptr(glo), ptr(ghi),
                                           ptr(kkt, ctypes.c_int32), ptr(ub)))
        out.update(flo=flo, fhi=fhi, glo=glo, ghi=ghi, kkt=kkt, ub=ub)
        return {k: v for k, v in out.items() if v is not None}


def eval_boxes_f32(module: "Module", params, lo, hi, rounding: str = "fast"):
    """FP32 value enclosures of explicit boxes (module built with f32=True)."""
    lo, hi = np.atleast_2d(f64(lo)), np.atleast_2d(f64(hi))
    n, d = lo.shape
    params = f64(params)
    flo, fhi = np.empty(n), np.empty(n)
    check(_lib.load().gpugo_eval_boxes_f32(module.handle, d, ptr(params), params.size, n, ptr(lo), ptr(hi),
                                           _lib.ROUND[rounding], ptr(flo), ptr(fhi)))
    return flo, fhi


def get_module(obj: Objective, options: str = "", roundings=("fast",), f32: bool = False,
               split: int = 0, windowed: bool = False) -> Module:
    roundings = tuple(sorted(set(roundings)))
    key = source_key(_source(obj, roundings, f32, split, windowed), options)
    with _mod_lock:
        m = _modules.get(key)
        if m is None:
            m = Module(obj, options, roundings, f32, split, windowed)
            _modules[key] = m
        return m


def ia_eval(op: str, alo, ahi, blo=None, bhi=None, s: float = 0.0, prec: str = "fp64",
            rounding: str = "exact"):
    """Batched interval op on the GPU (replaces IA.py's device functions)."""
    _lib.init()
    alo, ahi = f64(alo), f64(ahi)
    n = alo.size
    if blo is not None:
        blo, bhi = f64(blo), f64(bhi)
    olo, ohi = np.empty(n), np.empty(n)
    check(_lib.load().gpugo_ia_eval(_lib.OPS[op], _lib.PREC[prec], _lib.ROUND[rounding], n, ptr(alo),
                                    ptr(ahi), ptr(blo), ptr(bhi), float(s), ptr(olo), ptr(ohi)))
    return olo, ohi


def window_ranges(rows) -> np.ndarray:
    """The windowed modules' rule (host copy of the device function): rows
    of (win, c, d, klo, khi, nt) -> (n, lo0, hi0, lo1, hi1), the ranges of
    terms j in [0, nt) whose coordinates j + k, k in [klo, khi], meet the
    window {(win + t) mod d : t < c}.  Needs no GPU."""
    a = np.ascontiguousarray(np.atleast_2d(rows), dtype=np.int32)
    out = np.zeros((a.shape[0], 5), dtype=np.int32)
    check(_lib.load().gpugo_window_ranges(a.shape[0], ptr(a, ctypes.c_int32), ptr(out, ctypes.c_int32)))
    return out


class Problem:
    """One device-resident branch-and-bound solve."""

    def __init__(self, obj: Objective, params, lo, hi, *, tol: float = 1e-6, tol_x: float = 0.0,
                 grid: Sequence[int] = (4, 10), k_select: int = 1, n_samples: int = 1,
                 init_ub: float = np.inf, rounding: str = "fast", max_queue: int = 0,
                 max_parents: int = 0, max_iters: int = 0, module: Optional[Module] = None,
                 mean_value: bool = False, precision: str = "fp64", split: int = 0,
                 persistent: bool = False, windowed: bool = False, eager_compact: bool = False,
                 tables: bool = True, band: bool = True):
        """precision "fp32": children are first bounded in FP32 intervals
        (outward, one ulp per operation) and only those that survive get the
        FP64 value, gradient and sample; the enclosure is certified in FP64.
        split = G > 1: term-parallel kernels, G lanes share each child (for
        steps too small to fill the GPU; needs the objective's loop
        templates).  persistent (with split): the steps of a solve run in
        one cooperative launch with grid barriers between their phases.
        windowed: the large-d form (loop templates only; each template sum
        is kept per parent record outside its split window, so a child
        costs the terms its window touches; any n_samples; FP64, or the
        FP32-prune form with precision "fp32").
        eager_compact (k_select 1): compact the queue after every step
        instead of every few steps with per-chunk minimum keys between
        (same results; a test and measurement switch).
        tables (windowed): look window terms up in the step's term tables
        (one entry per term and digit combination, filled once per split
        record) instead of evaluating them per child; False evaluates them
        per child (same results; a test and measurement switch).
        band (k_select > 1): each compaction also keeps the entries below a
        threshold in a side list that the next select reads instead of the
        whole queue; False passes over the whole queue every step (same
        results; a test and measurement switch)."""
        if precision not in _lib.PREC:
            raise ValueError(f"precision must be one of {sorted(_lib.PREC)}")
        self.module = module or get_module(obj, roundings=(rounding,), f32=precision == "fp32", split=split,
                                           windowed=windowed)
        self.obj = obj
        self._lo, self._hi, self._params = f64(lo), f64(hi), f64(params)
        wid, cycle = grid
        self.desc = _lib.ProblemDesc(
            d=obj.d, wid=int(wid), cycle=int(cycle), k_select=int(k_select), n_samples=int(n_samples),
            precision=_lib.PREC[precision], rounding=_lib.ROUND[rounding],
            flags=((_lib.FLAG_MEAN_VALUE if mean_value else 0) | (_lib.FLAG_SPLIT if split else 0)
                   | (_lib.FLAG_PERSIST if persistent else 0) | (_lib.FLAG_EAGER_COMPACT if eager_compact else 0)
                   | (0 if tables else _lib.FLAG_NO_TABLES) | (0 if band else _lib.FLAG_NO_BAND)),
            lo=ptr(self._lo), hi=ptr(self._hi),
            params=ptr(self._params), n_params=self._params.size, tol_f=float(tol), tol_x=float(tol_x),
            init_ub=float(init_ub), max_queue=int(max_queue), max_parents=int(max_parents),
            max_iters=int(max_iters))
        h = ctypes.c_void_p()
        check(_lib.load().gpugo_problem_create(self.module.handle, ctypes.byref(self.desc), ctypes.byref(h)))
        self.handle = h

    def reset(self):
        """Back to the root box, keeping the device allocations."""
        check(_lib.load().gpugo_problem_reset(self.handle))

    def timing(self, enable: bool = True):
        """Time each split+bound launch with HIP events (result['bound_ms'])."""
        check(_lib.load().gpugo_problem_timing(self.handle, int(bool(enable))))

    def step(self, iters: int = 1) -> dict:
        r = _lib.Result()
        check(_lib.load().gpugo_step(self.handle, int(iters), ctypes.byref(r)))
        return r.as_dict()

    def solve(self) -> dict:
        r = _lib.Result()
        check(_lib.load().gpugo_solve(self.handle, ctypes.byref(r)))
        return r.as_dict()

    def boxes(self, max_boxes: int = 1024):
        d = self.obj.d
        lo, hi, lb = np.empty((max_boxes, d)), np.empty((max_boxes, d)), np.empty(max_boxes)
        n = ctypes.c_int64(0)
        check(_lib.load().gpugo_boxes_copy(self.handle, max_boxes, ptr(lo), ptr(hi), ptr(lb), ctypes.byref(n)))
        m = min(n.value, max_boxes)
        return lo[:m], hi[:m], lb[:m], n.value

    def close(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.gpugo_problem_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def minimize(objective: Objective, lo, hi, tol: float = 1e-6, grid: Sequence[int] = (4, 10),
             n_samples: int = 1, init_ub: float = np.inf, params=None, k_select: int = 1,
             tol_x: float = 0.0, rounding: str = "fast", max_iters: int = 0, return_boxes: int = 0,
             combine_tol: float = 0.0, **kw) -> dict:
    """Enclose the global minimum of `objective` over the box [lo, hi].
    return_boxes > 0 also returns up to that many live (or, with tol_x,
    output) boxes; combine_tol > 0 merges them into roots
    (gpugo_amd.results.roots, the reference's combine/connectivity)."""
    params = np.zeros(0) if params is None else params
    with Problem(objective, params, lo, hi, tol=tol, tol_x=tol_x, grid=grid, k_select=k_select,
                 n_samples=n_samples, init_ub=init_ub, rounding=rounding, max_iters=max_iters, **kw) as p:
        res = p.solve()
        if return_boxes:
            blo, bhi, blb, n = p.boxes(return_boxes)
            res["boxes"] = (blo, bhi, blb)
            res["n_boxes"] = n
            if combine_tol > 0 and len(blo):
                from .results import roots
                rlo, rhi, owner = roots(blo, bhi, combine_tol)
                res["roots"] = (rlo, rhi, owner)
    return res
