"""Batched and continuously batched solves: many independent problems
resident on one GPU (include/gpugo.h, gpugo_batch_*).

A batch module holds objective families in variable-dimension form
(problems.objective_dynamic); a Batch owns device-resident slots, admits
problems into free slots at any time, advances every running slot by
`steps_per_launch` steps per launch (one hipGraph), and reports finished
problems.  Each slot runs the per-problem loop of gpugo_solve, so a
problem's result does not depend on what else is in the batch.

The reference solves one problem per script run
(GPUGO_1.1/1-Ackley/ackley50.py:950-1395); this form has no counterpart
there.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence

import numpy as np

from . import _lib, problems as pb
from ._lib import GpugoError, check, f64, ptr
from .emit import batch_module_source, source_key
from .solver import compile_source_to_cache

_bmods = {}
_bmod_lock = threading.Lock()


class BatchModule:
    """hipRTC batch module: `families` at maximum dimension `dmax`."""

    def __init__(self, families: Sequence[str] = pb.FAMILIES, dmax: int = 32, roundings=("fast",),
                 options: str = "", mean_value: bool = False):
        _lib.init()
        self.families = tuple(families)
        self.dmax = int(dmax)
        self.roundings = tuple(roundings)
        self.mean_value = bool(mean_value)
        objs = [pb.objective_dynamic(f, self.dmax) for f in self.families]
        src = batch_module_source(objs, self.roundings, self.mean_value)
        path = compile_source_to_cache(src, f"batch_d{self.dmax}", options)
        with open(path, "rb") as fh:
            code = fh.read()
        h = ctypes.c_void_p()
        check(_lib.load().gpugo_module_load(code, len(code), ctypes.byref(h)))
        self.handle = h
        self.code_path = path

    def family_index(self, family: str) -> int:
        return self.families.index(family)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.gpugo_module_destroy(h)
            self.handle = None


def batch_source(families: Sequence[str] = pb.FAMILIES, dmax: int = 32, roundings=("fast",),
                 mean_value: bool = False) -> str:
    return batch_module_source([pb.objective_dynamic(f, dmax) for f in families], tuple(roundings), mean_value)


def get_batch_module(families: Sequence[str] = pb.FAMILIES, dmax: int = 32, roundings=("fast",),
                     options: str = "", mean_value: bool = False) -> BatchModule:
    roundings = tuple(sorted(set(roundings)))
    key = (tuple(families), int(dmax), roundings, options, bool(mean_value))
    with _bmod_lock:
        m = _bmods.get(key)
        if m is None:
            m = BatchModule(families, dmax, roundings, options, mean_value)
            _bmods[key] = m
        return m


@dataclass
class BatchProblem:
    """One problem of a batch (coefficients in the family's own d layout)."""
    family: str
    d: int
    params: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    tol: float
    grid: tuple = (2, 10)
    k_select: int = 16
    n_samples: int = 1
    init_ub: float = np.inf
    max_iters: int = 0
    tag: int = 0
    mean_value: bool = False
    extra: dict = field(default_factory=dict)

    @staticmethod
    def from_spec(spec: "pb.ProblemSpec", grid=None, k_select=None, tag: int = 0, **kw) -> "BatchProblem":
        return BatchProblem(spec.family, spec.d, spec.params, spec.lo, spec.hi, spec.tol,
                            grid=tuple(grid or spec.grid), k_select=int(k_select or spec.k_select), tag=tag,
                            **kw)


class Batch:
    """Device-resident slots with continuous admission."""

    def __init__(self, slots: int = 4096, *, families: Sequence[str] = pb.FAMILIES, dmax: int = 32,
                 rounding: str = "fast", steps_per_launch: Optional[int] = None, use_graph: bool = True,
                 max_queue: int = 0, max_parents: int = 0, bound_blocks: Optional[int] = None, grow: int = 2,
                 module: Optional[BatchModule] = None, promote_age: int = 0):
        """grow: larger tiers (0..3, each 8x the per-slot capacity of the one
        below) into which a problem that outgrows its slot moves, state and
        all, and resumes; 0 reports such a problem as overflow.
        steps_per_launch / bound_blocks: None takes the tuned launch shape
        of the schedule table (gpugo_amd.schedule.launch_for).
        promote_age (with grow): a tier-0 problem that has run this many
        steps moves to tier 1 (0: never), so tier 0's launches stay short
        for the problems admitted since; same results."""
        from .schedule import launch_for
        bb0, spl0 = launch_for()
        steps_per_launch = spl0 if steps_per_launch is None else steps_per_launch
        bound_blocks = bb0 if bound_blocks is None else bound_blocks
        self.module = module or get_batch_module(families, dmax, (rounding,))
        self.dmax = self.module.dmax
        self.desc = _lib.BatchDesc(slots=int(slots), steps_per_launch=int(steps_per_launch),
                                   rounding=_lib.ROUND[rounding], use_graph=int(bool(use_graph)),
                                   max_queue=int(max_queue), max_parents=int(max_parents),
                                   bound_blocks=int(bound_blocks), grow=int(grow), promote_age=int(promote_age))
        h = ctypes.c_void_p()
        check(_lib.load().gpugo_batch_create(self.module.handle, ctypes.byref(self.desc), ctypes.byref(h)))
        self.handle = h
        self._keep = {}

    @staticmethod
    def device_bytes(module: BatchModule, slots: int, **kw) -> int:
        desc = _lib.BatchDesc(slots=int(slots), steps_per_launch=int(kw.get("steps_per_launch") or 4),
                              rounding=_lib.ROUND[kw.get("rounding", "fast")], use_graph=1,
                              max_queue=int(kw.get("max_queue", 0)), max_parents=int(kw.get("max_parents", 0)),
                              bound_blocks=0, grow=int(kw.get("grow", 2)))
        n = ctypes.c_int64(0)
        check(_lib.load().gpugo_batch_bytes(module.handle, ctypes.byref(desc), ctypes.byref(n)))
        return n.value

    @staticmethod
    def max_slots(module: BatchModule, budget_bytes: int, limit: int = 4096, **kw) -> int:
        """The most slots (at most `limit`, at least 1) whose batch, growth
        tiers included, fits in budget_bytes of device memory."""
        lo, hi = 1, int(limit)
        if Batch.device_bytes(module, hi, **kw) <= budget_bytes:
            return hi
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if Batch.device_bytes(module, mid, **kw) <= budget_bytes:
                lo = mid
            else:
                hi = mid
        return lo

    def admit(self, prob: BatchProblem, tag: Optional[int] = None) -> int:
        """Admit into a free slot; raises GpugoError (E_CAPACITY) when full."""
        fam = self.module.family_index(prob.family)
        if not 1 <= prob.d <= self.dmax:
            raise GpugoError(_lib.E_INVALID, f"d={prob.d} outside [1, {self.dmax}] of this batch module")
        params = f64(pb.pad_params(prob.family, prob.d, prob.params, self.dmax))
        lo, hi = f64(prob.lo), f64(prob.hi)
        wid, cycle = prob.grid
        desc = _lib.ProblemDesc(d=int(prob.d), wid=int(wid), cycle=int(cycle), k_select=int(prob.k_select),
                                n_samples=int(prob.n_samples), precision=0, rounding=self.desc.rounding,
                                flags=_lib.FLAG_MEAN_VALUE if prob.mean_value else 0,
                                lo=ptr(lo), hi=ptr(hi), params=ptr(params), n_params=params.size,
                                tol_f=float(prob.tol), tol_x=0.0, init_ub=float(prob.init_ub), max_queue=0,
                                max_parents=0, max_iters=int(prob.max_iters))
        slot = ctypes.c_int32(-1)
        t = prob.tag if tag is None else tag
        check(_lib.load().gpugo_batch_admit(self.handle, fam, ctypes.byref(desc), int(t), ctypes.byref(slot)))
        return slot.value

    def run(self, launches: int = 1):
        check(_lib.load().gpugo_batch_run(self.handle, int(launches)))

    def advance(self) -> int:
        """Collect what has landed and launch every tier that has running
        slots and fewer than two launches in flight, without waiting;
        returns the launches made."""
        n = ctypes.c_int64(0)
        check(_lib.load().gpugo_batch_advance(self.handle, ctypes.byref(n)))
        return n.value

    def sync(self):
        check(_lib.load().gpugo_batch_sync(self.handle))

    def counts(self):
        r, f, p = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(_lib.load().gpugo_batch_counts(self.handle, ctypes.byref(r), ctypes.byref(f), ctypes.byref(p)))
        return r.value, f.value, p.value

    def results(self, max_n: int = 1 << 16) -> List[tuple]:
        _, _, pending = self.counts()
        n_max = min(max_n, pending)
        if n_max == 0:
            return []
        tags = np.empty(n_max, dtype=np.int64)
        res = (_lib.Result * n_max)()
        n = ctypes.c_int64(0)
        check(_lib.load().gpugo_batch_results(self.handle, n_max, ptr(tags, ctypes.c_int64), res,
                                              ctypes.byref(n)))
        return [(int(tags[i]), res[i].as_dict()) for i in range(n.value)]

    def solve_all(self, probs: Iterable[BatchProblem], launches_per_poll: int = 2) -> dict:
        """Run every problem to termination, admitting into slots as they
        free up; returns {tag: result}."""
        pending = list(probs)[::-1]
        out = {}
        while True:
            _, free, _ = self.counts()
            while pending and free > 0:
                self.admit(pending.pop())
                free -= 1
            running, _, _ = self.counts()
            if running == 0 and not pending:
                break
            self.run(launches_per_poll)
            for tag, r in self.results():
                out[tag] = r
            running, free, _ = self.counts()
            if running == 0 and pending == [] :
                self.sync()
        self.sync()
        for tag, r in self.results():
            out[tag] = r
        return out

    def close(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.gpugo_batch_destroy(h)
        self.handle = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
