"""Served form: a request loop with continuous batching on one GPU.

Requests (optimisation problems of any family, dimension and tolerance)
arrive over time.  Each dimension class has a Batch of device-resident
slots (queues, parent records and incumbents stay in HBM) with growth tiers
(a problem that outgrows its slot moves into a larger one and resumes); the
slot count of each class is the most whose batch, tiers included, fits the
device memory it is given.  A request is
admitted into a free slot of the smallest class that holds its dimension
as soon as one is free, so its work overlaps on the GPU with whatever else
is running; every batch advances on its own stream, and a loop turn never
waits for a batch: it collects the launches whose results have landed and
launches again where fewer than two launches are in flight
(gpugo_batch_advance).  Latency is measured from arrival to the host seeing
the result.

A slot runs the per-problem loop of gpugo_solve, so each result is the
same alone and in any mix (bit for bit, for the same schedule).

The reference has no served form (one problem per script run,
GPUGO_1.1/1-Ackley/ackley50.py:950-1395).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib, problems as pb
from .batch import Batch, BatchProblem, get_batch_module
from .schedule import fits, schedule_for

DMAX_CLASSES = (16, 32, 64, 100)
# steps after which a request leaves its class's tier 0 for tier 1: the
# long-running requests then no longer lengthen tier 0's launches, which
# every new request waits on
PROMOTE_AGE = 64
# steps per launch of the served batches: a request waits for the launch in
# flight and then its own, so latency grows with the launch length.
# Measured on MI355X with 3 s of Poisson arrivals (same results at every
# setting): 1 / 2 / 4 / 8 steps gave p50 2.2 / 4.5 / 8.5 / 16.6 ms at 100
# requests/s, and 10.4 / 17.3 / 34.1 ms at 1000 requests/s, where 1 step
# took 2% longer to serve the whole trace.  (The batched form keeps its
# tuned launch shape: it is measured on throughput.)
SERVE_STEPS_PER_LAUNCH = 1


@dataclass
class Request:
    rid: int
    spec: "pb.ProblemSpec"
    arrival: float
    admitted: float = -1.0
    done: float = -1.0
    result: Optional[dict] = None


class Server:
    """Continuous-batching solver for a stream of problems."""

    def __init__(self, slots: Optional[Dict[int, int]] = None, *, dmax_classes: Sequence[int] = DMAX_CLASSES,
                 mem_fraction: float = 0.5, max_queue: int = 1 << 18, max_parents: int = 8192,
                 max_iters: int = 2000, steps_per_launch: Optional[int] = SERVE_STEPS_PER_LAUNCH,
                 rounding: str = "fast", grow: int = 2, promote_age: int = PROMOTE_AGE):
        _lib.init()
        self.classes = sorted(dmax_classes)
        self.max_queue = max_queue
        self.max_iters = max_iters
        self.batches: Dict[int, Batch] = {}
        mods = {dm: get_batch_module(pb.FAMILIES, dm, (rounding,)) for dm in self.classes}
        if slots is None:
            # admission by device memory: split the given fraction of free
            # HBM evenly over the classes
            free = ctypes_mem_free()
            share = free * mem_fraction / len(self.classes)
            slots = {dm: Batch.max_slots(mods[dm], int(share), max_queue=max_queue, max_parents=max_parents,
                                         grow=grow) for dm in self.classes}
        self.slots = dict(slots)
        for dm in self.classes:
            self.batches[dm] = Batch(slots=self.slots[dm], module=mods[dm], rounding=rounding,
                                     steps_per_launch=steps_per_launch, max_queue=max_queue,
                                     max_parents=max_parents, grow=grow, promote_age=promote_age)
        self.waiting: List[Request] = []
        self.inflight: Dict[int, Request] = {}

    def dmax_for(self, d: int) -> int:
        for dm in self.classes:
            if d <= dm:
                return dm
        raise ValueError(f"d={d} exceeds the largest class {self.classes[-1]}")

    def problem(self, req: Request) -> BatchProblem:
        s = req.spec
        grid, k = schedule_for(s.family, s.d)
        if not fits((grid[0], grid[1], k), s.d, self.max_queue):
            k = max(1, self.max_queue // grid[0] ** grid[1])
        return BatchProblem.from_spec(s, grid=grid, k_select=k, max_iters=self.max_iters, tag=req.rid)

    def _admit_waiting(self, now: float):
        still = []
        for req in self.waiting:
            b = self.batches[self.dmax_for(req.spec.d)]
            if b.counts()[1] > 0:
                b.admit(self.problem(req))
                req.admitted = now
                self.inflight[req.rid] = req
            else:
                still.append(req)
        self.waiting = still

    def _collect(self, now: float, out: List[Request]):
        for b in self.batches.values():
            for tag, r in b.results():
                req = self.inflight.pop(tag)
                req.done = now
                req.result = r
                out.append(req)

    def run_trace(self, specs: Sequence["pb.ProblemSpec"], arrivals: Sequence[float],
                  time_limit: float = 120.0) -> List[Request]:
        """Serve requests arriving at the given times (seconds from start);
        returns the completed requests (arrival, admitted, done, result)."""
        reqs = [Request(i, s, float(t)) for i, s, t in zip(range(len(specs)), specs, arrivals)]
        order = sorted(reqs, key=lambda r: r.arrival)
        done: List[Request] = []
        nxt = 0
        t0 = time.perf_counter()
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            if now > time_limit:
                break
            while nxt < len(order) and order[nxt].arrival <= now:
                self.waiting.append(order[nxt])
                nxt += 1
            if self.waiting:
                self._admit_waiting(now)
            # every batch advances on its own: one with long steps in flight
            # does not hold up the others (gpugo_batch_advance never waits)
            busy = False
            for b in self.batches.values():
                if b.counts()[0] > 0:
                    b.advance()
                    busy = True
            self._collect(time.perf_counter() - t0, done)
            if not busy and not self.waiting and nxt < len(order):
                # idle until the next arrival
                delay = order[nxt].arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(min(delay, 0.01))
            elif not busy and self.waiting:
                time.sleep(0.0005)
        for b in self.batches.values():
            b.sync()
        self._collect(time.perf_counter() - t0, done)
        return done

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches = {}

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def ctypes_mem_free() -> int:
    import ctypes
    f, t = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.load().gpugo_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def latency_summary(done: List[Request]) -> dict:
    lat = np.array([r.done - r.arrival for r in done]) if done else np.zeros(0)
    widths = [r.result["width"] for r in done if r.result["status_name"] == "certified"]
    ok = [r for r in done if r.result["status_name"] == "certified"]
    tol_ok = all(r.result["width"] <= r.spec.tol for r in ok)
    return {
        "completed": len(done),
        "certified": len(ok),
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "max_ms": float(lat.max() * 1e3) if len(lat) else None,
        "width_max": float(max(widths)) if widths else None,
        "width_median": float(np.median(widths)) if widths else None,
        "all_within_tol": tol_ok,
    }
