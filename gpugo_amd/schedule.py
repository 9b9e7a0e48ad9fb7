"""Split schedules (subregion grid and boxes per step) and their autotuner.

A schedule is (wid, cycle, k_select): each step selects the k_select boxes
with the smallest lower bounds and splits each into wid^cycle children
(cycle coordinates, wid parts each; the window of split coordinates
advances by cycle every step).  The reference fixes one schedule per
script (4^10 children of one box, GPUGO_1.1/1-Ackley/ackley50.py:1006-1015);
here the batch and served forms pick one per problem class from a table
that `tune` measures on the GPU: for each class (family, dimension range)
every candidate schedule solves a sample of that class's problems in one
batch, and the candidate that certifies the most problems, then in the
least wall time, wins.

Odd wid keeps a box centred on the centre of the initial box at every
level (with wid = 2 the centre lies on a split plane of 2^d boxes).

Two launch knobs are tuned too: for batches, the split+bound grid (bound
workgroups) and the steps captured per graph launch, on the whole sample
mix at once (`tune_launch`, criterion certified problems per second); for
single solves, the lanes that share one child (`tune_lanes`: 1 = one lane
per child, G > 1 = the term-parallel kernels with G lanes per child).
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(_HERE, "schedules.json")

# dimension classes (inclusive upper ends)
D_EDGES = (4, 8, 16, 32, 64, 128)
DEFAULT = (3, 4, 32)
CANDIDATES = [(3, 4, 32), (3, 6, 8), (3, 3, 64), (3, 5, 16), (2, 6, 32), (5, 3, 16)]
# batch launch shapes: (bound workgroups, steps per launch); 0 workgroups =
# 8 per CU
DEFAULT_LAUNCH = (0, 4)
LAUNCH_CANDIDATES = [(0, 4), (1024, 4), (4096, 4), (0, 2), (0, 8), (4096, 8)]
LANE_CANDIDATES = (1, 4, 8)

_table = None


def d_class(d: int) -> int:
    for e in D_EDGES:
        if d <= e:
            return e
    return D_EDGES[-1]


def load_table(path: str = TABLE_PATH) -> Dict[str, list]:
    global _table, _launch
    if _table is None:
        data = {}
        if os.path.exists(path):
            with open(path) as fh:
                data = json.load(fh)
        _table = data.get("schedules", {})
        _launch = data.get("launch")
    return _table


_launch = None


def launch_for() -> Tuple[int, int]:
    """(bound workgroups, steps per launch) of the batch form (tuned, else
    the default)."""
    load_table()
    bb, spl = _launch if _launch else DEFAULT_LAUNCH
    return int(bb), int(spl)


def schedule_for(family: str, d: int) -> Tuple[Tuple[int, int], int]:
    """((wid, cycle), k_select) for a problem of this family and dimension;
    cycle is capped at d."""
    wid, cyc, k = load_table().get(f"{family}/{d_class(d)}", DEFAULT)
    return (int(wid), int(min(cyc, d))), int(k)


def fits(sched: Sequence[int], d: int, max_queue: int) -> bool:
    wid, cyc, k = sched
    return k * wid ** min(cyc, d) <= max_queue


def tune(classes: Dict[str, List], candidates: Iterable[Sequence[int]] = CANDIDATES, *, module=None,
         max_queue: int = 1 << 18, max_parents: int = 8192, max_iters: int = 2000,
         time_limit: float = 4.0, log=print) -> Dict[str, list]:
    """Measure every candidate schedule on each class's sample problems
    (ProblemSpec lists keyed by "family/dclass") and return the winners."""
    from .batch import Batch, BatchProblem
    out = {}
    # warm-up (first launches, graph instantiation) outside the measurements
    first = next(iter(classes.values()))[:4]
    with Batch(slots=len(first), module=module, max_queue=max_queue, max_parents=max_parents) as b:
        b.solve_all([BatchProblem.from_spec(s, grid=(3, min(2, s.d)), k_select=4, max_iters=20, tag=i)
                     for i, s in enumerate(first)])
    for key, specs in classes.items():
        best = None
        for cand in candidates:
            if not all(fits(cand, s.d, max_queue) for s in specs):
                continue
            wid, cyc, k = cand
            probs = [BatchProblem.from_spec(s, grid=(wid, min(cyc, s.d)), k_select=k, max_iters=max_iters, tag=i)
                     for i, s in enumerate(specs)]
            with Batch(slots=len(probs), module=module, max_queue=max_queue, max_parents=max_parents) as b:
                t0 = time.perf_counter()
                for p in probs:
                    b.admit(p)
                res = {}
                while True:
                    b.run(1)
                    for tag, r in b.results():
                        res[tag] = r
                    if b.counts()[0] == 0 or time.perf_counter() - t0 > time_limit:
                        break
                b.sync()
                for tag, r in b.results():
                    res[tag] = r
                el = time.perf_counter() - t0
            cert = sum(r["status_name"] == "certified" for r in res.values())
            score = (cert, -el)
            log(f"tune {key} {tuple(cand)}: certified {cert}/{len(probs)} in {el * 1e3:.1f} ms")
            if best is None or score > best[0]:
                best = (score, list(cand))
        if best is not None:
            out[key] = best[1]
    return out


def _run_mix(probs, slots: int, module, time_limit: float, **batch_kw) -> Tuple[dict, float]:
    """Admit the problems into one batch (as slots free up) and run it to
    the end or the time limit; (results by tag, wall seconds)."""
    from .batch import Batch
    res = {}
    with Batch(slots=slots, module=module, **batch_kw) as b:
        pending = list(probs)
        t0 = time.perf_counter()
        while True:
            while pending and b.counts()[0] < slots:
                b.admit(pending.pop(0))
            b.run(1)
            for tag, r in b.results():
                res[tag] = r
            if (not pending and b.counts()[0] == 0) or time.perf_counter() - t0 > time_limit:
                break
        b.sync()
        for tag, r in b.results():
            res[tag] = r
        el = time.perf_counter() - t0
    return res, el


def tune_launch(specs: List, candidates: Iterable[Sequence[int]] = LAUNCH_CANDIDATES, *, module=None,
                slots: int = 256, max_queue: int = 1 << 18, max_parents: int = 8192, max_iters: int = 2000,
                time_limit: float = 8.0, log=print) -> Tuple[list, list]:
    """Measure each (bound workgroups, steps per launch) on the whole mix of
    sample problems (each on its class's schedule) and return the one with
    the most certified problems per second, and the measurements."""
    from .batch import BatchProblem
    probs = []
    for i, s in enumerate(specs):
        grid, k = schedule_for(s.family, s.d)
        probs.append(BatchProblem.from_spec(s, grid=grid, k_select=k, max_iters=max_iters, tag=i))
    _run_mix(probs[:8], min(slots, 8), module, time_limit, max_queue=max_queue, max_parents=max_parents)  # warm-up
    best, rows = None, []
    for cand in candidates:
        bb, spl = (int(v) for v in cand)
        res, el = _run_mix(probs, slots, module, time_limit, max_queue=max_queue, max_parents=max_parents,
                           bound_blocks=bb, steps_per_launch=spl)
        cert = sum(r["status_name"] == "certified" for r in res.values())
        rate = cert / el if el > 0 else 0.0
        rows.append({"bound_blocks": bb, "steps_per_launch": spl, "certified": cert, "n": len(probs),
                     "seconds": el, "certified_per_s": rate})
        log(f"tune launch {(bb, spl)}: certified {cert}/{len(probs)} in {el:.3f} s ({rate:.1f}/s)")
        if best is None or rate > best[0]:
            best = (rate, [bb, spl])
    return best[1], rows


def tune_lanes(obj, lo, hi, lanes: Iterable[int] = LANE_CANDIDATES, *, roundings=("fast",), repeats: int = 3,
               log=print, **solve_kw) -> Tuple[int, list]:
    """Solve one problem with each lane count (1: one lane per child, else
    the term-parallel kernels with that many lanes per child; the objective
    needs loop templates for those) and return the fastest lane count that
    certifies (median of `repeats` timed solves after one untimed), and the
    measurements."""
    import statistics
    from .solver import get_module, minimize
    rows, best = [], None
    for g in lanes:
        g = int(g)
        split = g if g > 1 else 0
        mod = get_module(obj, roundings=roundings, split=split)
        kw = dict(solve_kw, module=mod, split=split, rounding=roundings[0])
        r = minimize(obj, lo, hi, **kw)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            r = minimize(obj, lo, hi, **kw)
            ts.append(time.perf_counter() - t0)
        med = statistics.median(ts)
        ok = r["status_name"] == "certified"
        rows.append({"lanes": g, "certified": ok, "seconds": med, "lb": r["lb"], "ub": r["ub"]})
        log(f"tune lanes {g}: {r['status_name']} in {med * 1e3:.2f} ms")
        if ok and (best is None or med < best[0]):
            best = (med, g)
    return (best[1] if best else None), rows


def save_table(table: Dict[str, list], meta: dict, path: str = TABLE_PATH, launch: Optional[list] = None):
    global _table
    data = {"schedules": table, "meta": meta}
    if launch is not None:
        data["launch"] = list(launch)
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    _table = None


def main():
    """Tune the table on problems of the batched workload (needs a GPU)."""
    import argparse
    from collections import defaultdict
    from . import problems as pb
    from ._lib import init
    from .batch import get_batch_module
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-class", type=int, default=16)
    ap.add_argument("--pool", type=int, default=4096)
    ap.add_argument("--dmax", type=int, default=32)
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--launch-sample", type=int, default=256, help="problems of the launch-shape measurement")
    a = ap.parse_args()
    init(0)
    classes = defaultdict(list)
    # sample problems drawn the way the batched workload draws them, but
    # from a separate seed range so tuning does not see the measured set
    k = 0
    while k < a.pool * 4 and (len(classes) == 0 or min(len(v) for v in classes.values()) < a.per_class):
        s = pb.batch_problem(10 ** 6 + k)
        key = f"{s.family}/{d_class(s.d)}"
        if s.d <= a.dmax and len(classes[key]) < a.per_class:
            classes[key].append(s)
        k += 1
    m = get_batch_module(pb.FAMILIES, a.dmax, ("fast",))
    table = tune(classes, module=m)
    save_table(table, {"per_class": a.per_class, "dmax": a.dmax, "candidates": CANDIDATES,
                       "criterion": "most certified, then least wall time, one batch per class"}, a.out)
    # launch shape on a mix drawn like the workload (table schedules in use)
    mix = [s for s in (pb.batch_problem(2 * 10 ** 6 + i) for i in range(4 * a.launch_sample))
           if s.d <= a.dmax][:a.launch_sample]
    launch, rows = tune_launch(mix, module=m)
    save_table(table, {"per_class": a.per_class, "dmax": a.dmax, "candidates": CANDIDATES,
                       "criterion": "most certified, then least wall time, one batch per class",
                       "launch_candidates": LAUNCH_CANDIDATES, "launch_rows": rows,
                       "launch_criterion": "certified problems per second on the whole mix"}, a.out, launch)
    print(json.dumps({"schedules": table, "launch": launch}, indent=1))


if __name__ == "__main__":
    main()
