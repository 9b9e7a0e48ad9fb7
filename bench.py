#!/usr/bin/env python3
"""Benchmark: time to a certified enclosure (tol 1e-6) of the d = 50
Ackley-type problem, FP32-prune / FP64-certify form.

Headline workload (ackley50): f(x) = -A exp(-B sqrt(mean x^2)) -
exp(mean cos(C x)) + A + e + sum_i q_i x_i^4 / 50, d = 50, coefficients
drawn from seed 2, box [-32, 32]^50, tol 1e-6 on the width of the value
enclosure, 3^12 children of one box per step, the midpoint as the sample
point; children are ruled out in FP32 intervals (outward, one ulp per
operation) and the enclosure is certified in FP64.  Windowed code lists: a
child differs from its parent in the 12 coordinates of the split window, so
it evaluates those terms of each sum; the other 38 are enclosed once per
parent (the unrolled form, every coordinate of every child, is measured
beside it in c2_forms).

One "step" = one complete solve, from the root box to a certified enclosure
[lb, ub] with ub - lb <= 1e-6, with the objective module loaded and the
problem's device buffers allocated beforehand (inputs resident in HBM); the
timed region covers reset + solve and ends after the solve's final device
synchronisation.  value = mean time to certified enclosure (seconds, lower is
better).  With the midpoint as sample point this problem certifies in one
step (the natural extension over the child that holds x* = 0 has lower bound
0 up to rounding and its midpoint is x*); the c2_s10 leg runs the same
problem with the reference's ten diagonal samples (61 steps).

Also reported: the roofline of the dominant kernel (algorithmic FLOP per
launch over its HIP-event duration, measured in a second pass of the same
solves so the events do not add to value), the reference method on the
host cores (cpu_baseline: the C restatement of the reference loop, one box
per step, natural extension, no mean-value bound), the other
configurations (c0, c1 on its 4^10 grid, c2 in the FP64 form and with ten
samples), the compile-inclusive (cold) time, the large-d windowed form, the
4096-problem batch and the Poisson-served workload.

    python bench.py [--steps K] [--warmup W] [--legs ...] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-certified enclosure (tol 1e-6) + boxes bounded/s; served p50/p99 latency"
# MI355X published vector peaks (TFLOP/s)
PEAK = {"f64": 78.6, "f32": 157.3}
FAMILY_INDEX = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}
TRAFFIC_FILE = "profiles/traffic_ackley50.json"
# the headline schedule: the reference's one box per step, 3^12 children
C2_GRID, C2_K = (3, 12), 1
# seconds of one-thread work per CPU baseline (a bounded sample unless a
# complete solve fits)
ONE_CORE_BUDGET = 10.0
# (precision, windowed code lists): the headline form and the others measured beside it
HEADLINE_FORM = ("fp32", True)
OTHER_FORMS = (("fp64", True), ("fp32", False), ("fp64", False))
# the d = 10 chain problem on its 4^10 grid: the windowed forms, whose window
# covers all ten coordinates (every term from term tables; the first is the
# one the CPU baseline is compared with), and the unrolled form
C1_FORMS = (("fp32", True), ("fp64", True), ("fp64", False))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip every CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=40.0,
                    help="seconds of host work per CPU baseline (complete solves when they fit)")
    ap.add_argument("--legs", default="c2_forms,c2_s10,c1,c1_tuned,c0,cold,large_d,xl_d,batch,served,band",
                    help="comma list besides the headline; 'none' for the headline only")
    ap.add_argument("--batch-n", type=int, default=4096)
    ap.add_argument("--seq-sample", type=int, default=256,
                    help="problems of the batch solved one at a time for the sequential baseline")
    ap.add_argument("--large-d-cases", default="", metavar="NAME:D,...",
                    help="objectives of the large_d leg (default: the eleven reference objectives at d=1000, "
                         "Levy at d=2000)")
    ap.add_argument("--large-d-reps", type=int, default=1, help="timed solves per large_d case (after one untimed)")
    ap.add_argument("--no-headline", action="store_true",
                    help="run only the legs named in --legs (profiling one workload at a time); value is null")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batch leg with stream launches instead of the captured graph (kernel traces)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (scales HBM byte counters)")
    return ap.parse_args()


def copy_probe(nbytes):
    """A plain copy kernel of known size, for scaling FETCH_SIZE/WRITE_SIZE."""
    import torch
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}))


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter profile
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a device copy of
    known size); None if that profile has no row for it."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as fh:
            return json.load(fh)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def host_threads():
    """Host threads for the CPU baselines: the cores this process may run
    on, at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ---------------------------------------------------------------------------
# GPU solves
# ---------------------------------------------------------------------------
def roofline(obj, results, *, precision="fp64", mean_value=False, n_samples=1, kernel="gpugo_bound_fast",
             windowed_c=0, wid=3, tables=True):
    """Algorithmic FLOP of the split+bound launches over their HIP-event
    time (per-op weights add/sub/scalar 2, sqr 2, pown(k) 2 ceil(log2 k),
    interval product 8, division 10, sqrt 2, exp/sin/cos 40, counted by
    gpugo_amd.emit from the code lists): every child gets the value code list; value survivors also
    get the gradient (the full gradient is charged: the monotonicity test
    may stop at its first pruning component) and n_samples sample
    evaluations (with the mean-value bound also the midpoint value and d
    interval products and additions).  In the FP32-prune form the FP32
    value list is charged for every child and the FP64 lists for the value
    survivors only.  frac is therefore an upper bound of the VALU work
    actually done over the peak.  Windowed modules read their window terms
    from the step's term tables: a child is charged the table folds and the
    non-sum part only (the tables are filled once per split record by the
    prep kernel, whose FLOP are reported as flop_table_per_record)."""
    from gpugo_amd import emit
    fl = (emit.window_flop_counts(obj, windowed_c, tables=tables, wid=wid, n_samples=n_samples) if windowed_c
          else emit.flop_counts(obj))
    boxes = sum(x["boxes_bounded"] for x in results)
    pv = sum(x["pruned_value"] for x in results)
    bound_ms = sum(x["bound_ms"] for x in results)
    launches = sum(x["bound_launches"] for x in results)
    extra = n_samples * fl["sample"] + fl["grad"] + (10 * obj.d if mean_value else 0)
    surv = boxes - pv
    if precision == "fp32":
        flops = boxes * fl["value"] + surv * (fl["value"] + extra)
    else:
        flops = boxes * fl["value"] + surv * extra
    dt = "f32" if precision == "fp32" else "f64"
    achieved = flops / (bound_ms * 1e-3) / 1e12 if bound_ms > 0 else 0.0
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": PEAK[dt], "unit": "TFLOP/s",
            "frac": achieved / PEAK[dt], "frac_is_upper_bound": True,
            "flop_per_launch": flops / max(launches, 1), "flop_per_child": flops / max(boxes, 1),
            "children_per_launch": boxes / max(launches, 1),
            "avg_launch_ms": bound_ms / max(launches, 1), "launches": launches,
            **({"term_tables": bool(tables), "flop_table_per_record": fl.get("table", 0)} if windowed_c else {}),
            "timing": "HIP events from each split+bound dispatch, second pass of the same solves"}


def gpu_solves(G, obj, spec, *, steps, warmup, grid, k_select, n_samples=1, precision="fp64",
               mean_value=False, split=0, windowed=False, module=None, init_ub=np.inf, max_iters=100000,
               max_queue=0):
    """`steps` complete solves after `warmup` untimed ones; then the same
    solves again with per-launch events (roofline).  Returns timing and the
    results of both passes."""
    p = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=grid, k_select=k_select,
                  n_samples=n_samples, precision=precision, mean_value=mean_value, split=split,
                  windowed=windowed, module=module, init_ub=init_ub, max_iters=max_iters, max_queue=max_queue)
    for _ in range(warmup):
        p.reset()
        p.solve()
    times, res = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        p.reset()
        t1 = time.perf_counter()
        res.append(p.solve())
        times.append(time.perf_counter() - t1)
    total = time.perf_counter() - t0
    p.timing(True)
    timed = []
    for _ in range(steps):
        p.reset()
        timed.append(p.solve())
    p.timing(False)
    p.close()
    same = all(all(a[k] == b[k] for k in ("lb", "ub", "iterations", "boxes_bounded")) for a, b in zip(res, timed))
    r = res[-1]
    return {"mean_s": total / steps, "median_s": statistics.median(times), "min_s": min(times),
            "certified": all(x["status_name"] == "certified" for x in res + timed),
            "lb": r["lb"], "ub": r["ub"], "width": r["width"], "iterations": r["iterations"],
            "boxes_bounded": r["boxes_bounded"], "pruned_value": r["pruned_value"],
            "queue_peak": r["queue_peak"], "boxes_bounded_per_s": r["boxes_bounded"] / statistics.median(times),
            "passes_identical": same, "_timed": timed}


def public(d):
    return {k: v for k, v in d.items() if not k.startswith("_")}


# ---------------------------------------------------------------------------
# CPU: the reference method on the host cores
# ---------------------------------------------------------------------------
def cpu_reference(spec, grid, k_select, n_samples, budget, gpu_iterations=None, one_core=True):
    """The oracle (C restatement of IA.py and of the reference loop: one box
    per step, wid^cycle children, n_samples diagonal samples, natural
    extension and monotonicity test, no mean-value bound) on the host.
    Complete solves (median of up to 5) when one fits the budget; otherwise
    a bounded number of steps from the root, reported as boxes/s.  The same
    on one host thread (median of up to 3 complete solves, or steps)."""
    from oracle import oracle as O
    O.build()
    fam = FAMILY_INDEX[spec.family]
    threads = host_threads()

    def run(th, max_iters):
        t0 = time.perf_counter()
        r, _ = O.solve(fam, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=grid[0], cycle=grid[1],
                       k_select=k_select, n_samples=n_samples, max_iters=max_iters, threads=th)
        return time.perf_counter() - t0, r

    def measure(th, reps):
        # probe two steps, then decide between complete solves and a sample
        t_probe, r_probe = run(th, 2)
        per_box = t_probe / max(r_probe["boxes_bounded"], 1)
        est = None
        if gpu_iterations is not None and r_probe["status"] == 0:
            est = per_box * r_probe["boxes_bounded"] / max(r_probe["iterations"], 1) * gpu_iterations
        if r_probe["status"] == 1 or (est is not None and est <= budget):
            ts, r = [], None
            t_start = time.perf_counter()
            while len(ts) < reps and (not ts or time.perf_counter() - t_start + ts[-1] <= budget):
                t, r = run(th, 1 << 40)
                ts.append(t)
            return {"value": statistics.median(ts), "unit": "s", "cores": th, "complete": True,
                    "sample": f"{len(ts)} complete solve(s), median", "certified": r["status"] == 1,
                    "lb": r["lb"], "ub": r["ub"], "iterations": r["iterations"],
                    "boxes_bounded": r["boxes_bounded"], "boxes_bounded_per_s": r["boxes_bounded"] / statistics.median(ts)}
        # bounded sample: as many steps from the root as fit the budget
        steps = max(2, int(budget / max(t_probe / 2, 1e-9)))
        if gpu_iterations is not None:
            # the certificate is taken at the select after the last split step
            steps = min(steps, gpu_iterations + 1)
        t, r = run(th, steps)
        rate = r["boxes_bounded"] / t
        if r["status"] == 1:  # the bounded run reached the certificate
            return {"value": t, "unit": "s", "cores": th, "complete": True, "sample": "1 complete solve",
                    "certified": True, "lb": r["lb"], "ub": r["ub"], "iterations": r["iterations"],
                    "boxes_bounded": r["boxes_bounded"], "boxes_bounded_per_s": rate}
        return {"value": None, "unit": "s", "cores": th, "complete": False,
                "sample": f"{r['iterations']} steps from the root, then stopped", "boxes_bounded_per_s": rate,
                "boxes_bounded": r["boxes_bounded"], "seconds": t}

    out = measure(threads, 5)
    out.update({"kind": "port", "cpu_model": cpu_model(),
                "method": f"reference loop: one box per step, {grid[0]}^{grid[1]} children, "
                          f"{n_samples} diagonal sample(s), natural extension + monotonicity test"})
    if one_core:
        budget = min(budget, ONE_CORE_BUDGET)
        out["one_core"] = measure(1, 3)
    return out


def with_ratio(cpu, gpu_s, gpu_boxes):
    """Speed-up of the GPU over the host: measured if the host solve
    completed, else estimated from its boxes/s for the GPU's boxes."""
    if cpu.get("complete"):
        cpu["speedup_gpu_vs_cpu"] = cpu["value"] / gpu_s
    else:
        est = gpu_boxes / cpu["boxes_bounded_per_s"]
        cpu["est_time_for_gpu_boxes_s"] = est
        cpu["speedup_gpu_vs_cpu_est"] = est / gpu_s
    oc = cpu.get("one_core")
    if oc:
        oc["speedup_gpu_vs_cpu"] = (oc["value"] / gpu_s) if oc.get("complete") else \
            (gpu_boxes / oc["boxes_bounded_per_s"]) / gpu_s
    return cpu


# ---------------------------------------------------------------------------
# legs
# ---------------------------------------------------------------------------
def form_solves(G, pb, args, spec, grid, k, form, *, n_samples=1, mean_value=False, split=0, max_iters=100000):
    """Solves of one form: form = (precision, windowed).  Windowed forms use
    the loop-form module (a child costs the terms its split window touches;
    the other terms are summed once per parent record)."""
    prec, windowed = form
    if windowed:
        obj = pb.objective_loops(spec.family, spec.d)
        module = G.get_module(obj, roundings=("fast",), f32=prec == "fp32", windowed=True)
    else:
        obj = pb.objective(spec.family, spec.d)
        module = None
    g = gpu_solves(G, obj, spec, steps=args.steps, warmup=args.warmup, grid=grid, k_select=k,
                   n_samples=n_samples, precision=prec, mean_value=mean_value, split=split, windowed=windowed,
                   module=module, max_iters=max_iters, max_queue=1 << 26)
    kname = "gpugo_bound_fast" + ("_mv" if mean_value else "") + ("_g" if split else "")
    g["roofline"] = roofline(obj, g["_timed"], precision=prec, mean_value=mean_value, n_samples=n_samples,
                             kernel=kname + (" (windowed)" if windowed else ""),
                             windowed_c=min(grid[1], spec.d) if windowed else 0, wid=grid[0])
    return g


def form_name(form):
    prec, windowed = form
    return f"{'windowed' if windowed else 'unrolled'}_{prec}"


def leg_config(G, pb, args, spec, grid, k, *, forms=(("fp64", False),), n_samples=1, mean_value=False, split=0,
               cpu=True, max_iters=100000):
    out = {"workload": f"{spec.name}, box [{spec.lo[0]:g},{spec.hi[0]:g}]^{spec.d}, tol {spec.tol:g}",
           "grid": list(grid), "k_select": k, "n_samples": n_samples, "mean_value": mean_value,
           "lanes_per_child": split or 1}
    first = None
    for form in forms:
        g = form_solves(G, pb, args, spec, grid, k, form, n_samples=n_samples, mean_value=mean_value, split=split,
                        max_iters=max_iters)
        out[form_name(form)] = public(g)
        first = first or g
    if cpu and not args.no_cpu:
        print(f"# cpu {spec.name} {grid} S={n_samples}", file=sys.stderr, flush=True)
        out["cpu_baseline"] = with_ratio(cpu_reference(spec, grid, k, n_samples, args.cpu_budget,
                                                       gpu_iterations=first["iterations"]),
                                         first["median_s"], first["boxes_bounded"])
    return out


def leg_cold(G, pb):
    """Compile-inclusive time: hipRTC compile of the headline module from its
    source (the on-disk cache bypassed), load, problem set-up and one solve.
    The reference's own timer also covers its JIT (1-Ackley/ackley50.py:960,
    :1382-1390)."""
    import ctypes
    from gpugo_amd import _lib, emit
    spec = pb.ackley50()
    obj = pb.objective_loops(spec.family, spec.d)
    src = emit.window_module_source(obj, ("fast",), f32=True)
    lib = _lib.load()
    t0 = time.perf_counter()
    h = ctypes.c_void_p()
    _lib.check(lib.gpugo_module_compile(src.encode(), b"ackley50_cold.hip", b"", ctypes.byref(h)))
    t1 = time.perf_counter()

    class _M:  # a Module-like wrapper around the freshly compiled handle
        handle = h
    p = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=C2_GRID, k_select=C2_K,
                  precision="fp32", module=_M, windowed=True, max_iters=1000)
    r = p.solve()
    t2 = time.perf_counter()
    p.close()
    lib.gpugo_module_destroy(h)
    return {"workload": "ackley50, windowed FP32-prune / FP64-certify form, compiled from source",
            "compile_s": t1 - t0, "cold_time_to_certified_s": t2 - t0, "solve_s": t2 - t1,
            "certified": r["status_name"] == "certified"}


LARGE_D = tuple((n, 1000) for n in ("levy", "rosenbrock", "fu", "rastrigin", "ackley", "belegundu", "breiman",
                                     "salomon", "griewank", "styblinski", "zabinsky")) + (("levy", 2000),)
# the reference's levy10000.py dimension: one solve, a queue of up to 2^31 boxes
XL_D = (("levy", 10000),)
# Breiman's minimiser x* = 0 lies on the wid-3 split planes of its box
# [-1, 2] (every level), so it runs on the reference's wid 4
LARGE_D_GRID = {"breiman": (4, 9)}


def leg_large_d(G, cases=LARGE_D, reps=2, cpu_seconds=2.0, max_queue=1 << 28, warm=True, no_cpu=False):
    """The windowed (large-d) form on reference objectives at d = 1000 to
    10,000 (the reference's boxes and initial upper bounds, 3^12 children of
    one box per step, 10 diagonal samples, tol 1e-6): time to a certified
    enclosure, boxes bounded per second and the roofline of the windowed
    split+bound kernel.  CPU beside it: the oracle (the reference method,
    every coordinate of every child) on Rosenbrock and Levy at d = 1000,
    steps of 3^8 children from the root, all host threads: its boxes/s and
    the time it would need for the GPU's boxes."""
    from gpugo_amd import emit, families as F, problems as pb
    from oracle import oracle as O
    k, S = 1, 10
    dims = "/".join(sorted({str(d) for _, d in cases}, key=int))
    out = {"workload": f"reference objectives d={dims}, reference boxes and initial upper bounds, grid 3^12 "
                       f"(Breiman 4^9), one box per step, {S} diagonal samples, tol 1e-6, windowed code lists, "
                       f"queue up to 2^{max_queue.bit_length() - 1} boxes"}
    for name, d in cases:
        grid = LARGE_D_GRID.get(name, (3, 12))
        s = F.spec(name, d)
        obj = F.objective_loops(name, d)
        m = G.get_module(obj, roundings=("exact", "fast"), windowed=True)
        p = G.Problem(obj, [], s.lo, s.hi, tol=1e-6, grid=grid, k_select=k, n_samples=S,
                      init_ub=s.extra["init_ub"], module=m, windowed=True, max_queue=max_queue, max_iters=40000)
        if warm:
            p.solve()
        p.timing(True)
        times, res = [], []
        for _ in range(reps):
            p.reset()
            t0 = time.perf_counter()
            res.append(p.solve())
            times.append(time.perf_counter() - t0)
        p.timing(False)
        p.close()
        r = res[-1]
        t = statistics.median(times)
        rl = roofline(obj, res, n_samples=S, kernel="gpugo_bound_fast (windowed)", windowed_c=grid[1], wid=grid[0])
        leg = {"time_to_certified_s": t, "certified": all(x["status_name"] == "certified" for x in res),
               "lb": r["lb"], "ub": r["ub"], "width": r["width"], "f_star": s.f_star,
               "contains_f_star": r["lb"] <= s.f_star <= r["ub"], "iterations": r["iterations"],
               "boxes_bounded": r["boxes_bounded"], "boxes_bounded_per_s": r["boxes_bounded"] / t,
               "queue_peak": r["queue_peak"], "roofline": rl}
        if name in ("rosenbrock", "levy") and d == 1000 and not no_cpu:
            # the oracle's restatements: Rosenbrock as family 1 (a = b = 0,
            # w = 100), Levy as family 3 (no coefficients)
            fam, params = (1, pb.rosenbrock(d).params) if name == "rosenbrock" else (3, np.zeros(0))
            threads = host_threads()
            O.build()
            t0 = time.perf_counter()
            nsteps, boxes = 0, 0
            while nsteps == 0 or time.perf_counter() - t0 < cpu_seconds:
                o, _ = O.solve(fam, params, s.lo, s.hi, tol_f=1e-6, wid=3, cycle=8, k_select=1,
                               n_samples=S, max_iters=1, threads=threads)
                boxes += o["boxes_bounded"]
                nsteps += 1
            rate = boxes / (time.perf_counter() - t0)
            leg["cpu_baseline"] = {"boxes_bounded_per_s": rate, "cores": threads, "kind": "port",
                                   "sample": f"{nsteps} x one step of 3^8 children from the root, d={d}, "
                                             f"{S} samples, every coordinate of every child",
                                   "est_time_for_gpu_boxes_s": r["boxes_bounded"] / rate,
                                   "speedup_vs_cpu_est": r["boxes_bounded"] / rate / t}
        out[name if d == 1000 else f"{name}{d}"] = leg
    return out


def run_batch(Batch, probs, **kw):
    b = Batch(slots=len(probs), **kw)
    t0 = time.perf_counter()
    for p in probs:
        b.admit(p)
    res = {}
    while True:
        b.run(1)
        for tag, r in b.results():
            res[tag] = r
        if b.counts()[0] == 0:
            break
    b.sync()
    for tag, r in b.results():
        res[tag] = r
    el = time.perf_counter() - t0
    b.close()
    return res, el


def overflow_kinds(results):
    """Overflowed problems by the capacity that ran out."""
    over = [r.get("overflow", 0) for r in results if r["status_name"] == "overflow"]
    return {"box_queue": sum(1 for o in over if o & 1), "parent_arena": sum(1 for o in over if o & 2)}


# batch slots: two growth tiers (8x and 64x the per-slot queue and arena);
# the one-at-a-time solves get capacities that no problem of the workload
# outgrows
GROW_TIERS = 2
SEQ_QUEUE, SEQ_PARENTS = 1 << 23, 1 << 18


def leg_batch(G, pb, n, seq_sample, use_graph=True):
    """n independent problems (families sepcos / sepcos_chain / ackley_q, d
    log-uniform in [2,32]) in one batch (schedules from the tuned table)
    against the same problems solved one at a time with the single-problem
    solver (same schedules and capacities).  Throughput counts certified
    problems on both sides."""
    from gpugo_amd.batch import Batch, BatchProblem, get_batch_module
    from gpugo_amd.schedule import schedule_for
    qcap, pcap, maxit = 1 << 18, 8192, 2000
    specs = [pb.batch_problem(k) for k in range(n)]
    probs = []
    for k, s in enumerate(specs):
        grid, ks = schedule_for(s.family, s.d)
        probs.append(BatchProblem.from_spec(s, grid=grid, k_select=ks, max_iters=maxit, tag=k))
    m = get_batch_module(pb.FAMILIES, 32, ("fast",))
    kw = dict(module=m, max_queue=qcap, max_parents=pcap, steps_per_launch=8, grow=GROW_TIERS, use_graph=use_graph)
    run_batch(Batch, probs[:64], **kw)  # warm-up
    res, el = run_batch(Batch, probs, **kw)
    cert = sum(r["status_name"] == "certified" for r in res.values())
    sample = probs[:seq_sample]
    mods = {}
    for p in sample:
        key = (p.family, p.d)
        if key not in mods:
            mods[key] = G.get_module(pb.objective(p.family, p.d), roundings=("fast",))
    seq = {}
    t0 = time.perf_counter()
    for p in sample:
        obj = pb.objective(p.family, p.d)
        with G.Problem(obj, p.params, p.lo, p.hi, tol=p.tol, grid=p.grid, k_select=p.k_select,
                       max_iters=maxit, max_queue=SEQ_QUEUE, max_parents=SEQ_PARENTS,
                       module=mods[(p.family, p.d)]) as q:
            try:
                seq[p.tag] = q.solve()
            except G.GpugoError:  # capacity exceeded: no certificate
                seq[p.tag] = {"status_name": "overflow"}
    t_seq = time.perf_counter() - t0

    def same_result(a, b):
        if a["status_name"] == "overflow" or b["status_name"] == "overflow":
            return a["status_name"] == b["status_name"]
        return all(a[key] == b[key] for key in ("lb", "ub", "width", "iterations", "boxes_bounded"))

    same = sum(same_result(seq[t], res[t]) for t in seq)
    seq_cert = sum(r["status_name"] == "certified" for r in seq.values())
    widths = [r["width"] for r in res.values() if r["status_name"] == "certified"]
    return {"workload": f"{n} problems (sepcos / sepcos_chain / ackley_q, d log-uniform [2,32], box [-h,h]^d, "
                        "h~U[2,32], tol in {1e-4,1e-6}), seeds [3,k]",
            "slots": n, "steps_per_launch": 8, "graph": use_graph, "max_queue": qcap, "max_parents": pcap,
            "growth_tiers": GROW_TIERS, "sequential_capacity": {"max_queue": SEQ_QUEUE, "max_parents": SEQ_PARENTS},
            "max_iters": maxit, "time_s": el, "problems_per_s": n / el, "certified": cert,
            "certified_per_s": cert / el,
            "status_counts": {k: sum(r["status_name"] == k for r in res.values())
                              for k in sorted({r["status_name"] for r in res.values()})},
            "overflow_kinds": overflow_kinds(res.values()),
            "width_max_certified": max(widths) if widths else None,
            "boxes_bounded": sum(r["boxes_bounded"] for r in res.values()),
            "sequential": {"problems": len(sample), "time_s": t_seq, "problems_per_s": len(sample) / t_seq,
                           "certified": seq_cert, "certified_per_s": seq_cert / t_seq,
                           "identical_to_batch": same},
            "speedup_certified_per_s": (cert / el) / max(seq_cert / t_seq, 1e-12)}


def leg_band(G, pb, k=44, grid=(3, 3), k_select=64, steps=3000):
    """A problem of the batched workload whose queue grows past 10^6 entries
    (sepcos_chain d = 30), solved alone for a fixed number of steps with K
    boxes per step: the select ranks the priority band, or (band=False) the
    whole queue.  Same steps and results either way; per-step time."""
    s = pb.batch_problem(k)
    obj = pb.objective(s.family, s.d)
    m = G.get_module(obj, roundings=("fast",))
    out = {"workload": f"batch problem {k}: {s.family} d={s.d}, box [-{s.hi[0]:.2f},{s.hi[0]:.2f}]^{s.d}, "
                       f"tol {s.tol:g}, grid {grid[0]}^{grid[1]}, {k_select} boxes per step, {steps} steps"}
    res = {}
    for band in (True, False):
        with G.Problem(obj, s.params, s.lo, s.hi, tol=s.tol, grid=grid, k_select=k_select, max_iters=steps,
                       max_queue=1 << 26, max_parents=1 << 22, module=m, band=band) as p:
            t0 = time.perf_counter()
            r = p.solve()
            el = time.perf_counter() - t0
        res[band] = r
        out["band" if band else "whole_queue"] = {
            "seconds": el, "us_per_step": el / max(r["iterations"], 1) * 1e6, "status": r["status_name"],
            "iterations": r["iterations"], "boxes_bounded": r["boxes_bounded"], "queue_peak": r["queue_peak"],
            "width": r["width"]}
    out["identical"] = all(res[True][x] == res[False][x]
                           for x in ("lb", "ub", "iterations", "boxes_bounded", "queue_size", "queue_peak"))
    out["speedup"] = out["whole_queue"]["seconds"] / out["band"]["seconds"]
    return out


def leg_served(pb, lams=(100.0, 1000.0), seconds=3.0):
    """Served workload: Poisson arrivals, continuous batching, p50/p99 latency."""
    from gpugo_amd.serve import Server, latency_summary
    out = {}
    with Server(mem_fraction=0.5) as srv:
        slots = dict(srv.slots)
        for lam in lams:
            n = int(lam * seconds)
            arr = pb.arrivals(lam, n, seed=40)
            specs = [pb.served_problem(k) for k in range(n)]
            t0 = time.perf_counter()
            done = srv.run_trace(specs, arr, time_limit=seconds * 10 + 30)
            el = time.perf_counter() - t0
            summ = latency_summary(done)
            summ.update({"lambda_per_s": lam, "requests": n, "wall_s": el,
                         "throughput_per_s": len(done) / el,
                         "certified_per_s": summ["certified"] / el,
                         "status_counts": {k: sum(r.result["status_name"] == k for r in done)
                                           for k in sorted({r.result["status_name"] for r in done})},
                         "overflow_kinds": overflow_kinds(r.result for r in done)})
            out[f"lambda_{int(lam)}"] = summ
    out["slots_per_dmax_class"] = {str(k): v for k, v in slots.items()}
    out["workload"] = ("sepcos / sepcos_chain / ackley_q, d log-uniform [2,100], tol log-uniform [1e-8,1e-4], "
                       "box [-h,h]^d h~U[2,32], seeds [4,k]; arrivals exponential(1/lambda), seed 40")
    return out


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a problem shards by independent boxes, a batch by independent problems)", file=sys.stderr)
        sys.exit(2)
    import gpugo_amd as G
    from gpugo_amd import problems as pb

    legs = set(args.legs.split(","))
    G.init(0)
    spec = pb.ackley50()
    if args.no_headline:
        out = {"metric": METRIC, "value": None, "unit": "s", "n_gpus": 1, "legs_only": sorted(legs)}
        run_legs(G, pb, args, legs, spec, out, None)
        print(json.dumps(out), flush=True)
        return
    head = form_solves(G, pb, args, spec, C2_GRID, C2_K, HEADLINE_FORM)
    print(f"# headline done ({args.steps} solves)", file=sys.stderr, flush=True)
    rl = head["roofline"]
    rl["traffic"] = measured_traffic("gpugo_bound_fast")
    rl["traffic_source"] = TRAFFIC_FILE
    out = {
        "metric": METRIC,
        "value": head["mean_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["mean_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 prune, f64 certify",
        "data": "synthetic (seeded coefficients)",
        "config": {"workload": "ackley50: ackley_q d=50, box [-32,32]^50, tol 1e-6, seed 2",
                   "precision": "fp32 prune, fp64 certify", "code_lists": "windowed",
                   "grid": list(C2_GRID), "k_select": C2_K,
                   "n_samples": 1, "rounding": "fast", "mean_value": False,
                   "note": "certifies in one step: the child holding x* = 0 has natural-extension lower "
                           "bound 0 and its midpoint is x*; see c2_s10 for the ten-sample run"},
        "time_to_certified_s": head["mean_s"],
        "median_s": head["median_s"],
        "certified": head["certified"],
        "lb": head["lb"], "ub": head["ub"], "width": head["width"], "iterations": head["iterations"],
        "boxes_bounded": head["boxes_bounded"], "boxes_bounded_per_s": head["boxes_bounded_per_s"],
        "passes_identical": head["passes_identical"],
        "roofline": rl,
    }
    if not args.no_cpu:
        print("# cpu baseline (headline workload)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = with_ratio(cpu_reference(spec, C2_GRID, C2_K, 1, args.cpu_budget,
                                                       gpu_iterations=head["iterations"]),
                                         head["mean_s"], head["boxes_bounded"])
    run_legs(G, pb, args, legs, spec, out, head)
    print(json.dumps(out), flush=True)


def run_legs(G, pb, args, legs, spec, out, head):
    if "c2_forms" in legs:
        print("# leg c2_forms", file=sys.stderr, flush=True)
        out["c2_forms"] = leg_config(G, pb, args, spec, C2_GRID, C2_K, forms=OTHER_FORMS, cpu=False)
        if head is not None:
            out["c2_forms"]["headline_speedup_over"] = {
                form_name(f): out["c2_forms"][form_name(f)]["mean_s"] / head["mean_s"] for f in OTHER_FORMS}
    if "c2_s10" in legs:
        print("# leg c2_s10", file=sys.stderr, flush=True)
        out["c2_s10"] = leg_config(G, pb, args, spec, C2_GRID, C2_K, forms=(HEADLINE_FORM,) + OTHER_FORMS,
                                   n_samples=10)
    if "c1" in legs:
        print("# leg c1", file=sys.stderr, flush=True)
        out["c1"] = leg_config(G, pb, args, pb.chain10(), (4, 10), 1, forms=C1_FORMS)
    if "c1_tuned" in legs:
        print("# leg c1_tuned", file=sys.stderr, flush=True)
        out["c1_tuned"] = leg_config(G, pb, args, pb.chain10(), (4, 5), 32, mean_value=True, split=4, cpu=False)
    if "c0" in legs:
        print("# leg c0", file=sys.stderr, flush=True)
        out["c0"] = leg_config(G, pb, args, pb.sepcos2(), (64, 2), 1)
    if "cold" in legs:
        print("# leg cold", file=sys.stderr, flush=True)
        out["cold"] = leg_cold(G, pb)
    if "large_d" in legs:
        print("# leg large_d", file=sys.stderr, flush=True)
        cases = LARGE_D
        if args.large_d_cases:
            cases = tuple((c.split(":")[0], int(c.split(":")[1])) for c in args.large_d_cases.split(","))
        out["large_d"] = leg_large_d(G, cases, reps=args.large_d_reps, no_cpu=args.no_cpu)
    if "xl_d" in legs:
        print("# leg xl_d", file=sys.stderr, flush=True)
        out["xl_d"] = leg_large_d(G, XL_D, reps=1, max_queue=1 << 31, warm=False, no_cpu=True)
    if "batch" in legs:
        print("# leg batch", file=sys.stderr, flush=True)
        out["batch"] = leg_batch(G, pb, args.batch_n, args.seq_sample, use_graph=not args.batch_no_graph)
    if "served" in legs:
        print("# leg served", file=sys.stderr, flush=True)
        out["served"] = leg_served(pb)
    if "band" in legs:
        print("# leg band", file=sys.stderr, flush=True)
        out["band"] = leg_band(G, pb)


if __name__ == "__main__":
    main()
