#!/usr/bin/env python3
"""Benchmark: time to a certified enclosure of configuration 1.

Workload (configs[1] of the benchmark set): FP64 interval branch and bound
on f(x) = sum_i a_i x_i^2 + b_i (1 - cos(c_i x_i)) + sum_{i<9} w_i (x_{i+1} -
x_i^2)^2 + (1 - x_i)^2, d = 10, seeded coefficients (seed 1), box
[-5, 10]^10, tol 1e-6 on the width of the value enclosure.

One "step" = one complete solve, from the root box to a certified
enclosure [lb, ub] with ub - lb <= 1e-6, with the objective module loaded
and the problem's device buffers allocated beforehand (inputs resident in
HBM); the timed region covers reset + solve and ends after the solve's
final device synchronisation.  value = mean time to certified enclosure
(seconds, lower is better).  Also reported: boxes bounded per second, the
roofline of the split+bound kernel (algorithmic FP64 FLOP per launch over
its HIP-event duration, against the MI355X FP64 vector peak) and the CPU
oracle on the same workload and schedule (cpu_baseline).

    python bench.py [--steps K] [--warmup W] [--schedule tuned|reference]
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
FP64_PEAK_TFLOPS = 78.6
SCHEDULES = {
    # (wid, cycle) subregion grid, boxes selected per step, mean-value bound
    "tuned": {"grid": (4, 5), "k_select": 256, "mean_value": True},
    "reference": {"grid": (4, 10), "k_select": 1, "mean_value": False},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--schedule", choices=sorted(SCHEDULES), default="tuned")
    ap.add_argument("--rounding", choices=["fast", "exact"], default="fast")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
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


def cpu_baseline(spec, sched, seconds):
    """The oracle (C restatement of the reference loop) on the same workload
    and schedule, all host threads, repeated for about `seconds`."""
    from oracle import oracle as O
    O.build()
    threads = os.cpu_count() or 1
    fam = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}[spec.family]
    times, boxes, res = [], [], None
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        res, _ = O.solve(fam, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=sched["grid"][0],
                         cycle=sched["grid"][1], k_select=sched["k_select"], n_samples=1,
                         max_iters=100000, threads=threads, mean_value=sched["mean_value"])
        times.append(time.perf_counter() - t0)
        boxes.append(res["boxes_bounded"])
        if times[-1] > seconds:
            break
    t = statistics.median(times)
    return {"value": t, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} complete solves of the same workload and schedule, median",
            "boxes_bounded_per_s": boxes[0] / t, "certified": res["status"] == 1,
            "lb": res["lb"], "ub": res["ub"]}


def cpu_reference_schedule(spec, steps=3):
    """Boxes/s of the oracle on the reference schedule (4^10 children, one
    box per step), over a bounded number of steps."""
    from oracle import oracle as O
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    res, _ = O.solve(1, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=4, cycle=10, k_select=1,
                     max_iters=steps, threads=threads)
    dt = time.perf_counter() - t0
    return {"boxes_bounded_per_s": res["boxes_bounded"] / dt, "steps": steps, "cores": threads}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(problems shard by independent boxes/problems, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import gpugo_amd as G
    from gpugo_amd import emit, problems as pb

    spec = pb.config1()
    sched = SCHEDULES[args.schedule]
    obj = pb.objective(spec.family, spec.d)
    G.init(0)
    prob = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=sched["grid"],
                     k_select=sched["k_select"], n_samples=1, rounding=args.rounding, max_iters=200000,
                     mean_value=sched["mean_value"])
    for _ in range(args.warmup):
        prob.reset()
        prob.solve()
    print(f"# warmup done ({args.warmup} solves)", file=sys.stderr, flush=True)

    prob.timing(True)
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.reset()
        results.append(prob.solve())
    t1 = time.perf_counter()
    prob.timing(False)
    per_solve = (t1 - t0) / args.steps
    r = results[-1]
    certified = all(x["status_name"] == "certified" for x in results)
    boxes = sum(x["boxes_bounded"] for x in results)
    pv = sum(x["pruned_value"] for x in results)
    bound_ms = sum(x["bound_ms"] for x in results)
    launches = sum(x["bound_launches"] for x in results)

    fl = emit.flop_counts(obj)
    # survivors of the value test also get the sample and the gradient; with
    # the mean-value bound they also get the midpoint value and d interval
    # products and additions (10 FLOP each)
    extra = fl["sample"] + fl["grad"] + (10 * spec.d if sched["mean_value"] else 0)
    flops = boxes * fl["value"] + (boxes - pv) * extra
    achieved = flops / (bound_ms * 1e-3) / 1e12 if bound_ms > 0 else 0.0

    out = {
        "metric": METRIC,
        "value": per_solve,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_solve * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded coefficients)",
        "config": {"workload": "config1: sepcos_chain d=10, box [-5,10]^10, tol 1e-6, seed 1",
                   "schedule": args.schedule, "grid": list(sched["grid"]),
                   "k_select": sched["k_select"], "n_samples": 1, "rounding": args.rounding,
                   "mean_value": sched["mean_value"]},
        "time_to_certified_s": per_solve,
        "boxes_bounded_per_s": boxes / (t1 - t0),
        "boxes_bounded_per_solve": boxes // args.steps,
        "iterations_per_solve": r["iterations"],
        "certified": certified,
        "lb": r["lb"], "ub": r["ub"], "width": r["width"],
        "roofline": {"bound": "valu",
                     "kernel": f"gpugo_bound_{args.rounding}" + ("_mv" if sched["mean_value"] else ""),
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flop_per_launch": flops / max(launches, 1),
                     "avg_launch_ms": bound_ms / max(launches, 1), "launches": launches},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, sched, args.cpu_seconds)
        out["cpu_reference_schedule"] = cpu_reference_schedule(spec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
