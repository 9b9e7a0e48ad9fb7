#!/usr/bin/env python3
"""Benchmark: time to a certified enclosure of the d = 10 sepcos_chain problem.

Workload: FP64 interval branch and bound
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
its HIP-event duration, measured in a second pass of the same solves so the
events do not add to value, against the MI355X FP64 vector peak) and the CPU
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
    # (wid, cycle) subregion grid, boxes selected per step, mean-value bound,
    # lanes per child of the term-parallel kernels (0: one lane per child)
    "tuned": {"grid": (4, 5), "k_select": 32, "mean_value": True, "split": 4},
    "reference": {"grid": (4, 10), "k_select": 1, "mean_value": False, "split": 0},
}


TRAFFIC_FILE = "profiles/traffic_chain10.json"


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter profile
    (rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by a device copy of
    known size); None if that profile has no row for it."""
    try:
        with open(os.path.join(ROOT, TRAFFIC_FILE)) as fh:
            return json.load(fh)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--schedule", choices=sorted(SCHEDULES), default="tuned")
    ap.add_argument("--rounding", choices=["fast", "exact"], default="fast")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--legs", default="fp32,c2,large_d,xl_d,batch,served",
                    help="comma list besides the headline: fp32 (headline workload in the FP32-prune form), "
                         "c2 (d=50), large_d (windowed form, d=1000 and 2000), xl_d (Levy d=10000, one solve), "
                         "batch (4096 problems), served (Poisson); 'none' for the headline only")
    ap.add_argument("--batch-n", type=int, default=4096)
    ap.add_argument("--seq-sample", type=int, default=256,
                    help="problems of the batch solved one at a time for the sequential baseline")
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
    # the same solve on one host thread (3 solves, median)
    t1 = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.solve(fam, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=sched["grid"][0], cycle=sched["grid"][1],
                k_select=sched["k_select"], n_samples=1, max_iters=100000, threads=1,
                mean_value=sched["mean_value"])
        t1.append(time.perf_counter() - t0)
    return {"value": t, "unit": "s", "cores": threads, "kind": "port",
            "sample": f"{len(times)} complete solves of the same workload and schedule, median",
            "boxes_bounded_per_s": boxes[0] / t, "certified": res["status"] == 1,
            "lb": res["lb"], "ub": res["ub"],
            "one_core": {"value": statistics.median(t1), "unit": "s", "cores": 1,
                         "sample": "3 complete solves on one host thread, median"}}


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


def bound_roofline(obj, results, mean_value=False, kernel="gpugo_bound_fast"):
    """Algorithmic FP64 FLOP of the split+bound launches over their
    HIP-event time: every child gets the value code list, survivors of the
    value test also the gradient and the sample (with the mean-value bound
    also the midpoint value and d interval products and additions)."""
    from gpugo_amd import emit
    fl = emit.flop_counts(obj)
    boxes = sum(x["boxes_bounded"] for x in results)
    pv = sum(x["pruned_value"] for x in results)
    bound_ms = sum(x["bound_ms"] for x in results)
    launches = sum(x["bound_launches"] for x in results)
    extra = fl["sample"] + fl["grad"] + (10 * obj.d if mean_value else 0)
    flops = boxes * fl["value"] + (boxes - pv) * extra
    achieved = flops / (bound_ms * 1e-3) / 1e12 if bound_ms > 0 else 0.0
    return {"bound": "valu", "kernel": kernel + ("_mv" if mean_value else ""),
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
            "flop_per_launch": flops / max(launches, 1), "flop_per_child": flops / max(boxes, 1),
            "avg_launch_ms": bound_ms / max(launches, 1), "launches": launches}


def leg_fp32(G, spec, obj, sched, args):
    """The headline workload in the FP32-prune / FP64-certify precision form,
    against the FP64 form with the same schedule (one lane per child: the
    FP32 pass has no term-parallel variant)."""
    def timed(precision):
        prob = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=sched["grid"],
                         k_select=sched["k_select"], n_samples=1, rounding=args.rounding, max_iters=200000,
                         mean_value=sched["mean_value"], precision=precision)
        for _ in range(args.warmup):
            prob.reset()
            prob.solve()
        res = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            prob.reset()
            res.append(prob.solve())
        t = (time.perf_counter() - t0) / args.steps
        prob.close()
        return t, res
    t64, r64s = timed("fp64")
    t32, r32s = timed("fp32")
    r, r64 = r32s[-1], r64s[-1]
    # the FP64 host oracle on the same workload and schedule
    from oracle import oracle as O
    O.build()
    o, _ = O.solve(1, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=sched["grid"][0], cycle=sched["grid"][1],
                   k_select=sched["k_select"], max_iters=100000, threads=os.cpu_count() or 1,
                   mean_value=sched["mean_value"])
    return {"precision": "fp32 prune, fp64 certify", "time_to_certified_s": t32, "fp64_time_s": t64,
            "oracle_fp64": {"lb": o["lb"], "ub": o["ub"], "width": o["ub"] - o["lb"], "certified": o["status"] == 1},
            "intersects_oracle": max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"]),
            "speedup_vs_fp64": t64 / t32,
            "certified": all(x["status_name"] == "certified" for x in r32s),
            "lb": r["lb"], "ub": r["ub"], "width": r["width"], "iterations": r["iterations"],
            "boxes_bounded": r["boxes_bounded"], "pruned_value": r["pruned_value"],
            "fp64_form": {"lb": r64["lb"], "ub": r64["ub"], "width": r64["width"]},
            "enclosures_intersect": max(r["lb"], r64["lb"]) <= min(r["ub"], r64["ub"])}


def leg_c2(G, pb, steps=5):
    """ackley_q d = 50, box [-32,32]^50, tol 1e-6: GPU time to certified enclosure
    against the oracle on all host cores with the same schedule."""
    from oracle import oracle as O
    spec = pb.ackley50()
    obj = pb.objective(spec.family, spec.d)
    grid, k = (3, 12), 1
    p = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=grid, k_select=k, max_iters=10000)
    p.solve()
    times, res = [], []
    for _ in range(steps):
        p.reset()
        t0 = time.perf_counter()
        p.solve()
        times.append(time.perf_counter() - t0)
    p.timing(True)  # roofline pass: the same solves with dispatch events
    for _ in range(steps):
        p.reset()
        res.append(p.solve())
    p.timing(False)
    p.close()
    r = res[-1]
    O.build()
    threads = os.cpu_count() or 1
    t0 = time.perf_counter()
    o, _ = O.solve(2, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=grid[0], cycle=grid[1], k_select=k,
                   max_iters=10000, threads=threads)
    tc = time.perf_counter() - t0
    gpu = statistics.median(times)
    # precision form: FP32 prune, FP64 certify
    p = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=grid, k_select=k, max_iters=10000,
                  precision="fp32")
    p.solve()
    t32 = []
    for _ in range(steps):
        p.reset()
        t0 = time.perf_counter()
        r32 = p.solve()
        t32.append(time.perf_counter() - t0)
    p.close()
    return {"workload": "ackley50: ackley_q d=50, box [-32,32]^50, tol 1e-6, seed 2",
            "fp32_prune": {"time_to_certified_s": statistics.median(t32), "lb": r32["lb"], "ub": r32["ub"],
                           "width": r32["width"], "certified": r32["status_name"] == "certified",
                           "intersects_oracle": max(r32["lb"], o["lb"]) <= min(r32["ub"], o["ub"]),
                           "pruned_value": r32["pruned_value"], "boxes_bounded": r32["boxes_bounded"],
                           "speedup_vs_fp64": gpu / statistics.median(t32),
                           "note": "one step from the root: an FP32 sample pass over all children gives "
                                   "that step its incumbent, then the FP32 value test rules children out"},
            "grid": list(grid), "k_select": k, "dtype": "f64",
            "time_to_certified_s": gpu, "certified": all(x["status_name"] == "certified" for x in res),
            "lb": r["lb"], "ub": r["ub"], "width": r["width"], "iterations": r["iterations"],
            "boxes_bounded": r["boxes_bounded"], "boxes_bounded_per_s": r["boxes_bounded"] / gpu,
            "roofline": bound_roofline(obj, res),
            "cpu_baseline": {"value": tc, "unit": "s", "cores": threads, "kind": "port",
                             "sample": "one complete solve, same workload and schedule",
                             "certified": o["status"] == 1, "lb": o["lb"], "ub": o["ub"]},
            "speedup_vs_cpu": tc / gpu}


LARGE_D = (("levy", 1000), ("rosenbrock", 1000), ("fu", 1000), ("rastrigin", 1000), ("levy", 2000))
# the reference's levy10000.py dimension: one solve (~30 s), a queue of up to 2^31 boxes
XL_D = (("levy", 10000),)


def leg_large_d(G, cases=LARGE_D, reps=2, cpu_seconds=2.0, max_queue=1 << 28, warm=True):
    """The windowed (large-d) form on reference objectives at d = 1000 to
    10,000 (the reference's boxes and initial upper bounds, 3^12 children of
    one box per step, 10 diagonal samples, tol 1e-6, a queue of up to
    max_queue boxes): time to a certified enclosure,
    boxes bounded per second and the roofline of the windowed split+bound
    kernel.  CPU beside it: the oracle (the reference method, every
    coordinate of every child) on Rosenbrock and Levy at d = 1000 (the ones
    of these it restates), steps of 3^8 children from the root, all host
    threads: its boxes/s and the time it would need for the GPU's boxes."""
    from gpugo_amd import emit, families as F, problems as pb
    from oracle import oracle as O
    grid, k, S = (3, 12), 1, 10
    dims = "/".join(sorted({str(d) for _, d in cases}, key=int))
    out = {"workload": f"reference objectives d={dims}, reference boxes and initial upper bounds, grid 3^12, "
                       f"one box per step, {S} diagonal samples, tol 1e-6, windowed code lists, "
                       f"queue up to 2^{max_queue.bit_length() - 1} boxes"}
    cpu_rate = None
    for name, d in cases:
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
        fl = emit.window_flop_counts(obj, grid[1])
        flops = sum(x["boxes_bounded"] * fl["value"]
                    + (x["boxes_bounded"] - x["pruned_value"]) * (fl["grad"] + S * fl["sample"]) for x in res)
        bound_ms = sum(x["bound_ms"] for x in res)
        launches = sum(x["bound_launches"] for x in res)
        achieved = flops / (bound_ms * 1e-3) / 1e12 if bound_ms > 0 else 0.0
        t = statistics.median(times)
        leg = {"time_to_certified_s": t, "certified": all(x["status_name"] == "certified" for x in res),
               "lb": r["lb"], "ub": r["ub"], "width": r["width"], "f_star": s.f_star,
               "contains_f_star": r["lb"] <= s.f_star <= r["ub"], "iterations": r["iterations"],
               "boxes_bounded": r["boxes_bounded"], "boxes_bounded_per_s": r["boxes_bounded"] / t,
               "queue_peak": r["queue_peak"],
               "roofline": {"bound": "valu", "kernel": "gpugo_bound_fast (windowed)", "achieved": achieved,
                            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                            "flop_per_child_value": fl["value"], "avg_launch_ms": bound_ms / max(launches, 1),
                            "launches": launches}}
        if name in ("rosenbrock", "levy") and d == 1000:
            # the oracle's restatements: Rosenbrock as family 1 (a = b = 0,
            # w = 100), Levy as family 3 (no coefficients)
            fam, params = (1, pb.rosenbrock(d).params) if name == "rosenbrock" else (3, np.zeros(0))
            threads = os.cpu_count() or 1
            O.build()
            t0 = time.perf_counter()
            nsteps, boxes = 0, 0
            while nsteps == 0 or time.perf_counter() - t0 < cpu_seconds:
                o, _ = O.solve(fam, params, s.lo, s.hi, tol_f=1e-6, wid=3, cycle=8, k_select=1,
                               n_samples=S, max_iters=1, threads=threads)
                boxes += o["boxes_bounded"]
                nsteps += 1
            cpu_rate = boxes / (time.perf_counter() - t0)
            leg["cpu_baseline"] = {"boxes_bounded_per_s": cpu_rate, "cores": threads, "kind": "port",
                                   "sample": f"{nsteps} x one step of 3^8 children from the root, d={d}, "
                                             f"{S} samples, every coordinate of every child",
                                   "est_time_for_gpu_boxes_s": r["boxes_bounded"] / cpu_rate,
                                   "speedup_vs_cpu_est": r["boxes_bounded"] / cpu_rate / t}
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


def leg_batch(G, pb, n, seq_sample):
    """n independent problems (families sepcos / sepcos_chain / ackley_q, d
    log-uniform in [2,32]) in one batch (schedules from the
    tuned table) against the same problems solved one at a time with the
    single-problem solver (same schedules and capacities)."""
    from gpugo_amd.batch import Batch, BatchProblem, get_batch_module
    from gpugo_amd.schedule import schedule_for
    qcap, pcap, maxit = 1 << 18, 8192, 2000
    specs = [pb.batch_problem(k) for k in range(n)]
    probs = []
    for k, s in enumerate(specs):
        grid, ks = schedule_for(s.family, s.d)
        probs.append(BatchProblem.from_spec(s, grid=grid, k_select=ks, max_iters=maxit, tag=k))
    m = get_batch_module(pb.FAMILIES, 32, ("fast",))
    kw = dict(module=m, max_queue=qcap, max_parents=pcap, steps_per_launch=8)
    run_batch(Batch, probs[:64], **kw)  # warm-up
    res, el = run_batch(Batch, probs, **kw)
    cert = sum(r["status_name"] == "certified" for r in res.values())
    # sequential: the first seq_sample problems, one gpugo_solve each
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
                       max_iters=maxit, max_queue=qcap, max_parents=pcap, module=mods[(p.family, p.d)]) as q:
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
            "slots": n, "steps_per_launch": 8, "graph": True, "max_queue": qcap, "max_parents": pcap,
            "max_iters": maxit, "time_s": el, "problems_per_s": n / el, "certified": cert,
            "certified_per_s": cert / el,
            "status_counts": {k: sum(r["status_name"] == k for r in res.values())
                              for k in sorted({r["status_name"] for r in res.values()})},
            "overflow_kinds": overflow_kinds(res.values()),
            "width_max_certified": max(widths) if widths else None,
            "boxes_bounded": sum(r["boxes_bounded"] for r in res.values()),
            "sequential": {"problems": len(sample), "time_s": t_seq, "problems_per_s": len(sample) / t_seq,
                           "certified": seq_cert, "identical_to_batch": same},
            "speedup_problems_per_s": (n / el) / (len(sample) / t_seq)}


def overflow_kinds(results):
    """Overflowed problems by the capacity that ran out."""
    over = [r.get("overflow", 0) for r in results if r["status_name"] == "overflow"]
    return {"box_queue": sum(1 for o in over if o & 1), "parent_arena": sum(1 for o in over if o & 2)}


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
    from gpugo_amd import emit, problems as pb

    legs = set(args.legs.split(","))
    spec = pb.chain10()
    sched = SCHEDULES[args.schedule]
    obj = pb.objective(spec.family, spec.d)
    G.init(0)
    prob = G.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=sched["grid"],
                     k_select=sched["k_select"], n_samples=1, rounding=args.rounding, max_iters=200000,
                     mean_value=sched["mean_value"], split=sched["split"])
    kernel = (f"gpugo_bound_{args.rounding}" + ("_mv" if sched["mean_value"] else "")
              + ("_g" if sched["split"] else ""))
    for _ in range(args.warmup):
        prob.reset()
        prob.solve()
    print(f"# warmup done ({args.warmup} solves)", file=sys.stderr, flush=True)

    # value: K solves without per-launch events (they add ~6 us a launch);
    # the roofline: the same K solves again, each split+bound launch timed by
    # events taken from its dispatch on the problem's stream
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prob.reset()
        results.append(prob.solve())
    t1 = time.perf_counter()
    per_solve = (t1 - t0) / args.steps
    prob.timing(True)
    timed = []
    for _ in range(args.steps):
        prob.reset()
        timed.append(prob.solve())
    prob.timing(False)
    r = results[-1]
    certified = all(x["status_name"] == "certified" for x in results + timed)
    same = all(all(a[k] == b[k] for k in ("lb", "ub", "iterations", "boxes_bounded")) for a, b in zip(results, timed))
    boxes = sum(x["boxes_bounded"] for x in timed)
    pv = sum(x["pruned_value"] for x in timed)
    bound_ms = sum(x["bound_ms"] for x in timed)
    launches = sum(x["bound_launches"] for x in timed)

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
        "config": {"workload": "chain10: sepcos_chain d=10, box [-5,10]^10, tol 1e-6, seed 1",
                   "schedule": args.schedule, "grid": list(sched["grid"]),
                   "k_select": sched["k_select"], "n_samples": 1, "rounding": args.rounding,
                   "mean_value": sched["mean_value"], "lanes_per_child": sched["split"] or 1},
        "time_to_certified_s": per_solve,
        "boxes_bounded_per_s": sum(x["boxes_bounded"] for x in results) / (t1 - t0),
        "boxes_bounded_per_solve": boxes // args.steps,
        "iterations_per_solve": r["iterations"],
        "certified": certified,
        "lb": r["lb"], "ub": r["ub"], "width": r["width"],
        "roofline": {"bound": "valu",
                     "kernel": kernel,
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(kernel),
                     "traffic_source": TRAFFIC_FILE,
                     "flop_per_launch": flops / max(launches, 1),
                     "avg_launch_ms": bound_ms / max(launches, 1), "launches": launches,
                     "timing": "a second pass of the same solves with dispatch events on each launch",
                     "passes_identical": same},
    }
    if "fp32" in legs:
        out["fp32_prune"] = leg_fp32(G, spec, obj, sched, args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(spec, sched, args.cpu_seconds)
        out["cpu_reference_schedule"] = cpu_reference_schedule(spec)
    if "c2" in legs:
        print("# leg c2", file=sys.stderr, flush=True)
        out["c2"] = leg_c2(G, pb)
    if "large_d" in legs:
        print("# leg large_d", file=sys.stderr, flush=True)
        out["large_d"] = leg_large_d(G)
    if "xl_d" in legs:
        print("# leg xl_d", file=sys.stderr, flush=True)
        out["xl_d"] = leg_large_d(G, XL_D, reps=1, max_queue=1 << 31, warm=False)
    if "batch" in legs:
        print("# leg batch", file=sys.stderr, flush=True)
        out["batch"] = leg_batch(G, pb, args.batch_n, args.seq_sample)
    if "served" in legs:
        print("# leg served", file=sys.stderr, flush=True)
        out["served"] = leg_served(pb)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
