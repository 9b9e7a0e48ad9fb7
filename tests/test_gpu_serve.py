"""Served form: continuous batching of problems that arrive over time.

Bar: every result of the served mix equals the same problem solved alone
(one slot, nothing else on the GPU) -- bit for bit, which is within any
tolerance; certified enclosures contain f* where it is known in closed form
and are no wider than the problem's tol; requests that arrive while others
run are admitted into free slots (some requests are admitted before others
finished).
"""
import numpy as np
import pytest

from gpugo_amd import problems as pb
from gpugo_amd.batch import Batch, get_batch_module
from gpugo_amd.serve import Server, latency_summary

pytestmark = pytest.mark.gpu
KEYS = ("lb", "ub", "width", "status", "iterations", "boxes_bounded", "queue_size")


def specs(n, dmax=32, dmin=1):
    """The first n served-workload problems with dmin <= d <= dmax."""
    out, k = [], 0
    while len(out) < n:
        s = pb.served_problem(k)
        k += 1
        if dmin <= s.d <= dmax:
            out.append(s)
    return out


@pytest.mark.parametrize("classes,slots,dmin,dmax,lam", [
    ((16, 32), {16: 3, 32: 2}, 1, 32, 500.0),
    ((64, 100), {64: 3, 100: 2}, 33, 100, 200.0),
], ids=["d<=32", "d33-100"])
def test_served_equals_alone(gpu, classes, slots, dmin, dmax, lam):
    """Poisson arrivals (the served workload's exponential gaps, seed 40)
    into fewer slots than requests."""
    ss = specs(24, dmax=dmax, dmin=dmin)
    arrivals = pb.arrivals(lam, len(ss), seed=40)
    with Server(slots=slots, dmax_classes=classes, max_iters=300) as srv:
        done = srv.run_trace(ss, arrivals, time_limit=90)
        assert len(done) == len(ss)
        byid = {r.rid: r for r in done}
        # some request was admitted while an earlier one was still running
        overlap = any(byid[i].admitted < byid[j].done and byid[i].arrival > byid[j].arrival
                      for i in byid for j in byid if i != j)
        assert overlap
        for rid, req in byid.items():
            prob = srv.problem(req)
            dm = srv.dmax_for(req.spec.d)
            with Batch(slots=1, module=get_batch_module(pb.FAMILIES, dm, ("fast",)),
                       max_queue=srv.max_queue, max_parents=8192) as b:
                alone = b.solve_all([prob])[rid]
            for key in KEYS:
                assert alone[key] == req.result[key], (rid, key)
            r = req.result
            if r["status_name"] == "certified":
                assert r["ub"] - r["lb"] <= req.spec.tol
                if req.spec.f_star is not None:
                    assert r["lb"] <= req.spec.f_star <= r["ub"]
    summ = latency_summary(done)
    assert summ["p50_ms"] <= summ["p99_ms"] and summ["all_within_tol"]
