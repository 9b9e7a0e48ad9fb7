"""Parent-record collection: a small parent arena gives the same solve.

Parent records are appended, one per split box, and queued boxes refer to
them by index.  When the arena has no room for a step's records, the select
kernel drops the records no queued box references and renumbers the rest in
order, so the (lb, parent, pos) order of the queue is unchanged.  Bar:
  - with an arena smaller than the records the solve writes in total, the
    EXACT rational Rosenbrock still follows the oracle (which never reuses
    records) step for step: same lb, ub, iterations, boxes bounded and queue
    size;
  - in FAST rounding a small arena and the default arena give identical
    results and identical live boxes, alone and in a batch;
  - an arena too small for the records in use is reported as overflow
    (batch slots without growth tiers), or the slot moves into a larger
    tier and its result equals the single solve (with tiers).
The cases are chosen from the oracle's own counts: the queue never holds
more boxes than 7/8 of the arena minus one step's records (so no overflow),
while the records written in total exceed the arena (so collection runs).
"""
import numpy as np
import pytest

from gpugo_amd import GpugoError, problems as pb
from gpugo_amd.batch import Batch, BatchProblem, get_batch_module

pytestmark = pytest.mark.gpu
KEYS = ("lb", "ub", "iterations", "boxes_bounded", "queue_size")


def gpu_problem(G, spec, **kw):
    return G.Problem(pb.objective(spec.family, spec.d), spec.params, spec.lo, spec.hi, **kw)


# (d, grid, k_select, arena): oracle at tol 1e-12 -- d=2 (2,1) K=1: 130
# steps, queue peak 47; d=2 (2,2) K=2: 39 steps, queue peak 56
@pytest.mark.parametrize("d,grid,k,arena", [(2, (2, 1), 1, 64), (2, (2, 2), 2, 72)])
def test_rational_small_arena_follows_oracle(gpu, oracle, d, grid, k, arena):
    spec = pb.rosenbrock(d)
    with gpu_problem(gpu, spec, tol=1e-12, grid=grid, k_select=k, rounding="exact", max_parents=arena) as p:
        prev = 0
        for steps in (10, 40, 60, 80, 100, 120, 140):
            r = p.step(steps - prev)
            prev = steps
            o, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, tol_f=1e-12, wid=grid[0], cycle=grid[1],
                                k_select=k, max_iters=steps)
            for key in KEYS:
                assert r[key] == o[key], (steps, key, r[key], o[key])
            if r["status_name"] != "running":
                break
    assert r["status_name"] == "certified" and o["status"] == 1
    assert r["iterations"] * k + 1 > arena  # the records written outgrew the arena


def test_fast_small_arena_equals_default(gpu):
    spec = pb.sepcos2()  # (2,2) K=1 tol 1e-12: 94 steps, queue peak 28
    out = []
    for arena in (40, 0):
        with gpu_problem(gpu, spec, tol=1e-12, grid=(2, 2), k_select=1, max_parents=arena) as p:
            r = p.solve()
            blo, bhi, blb, n = p.boxes(1000)
        out.append((r, blo[:n], bhi[:n], blb[:n]))
    (a, *ba), (b, *bb) = out
    for key in KEYS:
        assert a[key] == b[key], key
    assert a["status_name"] == "certified" and a["iterations"] + 1 > 40
    for x, y in zip(ba, bb):
        np.testing.assert_array_equal(x, y)


def test_arena_too_small_is_reported(gpu):
    spec = pb.sepcos2()
    with gpu_problem(gpu, spec, tol=1e-12, grid=(2, 2), k_select=1, max_parents=8) as p:
        with pytest.raises(GpugoError, match="CAPACITY"):
            p.solve()


def test_batch_small_arena_equals_single(gpu):
    bmod = get_batch_module(pb.FAMILIES, 8, ("exact", "fast"))
    cases = [(pb.sepcos2(), (2, 2), 1), (pb.rosenbrock(2), (2, 1), 1), (pb.rosenbrock(2), (2, 2), 2)]
    probs = [BatchProblem(s.family, s.d, s.params, s.lo, s.hi, 1e-12, grid=g, k_select=k, max_iters=100000, tag=i)
             for i, (s, g, k) in enumerate(cases)]
    with Batch(slots=4, module=bmod, rounding="fast", steps_per_launch=4, max_parents=72) as b:
        res = b.solve_all(probs)
    for (s, g, k), p in zip(cases, probs):
        with gpu_problem(gpu, s, tol=1e-12, grid=g, k_select=k, rounding="fast") as q:
            ref = q.solve()
        for key in KEYS:
            assert res[p.tag][key] == ref[key], (p.tag, key)
        assert res[p.tag]["status_name"] == "certified"
    # without growth tiers the slot reports the overflow; with them (the
    # default) it moves into a larger tier and certifies as alone
    with Batch(slots=1, module=bmod, rounding="fast", max_parents=8, grow=0) as b:
        r = b.solve_all(probs[:1])[0]
    assert r["status_name"] == "overflow"
    with Batch(slots=1, module=bmod, rounding="fast", max_parents=8) as b:
        r = b.solve_all(probs[:1])[0]
    s, g, k = cases[0]
    with gpu_problem(gpu, s, tol=1e-12, grid=g, k_select=k, rounding="fast") as q:
        ref = q.solve()
    assert r["status_name"] == "certified" and all(r[key] == ref[key] for key in KEYS)
