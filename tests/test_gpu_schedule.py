"""The autotuner on the GPU (gpugo_amd.schedule).

tune: every candidate schedule solves each class's sample in one batch;
the winner is a candidate that certifies the most problems, and solving
the sample again with it certifies them with enclosures that contain the
closed-form f* = 0 (sepcos, ackley_q) or intersect the single solve
(sepcos_chain).  tune_launch: each launch shape (bound workgroups, steps
per graph launch) runs the whole mix; every shape certifies every problem
with the same enclosures up to the order of ties, and the winner has the
highest certified rate of the rows it reports.  tune_lanes: one lane per
child and the term-parallel kernels (4 lanes per child) both certify the
same problem with intersecting enclosures, and the winner is the faster.
"""
import numpy as np
import pytest

from gpugo_amd import problems as pb, schedule as S

pytestmark = pytest.mark.gpu


def _sample(n, dmax=8, seed0=3 * 10 ** 6):
    out, k = [], 0
    while len(out) < n:
        s = pb.batch_problem(seed0 + k)
        k += 1
        if s.d <= dmax:
            out.append(s)
    return out


def _check(res, specs, single=None):
    for i, s in enumerate(specs):
        r = res[i]
        assert r["status_name"] == "certified", (s.name, r)
        assert r["ub"] - r["lb"] <= s.tol
        if s.f_star is not None:
            assert r["lb"] <= s.f_star <= r["ub"], (s.name, r)


def test_tune_picks_a_certifying_schedule(gpu):
    from gpugo_amd.batch import BatchProblem, get_batch_module
    m = get_batch_module(pb.FAMILIES, 32, ("fast",))
    specs = _sample(6, dmax=4)
    classes = {"mix/4": specs}
    cands = [(3, 4, 32), (3, 2, 64)]
    print("# tune 2 candidates x 6 problems", flush=True)
    table = S.tune(classes, cands, module=m, time_limit=10.0, log=lambda *a: None)
    assert list(table) == ["mix/4"] and tuple(table["mix/4"]) in cands
    wid, cyc, k = table["mix/4"]
    probs = [BatchProblem.from_spec(s, grid=(wid, min(cyc, s.d)), k_select=k, max_iters=2000, tag=i)
             for i, s in enumerate(specs)]
    res, _ = S._run_mix(probs, len(probs), m, 20.0, max_queue=1 << 18, max_parents=8192)
    _check(res, specs)


def test_tune_launch_shapes_agree(gpu):
    from gpugo_amd.batch import get_batch_module
    m = get_batch_module(pb.FAMILIES, 32, ("fast",))
    specs = _sample(24, dmax=8, seed0=4 * 10 ** 6)
    cands = [(0, 4), (1024, 2), (4096, 8)]
    print("# tune_launch 3 shapes x 24 problems", flush=True)
    best, rows = S.tune_launch(specs, cands, module=m, slots=16, time_limit=30.0, log=lambda *a: None)
    assert tuple(best) in cands and len(rows) == 3
    assert all(r["certified"] == r["n"] == len(specs) for r in rows), rows
    assert max(rows, key=lambda r: r["certified_per_s"])["bound_blocks"] == best[0]
    # every shape: the same certified enclosures of f*
    from gpugo_amd.batch import BatchProblem
    outs = []
    for bb, spl in cands:
        probs = []
        for i, s in enumerate(specs):
            grid, k = S.schedule_for(s.family, s.d)
            probs.append(BatchProblem.from_spec(s, grid=grid, k_select=k, max_iters=2000, tag=i))
        res, _ = S._run_mix(probs, 16, m, 30.0, max_queue=1 << 18, max_parents=8192, bound_blocks=bb,
                            steps_per_launch=spl)
        _check(res, specs)
        outs.append(res)
    for i in range(len(specs)):
        assert max(o[i]["lb"] for o in outs) <= min(o[i]["ub"] for o in outs), i


def test_tune_lanes(gpu):
    rng = np.random.default_rng([7, 4])
    params = pb._family_params(rng, "sepcos_chain", 4)
    obj = pb.objective("sepcos_chain", 4)
    lo, hi = np.full(4, -4.0), np.full(4, 4.0)
    print("# tune_lanes sepcos_chain d=4 lanes (1, 4)", flush=True)
    best, rows = S.tune_lanes(obj, lo, hi, lanes=(1, 4), params=params, tol=1e-6, grid=(3, 2), k_select=16,
                              max_iters=20000, repeats=2, log=lambda *a: None)
    assert best in (1, 4) and len(rows) == 2
    assert all(r["certified"] for r in rows)
    assert max(r["lb"] for r in rows) <= min(r["ub"] for r in rows)
    assert min(rows, key=lambda r: r["seconds"])["lanes"] == best
