"""Term-parallel split+bound (GPUGO_FLAG_SPLIT): the lanes of a group share
one child and split its sums (loop templates), then reduce.

Sums associate differently from the one-lane code lists, so bits may
differ; the bar is the mathematics: every certified enclosure has width
<= tol, contains f* (closed form) and intersects the one-lane form's and
the oracle's enclosures; runs are deterministic; the group reductions give
every lane of a group the same value (the same result for every launch
shape).
"""
import numpy as np
import pytest

from gpugo_amd import problems as pb

pytestmark = pytest.mark.gpu
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


@pytest.mark.parametrize("spec,grid,k,mv,rounding", [
    (pb.sepcos2(), (2, 2), 16, False, "fast"),
    (pb.rastrigin(3), (3, 3), 32, True, "fast"),
    (pb.rosenbrock(4), (2, 4), 64, False, "fast"),
    (pb.rosenbrock(4), (2, 4), 64, True, "exact"),
    (pb.ackley(4), (3, 4), 32, True, "fast"),
    (pb.chain10(), (4, 5), 32, True, "fast"),
    (pb.chain10(), (4, 5), 64, False, "fast"),
    (pb.ackley50(), (3, 12), 1, False, "fast"),
], ids=lambda v: getattr(v, "name", str(v)))
def test_split_certifies(gpu, oracle, spec, grid, k, mv, rounding):
    obj = pb.objective(spec.family, spec.d)
    kw = dict(tol=spec.tol, grid=grid, k_select=k, mean_value=mv, max_iters=20000, rounding=rounding)
    m = gpu.get_module(obj, roundings=(rounding,), split=4)
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, split=4, module=m, **kw) as p:
        r = p.solve()
        p.reset()
        r2 = p.solve()
    for key in ("lb", "ub", "iterations", "boxes_bounded", "queue_size"):
        assert r[key] == r2[key], key
    assert r["status_name"] == "certified" and r["ub"] - r["lb"] <= spec.tol, r
    if spec.f_star is not None:
        assert r["lb"] <= spec.f_star <= r["ub"]
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, module=m, **kw) as p:
        r1 = p.solve()
    assert max(r["lb"], r1["lb"]) <= min(r["ub"], r1["ub"])
    if spec.d <= 10:
        o, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=grid[0],
                            cycle=grid[1], k_select=k, max_iters=20000, mean_value=mv)
        assert o["status"] == 1 and max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])


def test_split_needs_templates(gpu):
    from gpugo_amd.expr import Objective, X, sqr
    obj = Objective(2, 0, sqr(X(0)) + sqr(X(1)), name="no_templates")
    with pytest.raises(ValueError):
        gpu.get_module(obj, roundings=("fast",), split=4)
