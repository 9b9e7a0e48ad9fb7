"""Persistent step loop (GPUGO_FLAG_PERSIST): one cooperative launch runs
the steps with grid barriers between select, split+bound and compaction.

The phases run the same code as the separate launches of the term-parallel
form, so results must be identical to it (lb, ub, width, iterations, boxes
bounded, queue size), also when a solve is split over several launches
(gpugo_step), and certified enclosures must contain f*.
"""
import pytest

from gpugo_amd import problems as pb

pytestmark = pytest.mark.gpu
KEYS = ("lb", "ub", "width", "iterations", "boxes_bounded", "queue_size", "status")


@pytest.mark.parametrize("spec,grid,k,mv,rounding", [
    (pb.sepcos2(), (2, 2), 16, False, "fast"),
    (pb.rosenbrock(4), (2, 4), 64, True, "fast"),
    (pb.rosenbrock(4), (2, 4), 64, False, "exact"),
    (pb.ackley(4), (3, 4), 32, True, "fast"),
    (pb.chain10(), (4, 5), 32, True, "fast"),
    (pb.ackley50(), (3, 12), 1, False, "fast"),
], ids=lambda v: getattr(v, "name", str(v)))
def test_persistent_equals_launches(gpu, spec, grid, k, mv, rounding):
    obj = pb.objective(spec.family, spec.d)
    m = gpu.get_module(obj, roundings=(rounding,), split=4)
    kw = dict(tol=spec.tol, grid=grid, k_select=k, mean_value=mv, max_iters=20000, rounding=rounding,
              module=m, split=4)
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, **kw) as p:
        ref = p.solve()
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, persistent=True, **kw) as p:
        r = p.solve()
        p.reset()
        r2 = p.solve()
        p.reset()
        part = p.step(3)
        assert part["iterations"] <= 3
        r3 = p.solve()
    for key in KEYS:
        assert r[key] == ref[key], (key, r[key], ref[key])
        assert r2[key] == ref[key], key
        assert r3[key] == ref[key], key
    assert r["status_name"] == "certified" and r["ub"] - r["lb"] <= spec.tol
    if spec.f_star is not None:
        assert r["lb"] <= spec.f_star <= r["ub"]
