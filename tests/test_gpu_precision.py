"""FP32-prune / FP64-certify precision form.

Children are first bounded in FP32 intervals (outward rounded, one ulp per
operation in FAST mode, coefficients rounded outward to FP32 intervals);
only children whose FP32 lower bound does not exceed the incumbent get the
FP64 value, gradient and sample.  Bar:
  - FP32 enclosures contain the true value (256-bit mpmath) at random
    points of random boxes, for every family;
  - FP32 enclosures are no tighter than possible: they contain the FP64
    enclosure's true range, so they intersect the oracle's enclosure, and
    their width is within a few FP32 ulps of the FP64 width;
  - solves certify enclosures of width <= tol that contain f* (closed form)
    or intersect the FP64 oracle's enclosure, and prune children in FP32.
"""
import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import problems as pb
from gpugo_amd.solver import eval_boxes_f32, get_module
from tests.test_oracle_objectives import f_true

pytestmark = pytest.mark.gpu
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


@pytest.mark.parametrize("spec", [pb.sepcos2(), pb.rosenbrock(4), pb.chain10(), pb.ackley50(d=6)],
                         ids=lambda s: s.name)
@pytest.mark.parametrize("rounding", ["fast"])
def test_f32_enclosure_contains_true_values(gpu, oracle, spec, rounding):
    m = get_module(pb.objective(spec.family, spec.d), roundings=(rounding,), f32=True)
    rng = np.random.default_rng(7)
    n = 200
    span = spec.hi - spec.lo
    lo = spec.lo + rng.uniform(0, 1, (n, spec.d)) * span
    w = rng.choice([1e-6, 1e-3, 0.1, 1.0], (n, 1)) * span
    hi = np.minimum(lo + rng.uniform(0, 1, (n, spec.d)) * w, spec.hi)
    flo, fhi = eval_boxes_f32(m, spec.params, lo, hi, rounding)
    o = oracle.eval_boxes(FAM[spec.family], spec.params, lo, hi, spec.lo, spec.hi, 1)
    assert np.all(flo <= fhi)
    # the FP64 oracle enclosure and the FP32 one both contain the range
    assert np.all(np.maximum(flo, o["flo"]) <= np.minimum(fhi, o["fhi"]))
    mp.mp.prec = 256
    for i in range(0, n, 5):
        for _ in range(3):
            x = lo[i] + rng.uniform(0, 1, spec.d) * (hi[i] - lo[i])
            t = f_true(spec.family, spec.params, x)
            assert flo[i] <= t <= fhi[i], (i, flo[i], t, fhi[i])
    # FP32 widths: at most FP64 width + a few FP32 ulps of the magnitude
    mag = np.maximum(np.abs(o["flo"]), np.abs(o["fhi"])) + 1.0
    assert np.all((fhi - flo) <= (o["fhi"] - o["flo"]) * 1.01 + 2e-4 * mag)


@pytest.mark.parametrize("spec,grid,k,mv", [
    (pb.sepcos2(), (2, 2), 16, False),
    (pb.rosenbrock(4), (2, 4), 64, False),
    (pb.chain10(), (4, 5), 256, True),
    (pb.chain10(), (4, 5), 256, False),
    (pb.ackley50(), (3, 12), 1, False),
], ids=lambda v: getattr(v, "name", str(v)))
def test_f32_prune_fp64_certify(gpu, oracle, spec, grid, k, mv):
    obj = pb.objective(spec.family, spec.d)
    kw = dict(tol=spec.tol, grid=grid, k_select=k, mean_value=mv, max_iters=20000)
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, precision="fp32", **kw) as p:
        r = p.solve()
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, **kw) as p:
        r64 = p.solve()
    assert r["status_name"] == "certified" and r["width"] <= spec.tol, r
    assert r["ub"] - r["lb"] <= spec.tol
    if spec.f_star is not None:
        assert r["lb"] <= spec.f_star <= r["ub"]
    if spec.d <= 10:
        o, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=grid[0],
                            cycle=grid[1], k_select=k, max_iters=20000, mean_value=mv)
        assert o["status"] == 1
        assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])
    assert max(r["lb"], r64["lb"]) <= min(r["ub"], r64["ub"])
    if r["iterations"] > 1:
        assert r["pruned_value"] > 0
    if spec.d == 50:
        # one step from the root: the FP32 sample pass gives that step an
        # incumbent, so FP32 rules children out there (the FP64 form, which
        # samples the children that survive by value, rules out none in its
        # first step)
        assert r["iterations"] == r64["iterations"] == 1
        assert r["pruned_value"] > 0 and r64["pruned_value"] == 0
