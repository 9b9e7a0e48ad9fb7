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


# ---- the windowed (large-d) form ------------------------------------------
# Each template sum of a child is the parent's FP64 rest sum (rounded
# outward to FP32) plus the window's terms in FP32.


def window_f32_module(obj):
    return get_module(obj, roundings=("exact", "fast"), f32=True, windowed=True)


@pytest.mark.parametrize("family,d", [("ackley_q", 6), ("sepcos_chain", 6)])
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_window_f32_enclosure_contains_true_values(gpu, oracle, family, d, rounding):
    if family == "ackley_q":
        spec = pb.ackley50(d=d)
    else:  # the chain family at d = 6 on the d = 10 chain's box
        spec = pb.ProblemSpec(family, d, pb._family_params(np.random.default_rng(3), family, d),
                              np.full(d, -5.0), np.full(d, 10.0), 1e-6)
    m = window_f32_module(pb.objective_loops(family, d))
    rng = np.random.default_rng(9)
    n = 120
    span = spec.hi - spec.lo
    lo = spec.lo + rng.uniform(0, 1, (n, d)) * span
    w = rng.choice([1e-6, 1e-3, 0.1, 1.0], (n, 1)) * span
    hi = np.minimum(lo + rng.uniform(0, 1, (n, d)) * w, spec.hi)
    flo, fhi = eval_boxes_f32(m, spec.params, lo, hi, rounding)
    o = oracle.eval_boxes(FAM[family], spec.params, lo, hi, spec.lo, spec.hi, 1)
    assert np.all(flo <= fhi)
    assert np.all(np.maximum(flo, o["flo"]) <= np.minimum(fhi, o["fhi"]))
    mp.mp.prec = 256
    for i in range(0, n, 4):
        for _ in range(3):
            x = lo[i] + rng.uniform(0, 1, d) * (hi[i] - lo[i])
            t = f_true(family, spec.params, x)
            assert flo[i] <= t <= fhi[i], (i, flo[i], t, fhi[i])
    mag = np.maximum(np.abs(o["flo"]), np.abs(o["fhi"])) + 1.0
    assert np.all((fhi - flo) <= (o["fhi"] - o["flo"]) * 1.01 + 2e-4 * mag)


@pytest.mark.parametrize("S", [1, 10])
def test_window_f32_ackley50_certifies(gpu, oracle, S):
    """The d = 50 problem in the windowed FP32-prune form: certified, contains
    f* = 0, intersects the windowed FP64 form and (S = 1) the oracle's
    enclosure under the same schedule; FP32 rules children out."""
    spec = pb.ackley50()
    obj = pb.objective_loops(spec.family, spec.d)
    kw = dict(tol=spec.tol, grid=(3, 12), k_select=1, n_samples=S, windowed=True, max_iters=5000,
              max_queue=1 << 24)
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, precision="fp32", module=window_f32_module(obj), **kw) as p:
        r = p.solve()
    with gpu.Problem(obj, spec.params, spec.lo, spec.hi, **kw) as p:
        r64 = p.solve()
    assert r["status_name"] == "certified" and r["ub"] - r["lb"] <= spec.tol, r
    assert r["lb"] <= 0.0 <= r["ub"]
    assert max(r["lb"], r64["lb"]) <= min(r["ub"], r64["ub"])
    assert r["pruned_value"] > 0
    if S == 1:
        o, _ = oracle.solve(2, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=3, cycle=12, k_select=1,
                            max_iters=100)
        assert o["status"] == 1 and max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])
        assert r["iterations"] == 1


@pytest.mark.parametrize("name", ["levy", "rastrigin"])
def test_window_f32_reference_d50_certifies(gpu, name):
    from gpugo_amd import families as F
    d = 50
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    with gpu.Problem(obj, [], s.lo, s.hi, tol=1e-6, grid=(3, 12), k_select=1, n_samples=10,
                     init_ub=s.extra["init_ub"], module=window_f32_module(obj), windowed=True, precision="fp32",
                     max_iters=20000, max_queue=1 << 24) as p:
        r = p.solve()
    assert r["status_name"] == "certified", r
    assert r["lb"] <= s.f_star <= r["ub"] and r["ub"] - r["lb"] <= 1e-6
