"""Per-box code lists on the GPU (gpugo_eval_boxes) against the oracle.

Reference code lists: <fn>Value_IA / <fn>Derivative_IA / <fn>Sample_IA
(1-Ackley/ackley50.py:116-479 and the Rastrigin / Rosenbrock equivalents).
Bar: for an objective whose value is rational in x (Rosenbrock: its cos
terms are multiplied by b = 0) EXACT mode equals the oracle bit for bit in
value, gradient, monotonicity flag and sample bound.  With exp/sin/cos the
device and host libraries' nearest values may differ by an ulp, so the
bar is: every enclosure contains the 256-bit value of f (and of its
partial derivatives) at points of the box, and differs from the oracle's by
at most the outward-rounding slack 1e-12 * (1 + |f|).
"""
import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import problems as pb
from tests.test_oracle_objectives import f_true, random_boxes

pytestmark = pytest.mark.gpu
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}


def run_both(gpu, oracle, spec, n=400, rounding="exact", seed=0):
    rng = np.random.default_rng(seed)
    lo, hi = random_boxes(rng, spec, n)
    m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=(rounding,))
    g = m.eval_boxes(spec.params, lo, hi, spec.lo, spec.hi, 1, rounding=rounding)
    o = oracle.eval_boxes(FAM[spec.family], spec.params, lo, hi, spec.lo, spec.hi, 1)
    return lo, hi, g, o


def test_rational_objective_bit_identical(gpu, oracle):
    spec = pb.rosenbrock(6)
    lo, hi, g, o = run_both(gpu, oracle, spec, n=2000)
    for k in ("flo", "fhi", "glo", "ghi", "kkt", "ub"):
        assert np.array_equal(g[k], o[k]), k


@pytest.mark.parametrize("spec", [pb.sepcos2(), pb.chain10(), pb.ackley50(d=6), pb.ackley50(d=50),
                                  pb.rastrigin(4), pb.ackley(5)], ids=lambda s: s.name)
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_enclosures_contain_truth_and_match_oracle(gpu, oracle, spec, rounding):
    lo, hi, g, o = run_both(gpu, oracle, spec, n=300, rounding=rounding)
    rng = np.random.default_rng(5)
    slack = 1e-12 * (1 + np.abs(o["flo"]))
    assert np.all(np.abs(g["flo"] - o["flo"]) <= slack)
    assert np.all(np.abs(g["fhi"] - o["fhi"]) <= 1e-12 * (1 + np.abs(o["fhi"])))
    assert np.all(np.abs(g["ub"] - o["ub"]) <= 1e-12 * (1 + np.abs(o["ub"])))
    fin = np.isfinite(o["glo"]) & np.isfinite(g["glo"])
    assert np.all(np.abs(g["glo"][fin] - o["glo"][fin]) <= 1e-10 * (1 + np.abs(o["glo"][fin])))
    # monotonicity decisions agree except where a gradient bound sits within
    # the slack of zero
    diff = np.flatnonzero(g["kkt"] != o["kkt"])
    for i in diff:
        near = np.minimum(np.abs(o["glo"][i]), np.abs(o["ghi"][i])) <= 1e-10 * (1 + np.abs(o["ghi"][i]))
        assert near.any(), i
    for i in range(0, lo.shape[0], 15):
        x = lo[i] + rng.uniform(0, 1, spec.d) * (hi[i] - lo[i])
        fx = f_true(spec.family, spec.params, x)
        assert mp.mpf(g["flo"][i]) <= fx <= mp.mpf(g["fhi"][i])
        mid = lo[i] + (hi[i] - lo[i]) / 2
        assert mp.mpf(g["ub"][i]) >= f_true(spec.family, spec.params, mid)


def test_point_boxes_closed_form(gpu, oracle):
    for spec in (pb.sepcos2(), pb.ackley50(d=50), pb.rastrigin(7)):
        z = np.zeros((1, spec.d))
        m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=("fast",))
        g = m.eval_boxes(spec.params, z, z, spec.lo, spec.hi, 1, rounding="fast")
        assert g["flo"][0] <= 0.0 <= g["fhi"][0] and g["fhi"][0] - g["flo"][0] < 1e-12
        assert g["kkt"][0] == 0
