"""Term tables of the windowed modules on the GPU.

Over a child, template-sum term j depends only on the digits of the window
coordinates in its stencil (the other coordinates are the parent record's),
so the step's prep kernel evaluates each term once per digit combination (at
the record's box and at its diagonal sample points, FP64 and FP32, and the
gradient components of local gradient templates) and the split+bound and
sample kernels add table entries instead of evaluating terms per child; a
gradient with template sums (all d components tested per child) reads each
component's local parts (the subexpressions in X(J + k) alone) from a parts
table instead.  The
table entry is the same interval operations on the same intervals, added in
the same order, so the bar is bit-identity with the per-child evaluation
(GPUGO_FLAG_NO_TABLES):
  - every reference objective at d = 3 with wrapping windows, several
    records per step (K = 16, 3^2 children each: a workgroup's children
    belong to many records), 1 and 10 samples, both roundings;
  - every reference objective at d = 50 (one box per step, 3^12 children,
    10 samples), the coefficient families at d = 6 (and d = 50 for the
    separable two) and the d = 10 chain problem on its 4^10 grid;
  - the FP32-prune form (FP32 tables, the FP32 sample pass) on the d = 50
    ackley_q headline problem and on Levy and Rastrigin at d = 50;
  - the 15-significant-digit IA (Levy d = 50);
  - Levy, Rastrigin and Ackley at d = 1000 (Ackley's gradient uses template
    sums, so every component is tested: its local parts come from the parts
    table).
Solves also certify and contain f* where it is known.
"""
import numpy as np
import pytest

from gpugo_amd import families as F, problems as pb

pytestmark = pytest.mark.gpu
WIN = sorted(F.TEMPLATES)
KEYS = ("status", "lb", "ub", "width", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt",
        "queue_size", "queue_peak")
REF_GRID = {"breiman": (4, 9)}


def both(G, obj, lo, hi, **kw):
    out = []
    for tables in (True, False):
        print(f"# solve {obj.name} tables={tables} {kw.get('grid')} K={kw.get('k_select')} "
              f"S={kw.get('n_samples')} {kw.get('rounding')} {kw.get('precision', 'fp64')}", flush=True)
        out.append(G.minimize(obj, lo, hi, windowed=True, tables=tables, **kw))
    a, b = out
    for key in KEYS:
        assert a[key] == b[key], (obj.name, key, a[key], b[key])
    return a


@pytest.mark.parametrize("name", WIN)
@pytest.mark.parametrize("S", [1, 10])
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_tables_equal_terms_small_d(gpu, name, S, rounding):
    d = 3
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    r = both(gpu, obj, s.lo, s.hi, tol=1e-6, grid=(3, 2), k_select=16, n_samples=S, init_ub=s.extra["init_ub"],
             module=m, rounding=rounding, max_iters=20000, max_queue=1 << 22)
    assert r["status_name"] == "certified" and r["lb"] <= s.f_star <= r["ub"], (name, r)


@pytest.mark.parametrize("name", WIN)
def test_tables_equal_terms_d50(gpu, name):
    d = 50
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    r = both(gpu, obj, s.lo, s.hi, tol=1e-6, grid=REF_GRID.get(name, (3, 12)), k_select=1, n_samples=10,
             init_ub=s.extra["init_ub"], module=m, rounding="fast", max_iters=20000, max_queue=1 << 24)
    assert r["status_name"] == "certified" and r["lb"] <= s.f_star <= r["ub"], (name, r)


@pytest.mark.parametrize("family,d,grid,k", [(f, 6, (3, 4), 16) for f in pb.FAMILIES]
                         + [("sepcos", 50, (3, 12), 1), ("ackley_q", 50, (3, 12), 1)])
def test_tables_equal_terms_coefficient_families(gpu, family, d, grid, k):
    rng = np.random.default_rng([7, d])
    params = pb._family_params(rng, family, d)
    lo, hi = np.full(d, -4.0), np.full(d, 4.0)
    obj = pb.objective_loops(family, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    r = both(gpu, obj, lo, hi, tol=1e-6, grid=grid, k_select=k, n_samples=1, params=params, module=m,
             rounding="fast", max_iters=20000, max_queue=1 << 24)
    assert r["status_name"] == "certified", r


def test_tables_equal_terms_chain10(gpu):
    """The d = 10 chain problem on its 4^10 grid (the window covers every
    coordinate; chain terms read two coordinates, 16 table entries each)."""
    spec = pb.chain10()
    obj = pb.objective_loops(spec.family, spec.d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    r = both(gpu, obj, spec.lo, spec.hi, tol=1e-6, grid=(4, 10), k_select=1, n_samples=1, params=spec.params,
             module=m, rounding="fast", max_iters=20000, max_queue=1 << 26)
    assert r["status_name"] == "certified", r


def test_tables_equal_terms_fp32_headline(gpu):
    """The d = 50 ackley_q problem in the FP32-prune / FP64-certify form:
    FP32 tables in the sample pass and the FP32 value test."""
    spec = pb.ackley50()
    obj = pb.objective_loops(spec.family, spec.d)
    m = gpu.get_module(obj, roundings=("fast",), f32=True, windowed=True)
    for S in (1, 10):
        r = both(gpu, obj, spec.lo, spec.hi, tol=1e-6, grid=(3, 12), k_select=1, n_samples=S, params=spec.params,
                 module=m, rounding="fast", precision="fp32", max_iters=20000, max_queue=1 << 24)
        assert r["status_name"] == "certified" and r["lb"] <= 0.0 <= r["ub"], r


@pytest.mark.parametrize("name", ["levy", "rastrigin"])
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_tables_equal_terms_fp32_reference(gpu, name, rounding):
    d = 50
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), f32=True, windowed=True)
    r = both(gpu, obj, s.lo, s.hi, tol=1e-6, grid=(3, 12), k_select=1, n_samples=10, init_ub=s.extra["init_ub"],
             module=m, rounding=rounding, precision="fp32", max_iters=20000, max_queue=1 << 24)
    assert r["status_name"] == "certified" and r["lb"] <= s.f_star <= r["ub"], r


def test_tables_equal_terms_digits15(gpu):
    d = 50
    s = F.spec("levy", d)
    obj = F.objective_loops("levy", d)
    m = gpu.get_module(obj, roundings=("digits15",), windowed=True)
    r = both(gpu, obj, s.lo, s.hi, tol=1e-6, grid=(3, 12), k_select=1, n_samples=10, init_ub=s.extra["init_ub"],
             module=m, rounding="digits15", max_iters=20000, max_queue=1 << 24)
    assert r["status_name"] == "certified" and r["lb"] <= s.f_star <= r["ub"], r


@pytest.mark.parametrize("name", ["levy", "rastrigin", "ackley"])
def test_tables_equal_terms_d1000(gpu, name):
    d = 1000
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    r = both(gpu, obj, s.lo, s.hi, tol=1e-6, grid=(3, 12), k_select=1, n_samples=10, init_ub=s.extra["init_ub"],
             module=m, rounding="fast", max_iters=20000, max_queue=1 << 29)
    assert r["status_name"] == "certified" and r["lb"] <= s.f_star <= r["ub"], r
