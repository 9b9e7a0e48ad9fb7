"""Windowed (large-d) modules on the GPU.

A windowed module is built from an objective's loop templates only; each
template sum is enclosed as (its terms outside the parent record's split
window, summed once per parent over the parent box and at its diagonal
sample points) + (the terms the window touches, over the child).  The
reference evaluates every coordinate for every child
(GPUGO_1.1/6-Levy/levy10000.py:160-380).  Bar:
  - explicit boxes through the windowed module's loop-form code lists:
    enclosures contain the 256-bit values and derivatives of the definitions
    (tests/test_families.py), both roundings; for the rational Rosenbrock in
    EXACT rounding the value enclosure is bit-identical to the unrolled
    module's (the same left fold);
  - solves at d = 3 with cycle 2 (windows wrap past d): certified, contain
    the closed-form f*, and intersect the unrolled module's certified
    enclosure under the same schedule, with 1 and 10 diagonal samples;
  - the coefficient families at d = 6 (cycle 4, wrapping): certified
    enclosures intersect the oracle's under the same schedule;
  - two solves are identical (deterministic);
  - with a parent arena small enough that records are collected (their rest
    sums move with them) the solve equals the default arena's;
  - d = 50 for all eleven reference families (the product objectives
    Griewank, Styblinski, Zabinsky through template products), d = 1000 for
    Ackley, Belegundu, Breiman, Fu, Levy, Rastrigin, Rosenbrock, Salomon and
    the product objectives, and d = 2000, 5000 and 10,000 for Levy (the
    reference's levy1000.py ... levy10000.py boxes and initial upper bounds,
    10 diagonal samples, 3^12 children per step, one box per step, a queue of
    up to 2^29 boxes, 2^31 at d = 10,000): certified enclosures of the
    closed-form f*;
  - one box per step, the default queue (compacted every few steps, chunk
    minima between) gives the same solve as compaction after every step,
    also with the capacity set to the eager solve's queue peak.
"""
import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import GpugoError, families as F, problems as pb
from tests.test_families import f_ref

pytestmark = pytest.mark.gpu
WIN = sorted(F.TEMPLATES)
RES_KEYS = ("lb", "ub", "width", "iterations", "boxes_bounded", "queue_size", "status")


def wmod(G, obj):
    return G.get_module(obj, roundings=("exact", "fast"), windowed=True)


def ref_solve(G, name, d, grid, k, S, module, windowed, tol=1e-6, **kw):
    s = F.spec(name, d, tol=tol)
    obj = F.objective_loops(name, d) if windowed else F.objective(name, d)
    print(f"# solve {name} d={d} grid={grid} K={k} S={S} windowed={windowed}", flush=True)
    return G.minimize(obj, s.lo, s.hi, tol=tol, grid=grid, k_select=k, n_samples=S,
                      init_ub=s.extra["init_ub"], module=module, rounding="fast", windowed=windowed, **kw), s


@pytest.mark.parametrize("name", WIN)
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_window_eval_contains_truth(gpu, name, rounding):
    d = 3
    fam = F.REFERENCE[name]
    m = wmod(gpu, F.objective_loops(name, d))
    rng = np.random.default_rng(21)
    n = 32
    span = fam.hi - fam.lo
    lo = fam.lo + rng.uniform(0, 1, (n, d)) * span
    w = rng.choice([1e-7, 1e-3, 0.05, 0.5], (n, 1)) * span
    hi = np.minimum(lo + rng.uniform(0, 1, (n, d)) * w, fam.hi)
    g = m.eval_boxes(np.zeros(0), lo, hi, np.full(d, fam.lo), np.full(d, fam.hi), 1, rounding=rounding)
    mp.mp.prec = 256
    for i in range(n):
        for _ in range(3):
            x = lo[i] + rng.uniform(0, 1, d) * (hi[i] - lo[i])
            t = f_ref(name, x)
            assert g["flo"][i] <= t <= g["fhi"][i], (name, i, g["flo"][i], t, g["fhi"][i])
        if i % 4 == 0:
            x = lo[i] + 0.5 * (hi[i] - lo[i])
            for j in range(d):
                def fj(s, j=j):
                    y = [mp.mpf(float(v)) for v in x]
                    y[j] = s
                    return f_ref(name, y)
                dj = mp.diff(fj, mp.mpf(float(x[j])))
                assert g["glo"][i, j] - 1e-30 <= dj <= g["ghi"][i, j] + 1e-30, (name, i, j)


@pytest.mark.parametrize("name", ["ackley", "levy", "rastrigin", "rosenbrock"])
def test_window_digits15_eval(gpu, name):
    """The paper's 15-significant-digit IA in a windowed module: on explicit
    boxes the value enclosure contains the 256-bit value of the definition
    and is no narrower than the EXACT-rounding enclosure of the same module
    form (up to an ulp-sized slack), and wider than it by at most 1e-12 of
    the magnitudes involved."""
    d = 3
    fam = F.REFERENCE[name]
    obj = F.objective_loops(name, d)
    m15 = gpu.get_module(obj, roundings=("digits15",), windowed=True)
    mx = wmod(gpu, obj)
    rng = np.random.default_rng(23)
    n = 48
    span = fam.hi - fam.lo
    lo = fam.lo + rng.uniform(0, 1, (n, d)) * span
    w = rng.choice([1e-7, 1e-3, 0.05, 0.5], (n, 1)) * span
    hi = np.minimum(lo + rng.uniform(0, 1, (n, d)) * w, fam.hi)
    a = m15.eval_boxes(np.zeros(0), lo, hi, rounding="digits15", want=("f",))
    b = mx.eval_boxes(np.zeros(0), lo, hi, rounding="exact", want=("f",))
    # each operation is held no narrower than EXACT and widens by about one
    # 15-digit step of its own result (terms of these objectives are below
    # ~1e2 where f is small); transcendental steps evaluated at other points
    # may nest only up to an ulp, so the comparison allows 1e-15 of slack
    mag = np.maximum(np.abs(b["flo"]), np.abs(b["fhi"])) + 100.0
    assert np.all(a["flo"] <= b["flo"] + 1e-15 * mag) and np.all(a["fhi"] >= b["fhi"] - 1e-15 * mag)
    assert np.all(a["fhi"] - a["flo"] >= b["fhi"] - b["flo"] - 1e-15 * mag)
    assert np.all(np.abs(b["flo"] - a["flo"]) <= 1e-12 * mag)
    assert np.all(np.abs(a["fhi"] - b["fhi"]) <= 1e-12 * mag)
    mp.mp.prec = 256
    for i in range(n):
        x = lo[i] + rng.uniform(0, 1, d) * (hi[i] - lo[i])
        t = f_ref(name, x)
        assert a["flo"][i] <= t <= a["fhi"][i], (name, i)


@pytest.mark.parametrize("name", ["levy", "rastrigin"])
def test_window_digits15_solve(gpu, name):
    """A windowed solve in the 15-digit IA (d = 50, the reference's box,
    initial upper bound and 10 samples): certified, contains f*, and its
    enclosure intersects the FAST-rounding solve's."""
    d = 50
    s = F.spec(name, d)
    obj = F.objective_loops(name, d)
    kw = dict(tol=1e-6, grid=(3, 12), k_select=1, n_samples=10, init_ub=s.extra["init_ub"], windowed=True,
              max_iters=20000, max_queue=1 << 24)
    print(f"# solve {name} d={d} digits15 windowed", flush=True)
    a = gpu.minimize(obj, s.lo, s.hi, rounding="digits15",
                     module=gpu.get_module(obj, roundings=("digits15",), windowed=True), **kw)
    b = gpu.minimize(obj, s.lo, s.hi, rounding="fast", module=wmod(gpu, obj), **kw)
    for r in (a, b):
        assert r["status_name"] == "certified", r
        assert r["lb"] <= s.f_star <= r["ub"] and r["width"] <= 1e-6
    assert max(a["lb"], b["lb"]) <= min(a["ub"], b["ub"])


def test_window_eval_rational_bit_identical(gpu):
    d = 3
    fam = F.REFERENCE["rosenbrock"]
    mw = wmod(gpu, F.objective_loops("rosenbrock", d))
    mu = gpu.get_module(F.objective("rosenbrock", d), roundings=("exact", "fast"))
    rng = np.random.default_rng(22)
    lo = rng.uniform(fam.lo, fam.hi, (64, d))
    hi = np.minimum(lo + rng.uniform(0, 5, (64, d)), fam.hi)
    a = mw.eval_boxes(np.zeros(0), lo, hi, rounding="exact", want=("f",))
    b = mu.eval_boxes(np.zeros(0), lo, hi, rounding="exact", want=("f",))
    assert np.array_equal(a["flo"], b["flo"]) and np.array_equal(a["fhi"], b["fhi"])


@pytest.mark.parametrize("name", WIN)
@pytest.mark.parametrize("S", [1, 10])
def test_window_solve_matches_unrolled(gpu, name, S):
    d, grid, k = 3, (3, 2), 16
    rw, s = ref_solve(gpu, name, d, grid, k, S, wmod(gpu, F.objective_loops(name, d)), True,
                      max_iters=20000, max_queue=1 << 22)
    ru, _ = ref_solve(gpu, name, d, grid, k, S, gpu.get_module(F.objective(name, d), roundings=("exact", "fast")),
                      False, max_iters=20000, max_queue=1 << 22)
    for r in (rw, ru):
        assert r["status_name"] == "certified", (name, r)
        assert r["lb"] <= s.f_star <= r["ub"] and r["ub"] - r["lb"] <= 1e-6, (name, r)
    assert max(rw["lb"], ru["lb"]) <= min(rw["ub"], ru["ub"])


@pytest.mark.parametrize("family", pb.FAMILIES)
def test_window_coefficient_families_match_oracle(gpu, oracle, family):
    d, grid, k = 6, (3, 4), 16
    rng = np.random.default_rng([7, d])
    params = pb._family_params(rng, family, d)
    lo, hi = np.full(d, -4.0), np.full(d, 4.0)
    obj = pb.objective_loops(family, d)
    print(f"# solve {family} d={d} windowed", flush=True)
    r = gpu.minimize(obj, lo, hi, tol=1e-6, grid=grid, k_select=k, params=params, module=wmod(gpu, obj),
                     windowed=True, max_iters=20000, max_queue=1 << 22)
    fam = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}[family]
    o, _ = oracle.solve(fam, params, lo, hi, tol_f=1e-6, wid=grid[0], cycle=grid[1], k_select=k, max_iters=20000)
    assert r["status_name"] == "certified" and o["status"] == 1, (r, o)
    assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"]), (r, o)
    if family != "sepcos_chain":
        assert r["lb"] <= 0.0 <= r["ub"]


@pytest.mark.parametrize("d", [3, 6])
def test_levy_matches_oracle(gpu, oracle, d):
    """Levy (the reference's box and initial upper bound, 10 diagonal
    samples) through the windowed and the unrolled modules against the
    oracle's restatement (family 3) under the same schedule: all certified,
    pairwise intersecting enclosures of f* = 0."""
    s = F.spec("levy", d)
    grid, k = (3, 2), 16
    o, _ = oracle.solve(3, np.zeros(0), s.lo, s.hi, tol_f=1e-6, wid=grid[0], cycle=grid[1], k_select=k,
                        n_samples=10, init_ub=s.extra["init_ub"], max_iters=20000)
    rw, _ = ref_solve(gpu, "levy", d, grid, k, 10, wmod(gpu, F.objective_loops("levy", d)), True,
                      max_iters=20000, max_queue=1 << 22)
    ru, _ = ref_solve(gpu, "levy", d, grid, k, 10, gpu.get_module(F.objective("levy", d), roundings=("exact", "fast")),
                      False, max_iters=20000, max_queue=1 << 22)
    assert o["status"] == 1 and o["lb"] <= 0.0 <= o["ub"]
    for r in (rw, ru):
        assert r["status_name"] == "certified" and r["lb"] <= 0.0 <= r["ub"], r
        assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])


def test_window_deterministic(gpu):
    m = wmod(gpu, F.objective_loops("levy", 50))
    a, _ = ref_solve(gpu, "levy", 50, (3, 12), 1, 10, m, True, max_queue=1 << 24)
    b, _ = ref_solve(gpu, "levy", 50, (3, 12), 1, 10, m, True, max_queue=1 << 24)
    assert all(a[key] == b[key] for key in RES_KEYS)


def test_window_small_arena_equals_default(gpu):
    """Collection of parent records moves their rest sums: for some arena
    that is outgrown (records written > arena) without overflow the result
    equals the default arena's; smaller arenas may report overflow."""
    d, grid = 6, (2, 2)
    m = wmod(gpu, F.objective_loops("rastrigin", d))
    base, _ = ref_solve(gpu, "rastrigin", d, grid, 1, 1, m, True, tol=1e-9, max_iters=20000)
    assert base["status_name"] == "certified"
    collected = 0
    for arena in (16, 24, 32, 48, 64, 96):
        r, _ = ref_solve(gpu, "rastrigin", d, grid, 1, 1, m, True, tol=1e-9, max_iters=20000, max_parents=arena)
        if r["status_name"] == "overflow":
            assert r["overflow"] & 2
            continue
        assert all(r[key] == base[key] for key in RES_KEYS), (arena, r, base)
        collected += r["iterations"] + 1 > arena
    assert collected, "no arena size exercised the collection"


@pytest.mark.parametrize("name,d,S,kw", [("levy", 50, 10, {}), ("rastrigin", 50, 10, {}),
                                         ("rosenbrock", 50, 1, {"max_queue": 1 << 24}),
                                         ("rastrigin", 6, 1, {"tol": 1e-9, "grid": (2, 2), "max_parents": 16})])
def test_window_lazy_queue_matches_eager(gpu, name, d, S, kw):
    """The default one-box-per-step queue (compacted every few steps, chunk
    minima between) against compaction after every step: the same solve
    (with a small parent arena: records are collected while stale entries
    are in the queue)."""
    kw = dict(kw)
    tol, grid = kw.pop("tol", 1e-6), kw.pop("grid", (3, 12))
    m = wmod(gpu, F.objective_loops(name, d))
    a, s = ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, **kw)
    b, _ = ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, eager_compact=True, **kw)
    assert a["status_name"] == "certified" and a["iterations"] > 2, a
    if "max_parents" in kw:
        assert a["iterations"] + 1 > kw["max_parents"], "the arena was never collected"
    for key in RES_KEYS + ("width", "pruned_value", "pruned_kkt"):
        assert a[key] == b[key], (key, a[key], b[key])
    assert a["queue_peak"] >= b["queue_peak"]


# (objective, d, samples, grid, tol): one-box-per-step solves; those whose
# eager queue peak leaves room for the lazy schedule (peak / 8 >= 2 wid^cycle)
# are compared at that capacity
CAP_CASES = [("rastrigin", 50, 1, (2, 4), 1e-6), ("rastrigin", 50, 1, (3, 4), 1e-6), ("levy", 50, 1, (3, 4), 1e-6),
             ("ackley", 50, 1, (2, 4), 1e-6)]


def test_lazy_queue_at_capacity_matches_eager(gpu):
    """Capacity set to the eagerly compacted solve's queue peak (the most
    entries it ever held, so it just fits): the lazy queue, which keeps
    consumed and ruled-out entries until its next compaction, fits too and
    gives the same solve; one entry less overflows the eager queue."""
    exercised = 0
    for name, d, S, grid, tol in CAP_CASES:
        m = wmod(gpu, F.objective_loops(name, d))
        e, _ = ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, eager_compact=True)
        assert e["status_name"] == "certified", e
        cap = int(e["queue_peak"])
        print(f"# eager queue peak {cap}, wid^cycle {grid[0] ** grid[1]}", flush=True)
        if cap // 8 // (grid[0] ** grid[1]) < 2:
            continue
        exercised += 1
        a, _ = ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, max_queue=cap)
        b, _ = ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, max_queue=cap,
                         eager_compact=True)
        for key in RES_KEYS + ("width", "pruned_value", "pruned_kkt"):
            assert a[key] == b[key] == e[key], (name, grid, key, a[key], b[key], e[key])
        assert a["queue_peak"] >= cap         # stale entries are held beside the live ones
        with pytest.raises(GpugoError, match="CAPACITY"):
            ref_solve(gpu, name, d, grid, 1, S, m, True, tol=tol, max_iters=20000, max_queue=cap - 1,
                      eager_compact=True)
    assert exercised >= 2, "too few cases had a queue large enough for the lazy schedule"


# Breiman's box is [-1, 2] (3-Breiman/cosine50.py:857) and its minimiser is
# x* = 0: with wid 3 the first split planes are -1, 0, 1, 2, so x* lies on a
# face of 2^d children and none of them can be ruled out.  The reference's
# own wid 4 (planes -1, -0.25, 0.5, 1.25, 2) keeps x* inside one child at
# every level; Breiman runs with 4^9 children per step.
REF_GRID = {"breiman": (4, 9)}


@pytest.mark.parametrize("name", WIN)
def test_window_d50_certifies(gpu, name):
    d = 50
    r, s = ref_solve(gpu, name, d, REF_GRID.get(name, (3, 12)), 1, 10, wmod(gpu, F.objective_loops(name, d)), True,
                     max_iters=20000, max_queue=1 << 24)
    assert r["status_name"] == "certified", (name, r)
    assert r["lb"] <= s.f_star <= r["ub"] and r["ub"] - r["lb"] <= 1e-6, (name, r)


@pytest.mark.parametrize("name,d", [("levy", 1000), ("rosenbrock", 1000), ("fu", 1000), ("rastrigin", 1000),
                                    ("ackley", 1000), ("belegundu", 1000), ("breiman", 1000), ("salomon", 1000),
                                    ("griewank", 1000), ("styblinski", 1000), ("zabinsky", 1000),
                                    ("levy", 2000), ("levy", 5000), ("levy", 10000)])
def test_window_large_d_certifies(gpu, name, d):
    r, s = ref_solve(gpu, name, d, REF_GRID.get(name, (3, 12)), 1, 10, wmod(gpu, F.objective_loops(name, d)), True,
                     max_iters=20000, max_queue=1 << (31 if d > 5000 else 29))
    assert r["status_name"] == "certified", (name, r)
    assert r["lb"] <= s.f_star <= r["ub"] and r["ub"] - r["lb"] <= 1e-6, (name, r)
