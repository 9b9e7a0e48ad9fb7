"""Batched solves: many problems resident on the GPU at once.

Each slot runs the per-problem loop (same select, split+bound, compaction),
so the bar is identity with the single-problem solver, which the other GPU
tests pin to the oracle:
  - every problem's result in a batch (lb, ub, width, status, iterations,
    boxes bounded, queue size) equals gpugo_solve with the same schedule;
  - the same problem gives the same result alone and in a mix, with slots
    admitted while others run (continuous batching) and with or without
    hipGraph capture;
  - EXACT rounding on the rational Rosenbrock follows the oracle exactly;
  - certified enclosures contain f* where it is known in closed form.
"""
import numpy as np
import pytest

from gpugo_amd import GpugoError, problems as pb
from gpugo_amd.batch import Batch, BatchProblem, get_batch_module

pytestmark = pytest.mark.gpu
DMAX = 8
KEYS = ("lb", "ub", "width", "status", "iterations", "boxes_bounded", "queue_size")


@pytest.fixture(scope="module")
def bmod(gpu):
    return get_batch_module(pb.FAMILIES, DMAX, ("exact", "fast"))


def mixed_problems(n, seed=0):
    """Small problems of all families and dimensions 1..DMAX."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        fam = pb.FAMILIES[k % 3]
        d = int(rng.integers(2, DMAX + 1))
        h = float(rng.uniform(2.0, 8.0))
        params = pb._family_params(rng, fam, d)
        tol = (1e-4, 1e-6)[k % 2]
        c = min(d, 6)
        out.append(BatchProblem(fam, d, params, np.full(d, -h), np.full(d, h), tol, grid=(2, c),
                                k_select=8 + 8 * (k % 3), max_iters=4000, tag=k))
    return out


def single(gpu, p: BatchProblem, rounding="fast"):
    obj = pb.objective(p.family, p.d)
    with gpu.Problem(obj, p.params, p.lo, p.hi, tol=p.tol, grid=p.grid, k_select=p.k_select,
                     n_samples=p.n_samples, rounding=rounding, max_iters=p.max_iters) as q:
        return q.solve()


def test_batch_equals_single(gpu, bmod):
    probs = mixed_problems(12)
    with Batch(slots=16, module=bmod, rounding="fast", steps_per_launch=4) as b:
        res = b.solve_all(probs)
    assert sorted(res) == [p.tag for p in probs]
    for p in probs:
        s = single(gpu, p)
        r = res[p.tag]
        for key in KEYS:
            assert r[key] == s[key], (p.tag, p.family, p.d, key, r[key], s[key])
        if r["status_name"] == "certified":
            assert r["ub"] - r["lb"] <= p.tol
            if p.family != "sepcos_chain":
                assert r["lb"] <= 0.0 <= r["ub"]


def test_max_iters_equals_single(gpu, bmod):
    """A slot stopped by its step limit reports what gpugo_solve reports
    after as many steps: the smallest live lower bound and ub - lb rounded
    up, bit for bit."""
    probs = mixed_problems(6, seed=5)
    for p in probs:
        p.max_iters = 3 + p.tag
    with Batch(slots=8, module=bmod, rounding="fast", steps_per_launch=4) as b:
        res = b.solve_all(probs)
    for p in probs:
        s = single(gpu, p)
        assert res[p.tag]["status_name"] == s["status_name"], p.tag
        for key in KEYS:
            assert res[p.tag][key] == s[key], (p.tag, key, res[p.tag][key], s[key])
    assert sum(r["status_name"] == "max_iters" for r in res.values()) >= 3


def test_alone_equals_mix_continuous(gpu, bmod):
    """Fewer slots than problems: slots are refilled while others run."""
    probs = mixed_problems(30, seed=1)
    with Batch(slots=7, module=bmod, rounding="fast", steps_per_launch=3) as b:
        mix = b.solve_all(probs, launches_per_poll=1)
    with Batch(slots=1, module=bmod, rounding="fast", steps_per_launch=5, use_graph=False) as b:
        for p in probs[:6]:
            alone = b.solve_all([p])
            for key in KEYS:
                assert alone[p.tag][key] == mix[p.tag][key], (p.tag, key)
    assert all(mix[p.tag]["status_name"] in ("certified", "max_iters") for p in probs)


def test_rational_exact_matches_oracle(gpu, bmod, oracle):
    spec = pb.rosenbrock(4)
    p = BatchProblem.from_spec(spec, grid=(2, 4), k_select=16, tag=5)
    with Batch(slots=4, module=bmod, rounding="exact", steps_per_launch=2) as b:
        r = b.solve_all([p])[5]
    o, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=2, cycle=4, k_select=16,
                        max_iters=20000)
    assert r["status_name"] == "certified" and o["status"] == 1
    for key in ("lb", "ub", "iterations", "boxes_bounded"):
        assert r[key] == o[key], (key, r[key], o[key])


def test_max_iters_and_capacity(gpu, bmod):
    spec = pb.ackley50(d=6)
    p = BatchProblem.from_spec(spec, grid=(2, 6), k_select=8, max_iters=5, tag=1)
    with Batch(slots=2, module=bmod, rounding="fast") as b:
        r = b.solve_all([p])[1]
        assert r["status_name"] == "max_iters" and r["iterations"] == 5
        assert r["lb"] <= 0.0 <= r["ub"]
        assert r["width"] >= r["ub"] - r["lb"]
        b.admit(p)
        b.admit(p)
        with pytest.raises(GpugoError):
            b.admit(p)
        b.solve_all([])
    with Batch(slots=1, module=bmod, rounding="fast", max_queue=64) as b:
        with pytest.raises(GpugoError):
            b.admit(BatchProblem.from_spec(spec, grid=(2, 6), k_select=2))  # 2 * 64 > 64
        with pytest.raises(GpugoError):
            b.admit(BatchProblem("sepcos", DMAX + 1, np.ones(3 * (DMAX + 1)), -np.ones(DMAX + 1),
                                 np.ones(DMAX + 1), 1e-6))


def test_mean_value_batch_equals_single(gpu):
    m = get_batch_module(pb.FAMILIES, DMAX, ("fast",), mean_value=True)
    probs = mixed_problems(9, seed=3)
    for p in probs:
        p.mean_value = True
    with Batch(slots=9, module=m, rounding="fast") as b:
        res = b.solve_all(probs)
    for p in probs:
        obj = pb.objective(p.family, p.d)
        with gpu.Problem(obj, p.params, p.lo, p.hi, tol=p.tol, grid=p.grid, k_select=p.k_select,
                         rounding="fast", max_iters=p.max_iters, mean_value=True) as q:
            s = q.solve()
        for key in KEYS:
            assert res[p.tag][key] == s[key], (p.tag, key)


def test_growth_equals_single(gpu, bmod):
    """Slots far too small for these problems (a 4096-entry queue and 64
    parent records) with two growth tiers above them: problems that outgrow
    a slot move into the next tier and resume, and every result equals the
    single-problem solve with capacities large enough from the start.  The
    same batch without tiers reports overflow, so problems did move."""
    probs = mixed_problems(16, seed=11)
    kw = dict(module=bmod, rounding="fast", steps_per_launch=4, max_queue=4096, max_parents=64)
    with Batch(slots=6, grow=2, **kw) as b:
        res = b.solve_all(probs)
    with Batch(slots=6, grow=0, **kw) as b:
        flat = b.solve_all(probs)
    assert any(r["status_name"] == "overflow" for r in flat.values())
    for p in probs:
        s = single(gpu, p)
        r = res[p.tag]
        assert r["status_name"] != "overflow", p.tag
        for key in KEYS:
            assert r[key] == s[key], (p.tag, p.family, p.d, key, r[key], s[key])


@pytest.mark.parametrize("promote_age", [0, 16])
def test_c3_workload_equals_single(gpu, promote_age):
    """The 4096-problem workload's first 64 problems as configured
    (problems.batch_problem, seeds [3, k]: the three families, d
    log-uniform in [2, 32], box [-h, h]^d with h ~ U[2, 32], tol in
    {1e-4, 1e-6}) with their tuned schedules on the dmax-32 batch module, the
    bench's per-slot capacities (queue 2^18, 8192 parent records) and two
    growth tiers: every result equals gpugo_solve with capacities large
    enough from the start, nothing overflows, and certified enclosures are
    no wider than tol and contain f* where it is known in closed form.
    promote_age 16: every problem that runs 16 steps also moves to tier 1
    (the served form's promotion), with the same results."""
    from gpugo_amd.schedule import schedule_for
    m = get_batch_module(pb.FAMILIES, 32, ("fast",))
    probs = []
    for k in range(64):
        s = pb.batch_problem(k)
        grid, ks = schedule_for(s.family, s.d)
        probs.append(BatchProblem.from_spec(s, grid=grid, k_select=ks, max_iters=2000, tag=k))
        probs[-1].extra["f_star"] = s.f_star
    with Batch(slots=24, module=m, rounding="fast", steps_per_launch=8, max_queue=1 << 18, max_parents=8192,
               grow=2, promote_age=promote_age) as b:
        res = b.solve_all(probs)
    assert sorted(res) == list(range(64))
    for p in probs:
        obj = pb.objective(p.family, p.d)
        with gpu.Problem(obj, p.params, p.lo, p.hi, tol=p.tol, grid=p.grid, k_select=p.k_select,
                         rounding="fast", max_iters=p.max_iters, max_queue=1 << 23, max_parents=1 << 18) as q:
            s = q.solve()
        r = res[p.tag]
        assert r["status_name"] != "overflow", p.tag
        for key in KEYS:
            assert r[key] == s[key], (p.tag, p.family, p.d, key, r[key], s[key])
        if r["status_name"] == "certified":
            assert r["ub"] - r["lb"] <= p.tol
            fs = p.extra["f_star"]
            if fs is not None:
                assert r["lb"] <= fs <= r["ub"], p.tag
