"""Device-resident branch and bound against the oracle.

Reference loop: 1-Ackley/ackley50.py:1033-1380 (sample -> rule out ->
list prune -> pop the smallest lower bound -> cycle the split window).
Bar:
  - Rosenbrock (rational) in EXACT mode follows the oracle's path exactly:
    same lb, ub, iterations, boxes bounded, pruned counts and queue sizes
    after every number of steps, and the same live boxes;
  - every certified enclosure contains f* (closed form where known, the
    oracle's enclosure otherwise) and is no wider than tol;
  - runs are deterministic;
  - the reference x-width rule returns boxes containing the minimiser;
  - a queue that overflows its capacity is reported, never silently dropped.
"""
import numpy as np
import pytest

from gpugo_amd import GpugoError, problems as pb

pytestmark = pytest.mark.gpu
FAM = {"sepcos": 0, "sepcos_chain": 1, "ackley_q": 2}
# pruned_value / pruned_kkt are not compared: the fused kernel rules children
# out against the incumbent at the start of the step and the list prune
# removes the rest, while the reference order samples first; the surviving
# sets are the same, the split between the two counters is not.
KEYS = ("lb", "ub", "iterations", "boxes_bounded", "queue_size")


def gpu_problem(G, spec, **kw):
    obj = pb.objective(spec.family, spec.d)
    return G.Problem(obj, spec.params, spec.lo, spec.hi, **kw)


@pytest.mark.parametrize("grid,k", [((4, 3), 1), ((2, 2), 8), ((3, 4), 32)])
def test_rational_path_identical_to_oracle(gpu, oracle, grid, k):
    spec = pb.rosenbrock(4)
    with gpu_problem(gpu, spec, tol=1e-6, grid=grid, k_select=k, rounding="exact") as p:
        for steps in (1, 2, 3, 5, 8, 13):
            r = p.step(steps - (0 if steps == 1 else prev))
            prev = steps
            o, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=grid[0], cycle=grid[1],
                                k_select=k, max_iters=steps)
            if r["status_name"] == "certified" or o["status"] == 1:
                assert r["status_name"] == "certified" and o["status"] == 1
                assert (r["lb"], r["ub"]) == (o["lb"], o["ub"])
                break
            for key in KEYS:
                assert r[key] == o[key], (steps, key, r[key], o[key])
        glo, ghi, glb, n = p.boxes(100000)
    assert n == r["queue_size"]


@pytest.mark.parametrize("spec,grid,k", [
    (pb.sepcos2(), (64, 2), 1),
    (pb.sepcos2(), (2, 2), 64),
    (pb.chain10(), (2, 5), 256),
    (pb.rastrigin(3), (2, 3), 64),
    (pb.rosenbrock(4), (2, 4), 64),
    (pb.ackley(4), (2, 4), 64),
    (pb.ackley50(d=50), (3, 12), 1),
], ids=lambda v: getattr(v, "name", str(v)))
@pytest.mark.parametrize("rounding", ["fast", "exact"])
def test_certified_enclosure(gpu, oracle, spec, grid, k, rounding):
    if rounding == "exact" and spec.d > 10:
        # d = 50 in EXACT rounding: the windowed module of the same objective
        # (its unrolled EXACT module is only used on explicit boxes)
        obj = pb.objective_loops(spec.family, spec.d)
        m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
        print(f"# solve {spec.name} exact windowed grid={grid} K={k}", flush=True)
        with gpu.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=grid, k_select=k,
                         rounding="exact", module=m, windowed=True, max_iters=20000) as p:
            r = p.solve()
    else:
        with gpu_problem(gpu, spec, tol=spec.tol, grid=grid, k_select=k, rounding=rounding,
                         max_iters=20000) as p:
            r = p.solve()
    assert r["status_name"] == "certified", r
    assert r["ub"] - r["lb"] <= spec.tol
    if spec.f_star is not None:
        assert r["lb"] <= spec.f_star <= r["ub"]
    else:
        o, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=spec.tol, wid=grid[0],
                            cycle=grid[1], k_select=k, max_iters=20000)
        assert o["status"] == 1
        assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])


@pytest.mark.parametrize("spec,grid,k", [(pb.rosenbrock(4), (2, 4), 16), (pb.chain10(), (2, 5), 256)],
                         ids=lambda v: getattr(v, "name", str(v)))
def test_mean_value_mode_matches_oracle(gpu, oracle, spec, grid, k):
    """The optional mean-value lower bound (GPUGO_FLAG_MEAN_VALUE) against the
    oracle's same mode: identical path for the rational objective, certified
    and overlapping enclosures otherwise, and no more steps than without it."""
    with gpu_problem(gpu, spec, tol=1e-6, grid=grid, k_select=k, rounding="exact", mean_value=True,
                     max_iters=20000) as p:
        r = p.solve()
    o, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=grid[0],
                        cycle=grid[1], k_select=k, max_iters=20000, mean_value=True)
    assert r["status_name"] == "certified" and o["status"] == 1
    if spec.family == "sepcos_chain" and spec.params[:2 * spec.d].max() == 0:
        for key in KEYS:
            assert r[key] == o[key], key
    assert max(r["lb"], o["lb"]) <= min(r["ub"], o["ub"])
    plain, _ = oracle.solve(FAM[spec.family], spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=grid[0],
                            cycle=grid[1], k_select=k, max_iters=20000)
    assert r["iterations"] <= plain["iterations"]
    # the plain (natural-extension) enclosure is pinned independently of the
    # mean-value code (closed form, L-BFGS witness): both must meet it
    assert plain["status"] == 1
    assert max(r["lb"], plain["lb"]) <= min(r["ub"], plain["ub"])
    if spec.f_star is not None:
        assert r["lb"] <= spec.f_star <= r["ub"]


def test_deterministic(gpu):
    spec = pb.chain10()
    out = []
    for _ in range(2):
        with gpu_problem(gpu, spec, tol=1e-6, grid=(2, 5), k_select=128) as p:
            out.append(p.solve())
    for key in KEYS + ("queue_peak",):
        assert out[0][key] == out[1][key], key


def test_x_width_rule(gpu):
    spec = pb.rastrigin(2)
    with gpu_problem(gpu, spec, tol=0.0, tol_x=1e-3, grid=(4, 2), k_select=8, max_iters=100000) as p:
        r = p.solve()
        olo, ohi, olb, n = p.boxes(4096)
    assert r["status_name"] == "x_converged" and n >= 1
    assert np.all(ohi - olo < 1e-3)
    assert any(np.all(olo[i] <= 0) and np.all(ohi[i] >= 0) for i in range(len(olb)))
    assert r["lb"] <= 0.0 <= r["ub"]
    # the output boxes around the one minimiser merge into one root
    # (combine/connectivity, 11-Rosenbrock/rosenbrock50.py:926-1049)
    from gpugo_amd.results import record, roots
    rlo, rhi, owner = roots(olo, ohi, 1e-6)
    assert len(rlo) == 1 and np.all(rlo[0] <= 0) and np.all(rhi[0] >= 0)
    rec = record(r, (olo, ohi, olb), tol=1e-6)
    assert rec["status"] == "x_converged" and len(rec["roots"]["lo"]) == 1


def test_queue_overflow_is_reported(gpu):
    spec = pb.chain10()
    with gpu_problem(gpu, spec, tol=1e-6, grid=(4, 10), k_select=1, max_queue=16) as p:
        with pytest.raises(GpugoError, match="CAPACITY"):
            p.solve()


def test_reference_grid_matches_oracle_first_steps(gpu, oracle):
    """Config 1 on the reference schedule (wid 4, 10 coordinates, one box per
    step): same incumbent and lower bound as the oracle within slack."""
    spec = pb.chain10()
    with gpu_problem(gpu, spec, tol=1e-6, grid=(4, 10), k_select=1, rounding="exact") as p:
        r = p.step(3)
    o, _ = oracle.solve(1, spec.params, spec.lo, spec.hi, tol_f=1e-6, wid=4, cycle=10, k_select=1, max_iters=3)
    assert r["boxes_bounded"] == o["boxes_bounded"] == 3 * 4 ** 10
    assert abs(r["ub"] - o["ub"]) <= 1e-12 * abs(o["ub"])
    assert abs(r["lb"] - o["lb"]) <= 1e-12 * (1 + abs(o["lb"]))
    assert abs(r["queue_size"] - o["queue_size"]) <= max(2, 0.001 * o["queue_size"])


def test_reference_grid_certifies(gpu):
    """Config 1 on its configured grid to the end (4^10 children of one box
    per step, the reference's schedule, natural extension): certified, the
    enclosure intersects the tuned-schedule solve's (4^5, K = 32, mean-value
    bound, 4 lanes per child), and its lower bound is below the value of an
    L-BFGS witness evaluated in 256-bit arithmetic."""
    from scipy.optimize import minimize as sp_minimize
    from tests.test_oracle_objectives import f_true
    spec = pb.chain10()
    print("# solve chain10 grid 4^10 K=1", flush=True)
    with gpu_problem(gpu, spec, tol=1e-6, grid=(4, 10), k_select=1, max_iters=20000) as p:
        r = p.solve()
    with gpu_problem(gpu, spec, tol=1e-6, grid=(4, 5), k_select=32, mean_value=True, split=4) as p:
        t = p.solve()
    assert r["status_name"] == "certified" and r["width"] <= 1e-6, r
    assert t["status_name"] == "certified"
    assert max(r["lb"], t["lb"]) <= min(r["ub"], t["ub"])
    rng = np.random.default_rng(0)
    f = lambda x: float(f_true(spec.family, spec.params, x))
    best = min((sp_minimize(f, rng.uniform(spec.lo, spec.hi), method="L-BFGS-B",
                            bounds=list(zip(spec.lo, spec.hi))) for _ in range(10)), key=lambda z: z.fun)
    assert r["lb"] <= float(f_true(spec.family, spec.params, best.x)) + 1e-12


@pytest.mark.parametrize("spec,grid,k,kw", [
    (pb.rosenbrock(4), (2, 4), 16, {"rounding": "exact"}),
    (pb.sepcos2(), (2, 2), 16, {}),
    (pb.chain10(), (4, 5), 32, {"mean_value": True, "split": 4}),
], ids=lambda v: getattr(v, "name", None))
def test_fused_select_matches_separate(gpu, spec, grid, k, kw):
    """A solve runs each next step's select inside the compaction kernel;
    one step per call never does (the call's last step keeps them apart).
    Both give the same solve."""
    with gpu_problem(gpu, spec, tol=1e-6, grid=grid, k_select=k, **kw) as p:
        a = p.solve()
    with gpu_problem(gpu, spec, tol=1e-6, grid=grid, k_select=k, **kw) as p:
        for _ in range(100000):
            b = p.step(1)
            if b["status_name"] != "running":
                break
    for key in ("lb", "ub", "width", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt", "status"):
        assert a[key] == b[key], (key, a[key], b[key])


def test_lazy_queue_boxes_between_calls(gpu):
    """A call's last step compacts, so between gpugo_step calls the live
    boxes (gpugo_boxes_copy) and counters of the default one-box-per-step
    queue equal those of compaction after every step."""
    spec = pb.rastrigin(3)
    kw = dict(tol=1e-6, grid=(4, 3), k_select=1, max_iters=20000)

    def rows(p):
        lo, hi, lb, n = p.boxes(1 << 16)
        assert n == len(lb)
        a = np.column_stack([lb, lo, hi])
        return a[np.lexsort(a.T[::-1])]

    with gpu_problem(gpu, spec, **kw) as a, gpu_problem(gpu, spec, eager_compact=True, **kw) as b:
        calls = 0
        for n in (2, 2, 3, 3, 5, 8, 13):
            ra, rb = a.step(n), b.step(n)
            for key in ("lb", "ub", "iterations", "boxes_bounded", "queue_size", "status"):
                assert ra[key] == rb[key], (n, key, ra[key], rb[key])
            assert np.array_equal(rows(a), rows(b)), n
            calls += 1
            if ra["status_name"] != "running":
                break
        assert calls >= 3 and ra["status_name"] == "certified"


def _live_rows(p, cap=1 << 16):
    lo, hi, lb, n = p.boxes(cap)
    a = np.column_stack([lb, lo, hi])
    return a[np.lexsort(a.T[::-1])], n


def test_stop_final_compaction_small_queue(gpu):
    """The stop step's final compaction, run by the workgroup whose select
    stopped the problem (a queue of at most 65,536 entries): gpugo_solve
    (first step alone, then steps two at a time) leaves the same live boxes,
    queue size and enclosure as steps taken in calls of 3, whose last step
    always compacts in a launch of its own."""
    spec = pb.rastrigin(3)
    kw = dict(tol=1e-6, grid=(4, 3), k_select=1, max_iters=20000)
    with gpu_problem(gpu, spec, **kw) as a, gpu_problem(gpu, spec, **kw) as b:
        ra = a.solve()
        while True:
            rb = b.step(3)
            if rb["status_name"] != "running":
                break
        assert ra["status_name"] == rb["status_name"] == "certified"
        for key in ("lb", "ub", "width", "iterations", "boxes_bounded", "queue_size"):
            assert ra[key] == rb[key], (key, ra[key], rb[key])
        xa, na = _live_rows(a)
        xb, nb = _live_rows(b)
        assert na == nb == ra["queue_size"] and 0 < na <= 65536
        assert np.array_equal(xa, xb)
        assert np.all(xa[:, 0] <= ra["ub"])


def test_stop_final_compaction_large_queue(gpu):
    """A stop whose queue holds more than 65,536 entries (live or stale):
    the final compaction runs in a launch of its own (enqueued by the host
    when no later step is in flight).  Levy d = 1000 (windowed, the
    reference's box, 10 samples; the queue peaks near 5e7 entries): the
    solve's queue size equals the live boxes the library reports, every one
    with lb <= ub, and the eager-compaction solve's."""
    from gpugo_amd import families as F
    d = 1000
    s = F.spec("levy", d)
    obj = F.objective_loops("levy", d)
    m = gpu.get_module(obj, roundings=("exact", "fast"), windowed=True)
    kw = dict(tol=1e-6, grid=(3, 12), k_select=1, n_samples=10, init_ub=s.extra["init_ub"], module=m,
              windowed=True, max_iters=20000, max_queue=1 << 28)
    with gpu.Problem(obj, np.zeros(0), s.lo, s.hi, **kw) as a:
        ra = a.solve()
        lo, hi, lb, n = a.boxes(4096)
    with gpu.Problem(obj, np.zeros(0), s.lo, s.hi, eager_compact=True, **kw) as b:
        rb = b.solve()
    assert ra["status_name"] == rb["status_name"] == "certified"
    assert ra["queue_peak"] > 65536, ra["queue_peak"]
    assert n == ra["queue_size"] == rb["queue_size"]
    assert np.all(lb <= ra["ub"]) and np.all(lo <= hi)
    for key in ("lb", "ub", "width", "iterations", "boxes_bounded"):
        assert ra[key] == rb[key], (key, ra[key], rb[key])


def test_unknown_flag_rejected(gpu):
    """A description with a flag bit the header does not define is refused
    with GPUGO_E_INVALID and a message; the eager-compaction, no-tables and
    no-band bits are known."""
    import ctypes
    from gpugo_amd import _lib
    spec = pb.sepcos2()
    with gpu_problem(gpu, spec, tol=1e-6, grid=(2, 2), k_select=1, eager_compact=True) as p:
        lib = _lib.load()
        d2 = _lib.ProblemDesc.from_buffer_copy(p.desc)
        d2.flags = 128
        h = ctypes.c_void_p()
        assert lib.gpugo_problem_create(p.module.handle, ctypes.byref(d2), ctypes.byref(h)) == _lib.E_INVALID
        assert b"unknown flags" in lib.gpugo_last_error()
        for f in (_lib.FLAG_NO_TABLES, _lib.FLAG_NO_BAND):
            d2.flags = f
            assert lib.gpugo_problem_create(p.module.handle, ctypes.byref(d2), ctypes.byref(h)) == _lib.OK
            lib.gpugo_problem_destroy(h)
        assert p.solve()["status_name"] == "certified"


LAZY_KEYS = ("lb", "ub", "width", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt", "status",
             "queue_size", "n_out")


@pytest.mark.parametrize("spec,grid,kw", [
    (pb.sepcos2(), (64, 2), {}),
    (pb.sepcos2(), (64, 2), {"max_queue": 1 << 16}),   # compaction every 2 steps
    (pb.chain10(), (2, 5), {"rounding": "exact"}),
    (pb.rosenbrock(4), (4, 3), {"rounding": "exact", "mean_value": True}),
    (pb.rastrigin(2), (4, 2), {"tol_x": 1e-3, "tol": 0.0}),  # x-width outputs
    (pb.chain10(), (2, 5), {"max_parents": 64}),        # parent records collected
], ids=lambda v: getattr(v, "name", None))
def test_lazy_queue_matches_eager(gpu, spec, grid, kw):
    """One box per step: the default queue (compacted every few steps,
    per-chunk minimum keys between, consumed and ruled-out entries left in
    place) selects the same boxes as compaction after every step, so the
    solve is the same; its queue_peak also counts the stale entries."""
    kw = dict(kw)
    tol = kw.pop("tol", 1e-6)
    out = []
    for eager in (False, True):
        with gpu_problem(gpu, spec, tol=tol, grid=grid, k_select=1, max_iters=20000, eager_compact=eager,
                         **kw) as p:
            try:
                out.append(p.solve())
            except GpugoError as e:
                out.append({"error": str(e)})
    a, b = out
    assert ("error" in a) == ("error" in b), (a, b)
    if "error" in a:
        assert a["error"] == b["error"]
        return
    assert a["status_name"] in ("certified", "x_converged"), a
    assert a["iterations"] > 2
    for key in LAZY_KEYS:
        assert a[key] == b[key], (key, a[key], b[key])
    assert a["queue_peak"] >= b["queue_peak"]


BAND_KEYS = ("lb", "ub", "width", "iterations", "boxes_bounded", "queue_size", "queue_peak", "pruned_value",
             "pruned_kkt", "status")


@pytest.mark.parametrize("spec,grid,k,kw,big", [
    (pb.chain10(), (3, 3), 64, {}, False),
    (pb.chain10(), (2, 5), 256, {"mean_value": True}, False),
    (pb.batch_problem(6), (3, 3), 64, {}, True),  # d = 20: 4475 steps, queue up to 1.1e5 entries
    (pb.batch_problem(6), (3, 3), 1024, {}, True),  # 27,648 appends a step: the band outgrows its 16384 entries
    (pb.batch_problem(6), (3, 3), 64, {"split": 4, "persistent": True}, True),
], ids=lambda v: getattr(v, "name", str(v)))
def test_band_select_matches_full_pass(gpu, spec, grid, k, kw, big):
    """k_select > 1: the select takes the K smallest boxes from the priority
    band the compaction keeps (entries below a threshold) while it holds at
    least K, instead of passing over the whole queue; queues of at most 32768
    entries keep no band.  Same selections, so the same results after every
    number of steps (the plain select between step calls, the fused one
    within them), after a reset, and the same live boxes, as with the band
    switched off (GPUGO_FLAG_NO_BAND)."""
    obj = pb.objective(spec.family, spec.d)
    if kw.get("split"):
        kw = dict(kw, module=gpu.get_module(obj, roundings=("fast",), split=kw["split"]))
    runs = []
    for band in (True, False):
        out = []
        with gpu.Problem(obj, spec.params, spec.lo, spec.hi, tol=spec.tol, grid=grid, k_select=k, band=band,
                         max_iters=20000, **kw) as p:
            for n in (1, 2, 5, 17):
                out.append(p.step(n))
            lo, hi, lb, nb = p.boxes(1 << 18)
            if nb <= 1 << 18:  # the live boxes as a set (queue order is not fixed)
                order = np.lexsort((lb, *hi.T, *lo.T))
                out.append((nb, lo[order], hi[order], lb[order]))
            else:
                out.append((nb,))
            out.append(p.solve())
            p.reset()
            out.append(p.solve())
        runs.append(out)
    a, b = runs
    assert a[-1]["status_name"] == "certified", a[-1]
    if big:  # a queue above 32768 entries: the band holds the entries below a threshold
        assert a[5]["queue_peak"] > 32768, a[5]["queue_peak"]
    for ra, rb in zip(a[:4] + a[5:], b[:4] + b[5:]):
        for key in BAND_KEYS:
            assert ra[key] == rb[key], (key, ra[key], rb[key])
    assert a[4][0] == b[4][0]
    for x, y in zip(a[4][1:], b[4][1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("spec,grid,k", [
    (pb.sepcos2(), (64, 2), 1),       # one box per step: lazy queue, stop in a lazy tail
    (pb.sepcos2(), (2, 2), 64),
    (pb.rosenbrock(4), (2, 4), 64),   # a few hundred steps: the stop is seen with a later batch in flight
    (pb.ackley(4), (3, 4), 16),
    (pb.chain10(), (2, 5), 256),
], ids=lambda v: getattr(v, "name", str(v)))
def test_pinned_final_state_equals_readback(gpu, spec, grid, k):
    """A solve that stops in a fused compaction + select gets its final state
    from the pinned copy the device writes before the status word; the
    static-arena solve (no pinned words) reads it back with a copy.  Every
    result field is the same."""
    from tests.test_zz_no_alloc import KEYS as SKEYS, desc_for, solve_created, solve_static
    keep = []
    m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=("fast",))
    desc = desc_for(spec, grid, k, 1 << 20, 1 << 14, keep)
    rc, r = solve_static(m, desc)
    assert rc == 0, rc
    c = solve_created(m, desc)
    assert c["status_name"] == "certified", c
    for key in SKEYS:
        assert r[key] == c[key], (key, r[key], c[key])
