"""gpugo_solve_static: a whole solve whose device memory is the library's
static arena (a file-scope device array, 128 MiB per GPU), for callers that
cannot allocate.  Bar: through the C ABI, two calls in a row (the second on
a smaller problem, so the arena holds the first call's data) give results
identical to gpugo_problem_create + gpugo_solve with the same capacities,
field by field; a problem larger than the arena is refused with
GPUGO_E_CAPACITY before any GPU work."""
import ctypes

import numpy as np
import pytest

from gpugo_amd import _lib, problems as pb
from gpugo_amd._lib import f64, ptr

pytestmark = pytest.mark.gpu
KEYS = ("lb", "ub", "width", "status", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt",
        "queue_size", "queue_peak", "n_out")


def desc_for(spec, grid, k, max_queue, max_parents, keep):
    lo, hi, params = f64(spec.lo), f64(spec.hi), f64(spec.params)
    keep += [lo, hi, params]
    return _lib.ProblemDesc(d=spec.d, wid=grid[0], cycle=grid[1], k_select=k, n_samples=1, precision=0,
                            rounding=_lib.ROUND["fast"], flags=0, lo=ptr(lo), hi=ptr(hi), params=ptr(params),
                            n_params=params.size, tol_f=spec.tol, tol_x=0.0, init_ub=float("inf"),
                            max_queue=max_queue, max_parents=max_parents, max_iters=100000)


def solve_static(module, desc):
    res = _lib.Result()
    rc = _lib.load().gpugo_solve_static(module.handle, ctypes.byref(desc), None, ctypes.byref(res))
    return rc, res.as_dict()


def solve_created(module, desc):
    lib = _lib.load()
    h = ctypes.c_void_p()
    _lib.check(lib.gpugo_problem_create(module.handle, ctypes.byref(desc), ctypes.byref(h)))
    res = _lib.Result()
    try:
        _lib.check(lib.gpugo_solve(h, ctypes.byref(res)))
    finally:
        lib.gpugo_problem_destroy(h)
    return res.as_dict()


CASES = [
    # (spec, grid, k_select, max_queue, max_parents): the second is smaller
    (pb.chain10(), (2, 5), 256, 1 << 20, 1 << 14),
    (pb.sepcos2(), (64, 2), 1, 1 << 16, 1 << 10),
]


def test_static_solve_equals_created(gpu):
    keep = []
    out = []
    for spec, grid, k, mq, mp in CASES:
        m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=("fast",))
        desc = desc_for(spec, grid, k, mq, mp, keep)
        print(f"# static solve {spec.name} grid={grid} K={k}", flush=True)
        rc, r = solve_static(m, desc)
        assert rc == 0, _lib.load().gpugo_last_error()
        c = solve_created(m, desc)
        for key in KEYS:
            assert r[key] == c[key], (spec.name, key, r[key], c[key])
        assert r["status_name"] == "certified"
        out.append(r)
    # the same first problem again after the smaller one: still identical
    spec, grid, k, mq, mp = CASES[0]
    m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=("fast",))
    rc, r = solve_static(m, desc_for(spec, grid, k, mq, mp, keep))
    assert rc == 0 and all(r[key] == out[0][key] for key in KEYS)


def test_static_solve_too_large_is_refused(gpu):
    keep = []
    spec = pb.chain10()
    m = gpu.get_module(pb.objective(spec.family, spec.d), roundings=("fast",))
    rc, _ = solve_static(m, desc_for(spec, (2, 5), 256, 1 << 26, 1 << 14, keep))
    assert rc == _lib.E_CAPACITY
    assert b"static arena" in _lib.load().gpugo_last_error()
