"""The reference's benchmark objectives on the GPU (gpugo_amd/families.py).

These objectives have no hand-written oracle code lists; they are pinned by
mathematics: value and gradient enclosures on random boxes must contain the
256-bit values (independent formulas in tests/test_families.py) at random
points, in both rounding modes, and certified solves must contain the
closed-form global minimum f*.
"""
import mpmath as mp
import numpy as np
import pytest

from gpugo_amd import families as F
from tests.test_families import f_ref

pytestmark = pytest.mark.gpu
SOLVE_D = {"ackley": 3, "belegundu": 2, "breiman": 3, "fu": 2, "griewank": 2, "levy": 3, "rastrigin": 3,
           "salomon": 2, "styblinski": 3, "zabinsky": 3, "rosenbrock": 3}


@pytest.mark.parametrize("name", F.names())
@pytest.mark.parametrize("rounding", ["exact", "fast"])
def test_enclosures_contain_truth(gpu, name, rounding):
    d = 3
    fam = F.REFERENCE[name]
    m = gpu.get_module(F.objective(name, d), roundings=("exact", "fast"))
    rng = np.random.default_rng(11)
    n = 48
    span = fam.hi - fam.lo
    lo = fam.lo + rng.uniform(0, 1, (n, d)) * span
    w = rng.choice([1e-7, 1e-3, 0.05, 0.5], (n, 1)) * span
    hi = np.minimum(lo + rng.uniform(0, 1, (n, d)) * w, fam.hi)
    dlo, dhi = np.full(d, fam.lo), np.full(d, fam.hi)
    g = m.eval_boxes(np.zeros(0), lo, hi, dlo, dhi, 1, rounding=rounding)
    mp.mp.prec = 256
    for i in range(n):
        for _ in range(3):
            x = lo[i] + rng.uniform(0, 1, d) * (hi[i] - lo[i])
            t = f_ref(name, x)
            assert g["flo"][i] <= t <= g["fhi"][i], (name, i, g["flo"][i], t, g["fhi"][i])
            assert g["ub"][i] >= g["flo"][i]
        if i % 4 == 0:
            x = lo[i] + 0.5 * (hi[i] - lo[i])
            for j in range(d):
                def fj(s, j=j):
                    y = [mp.mpf(float(v)) for v in x]
                    y[j] = s
                    return f_ref(name, y)
                dj = mp.diff(fj, mp.mpf(float(x[j])))
                assert g["glo"][i, j] - 1e-30 <= dj <= g["ghi"][i, j] + 1e-30, (name, i, j)


@pytest.mark.parametrize("name", F.names())
def test_certified_minimum(gpu, name):
    d = SOLVE_D[name]
    s = F.spec(name, d, tol=1e-6)
    obj = F.objective(name, d)
    m = gpu.get_module(obj, roundings=("exact", "fast"))
    res = gpu.minimize(obj, s.lo, s.hi, tol=1e-6, grid=(3, d), k_select=64, init_ub=s.extra["init_ub"],
                       max_iters=20000, max_queue=1 << 22, module=m, rounding="fast")
    assert res["status_name"] == "certified", (name, res)
    assert res["lb"] <= s.f_star <= res["ub"], (name, res["lb"], s.f_star, res["ub"])
    assert res["ub"] - res["lb"] <= 1e-6
