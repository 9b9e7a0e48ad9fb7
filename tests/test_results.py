"""Root merging of output boxes and the result record (CPU)."""
import json

import numpy as np

from gpugo_amd import results as R


def test_connectivity_rule():
    a_lo, a_hi = np.array([0.0, 0.0]), np.array([1.0, 1.0])
    # inside, touching, within tol of touching, and apart in one coordinate
    assert R.connected(a_lo, a_hi, np.array([0.2, 0.2]), np.array([0.5, 0.5]), 1e-6)
    assert R.connected(a_lo, a_hi, np.array([1.0, 0.5]), np.array([2.0, 0.7]), 1e-6)
    assert R.connected(a_lo, a_hi, np.array([1.0 + 5e-7, 0.5]), np.array([2.0, 0.7]), 1e-6)
    assert not R.connected(a_lo, a_hi, np.array([1.1, 0.5]), np.array([2.0, 0.7]), 1e-6)
    assert not R.connected(a_lo, a_hi, np.array([0.5, 1.5]), np.array([0.7, 2.0]), 1e-6)


def test_greedy_combine_like_reference_order():
    # boxes along a line merge into one root; a far box makes a second root
    lo = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [10.0, 10.0], [3.0, 0.0]])
    hi = lo + 1.0
    rlo, rhi, owner = R.combine_roots(lo, hi, 1e-9)
    assert len(rlo) == 2
    assert np.array_equal(rlo[0], [0.0, 0.0]) and np.array_equal(rhi[0], [4.0, 1.0])
    assert list(owner) == [0, 0, 0, 1, 0]


def test_merge_joins_roots_connected_later():
    # box 2 connects roots 0 and 1, which the greedy pass keeps apart
    lo = np.array([[0.0, 0.0], [2.0, 0.0], [1.0, 0.0]])
    hi = lo + 1.0
    rlo, rhi, owner = R.combine_roots(lo, hi, 1e-9)
    assert len(rlo) == 2
    mlo, mhi, mown = R.merge_roots(rlo, rhi, owner, 1e-9)
    assert len(mlo) == 1 and np.array_equal(mhi[0], [3.0, 1.0]) and set(mown) == {0}


def test_record_and_save(tmp_path):
    res = {"lb": -1e-9, "ub": 1e-9, "width": 2e-9, "status_name": "certified", "iterations": 7,
           "boxes_bounded": 100, "ms": 1.0}
    lo = np.array([[-1e-3, -1e-3], [0.0, -1e-3]])
    hi = lo + 1e-3
    rec = R.record(res, (lo, hi, np.zeros(2)), tol=1e-6, meta={"objective": "test"})
    assert rec["status"] == "certified" and len(rec["roots"]["lo"]) == 1
    R.save(rec, str(tmp_path / "r.json"))
    back = json.load(open(tmp_path / "r.json"))
    assert back["roots"]["box_root"] == [0, 0]
    R.save(rec, str(tmp_path / "r.npz"))
    z = np.load(tmp_path / "r.npz")
    assert z["roots_lo"].shape == (1, 2) and float(z["lb"]) == -1e-9
