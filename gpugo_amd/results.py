"""Result records: output boxes merged into roots, and a structured export.

The reference prints its output regions and, in the Rosenbrock driver,
merges them into "roots" with combine/connectivity
(GPUGO_1.1/11-Rosenbrock/rosenbrock50.py:926-1049, call :1439): each new
output region is merged into the first existing root it is connected to
(the root becomes the coordinate-wise hull of both), or becomes a new root.

Connectivity here follows the rule that code's docstring states: two boxes
are connected when, in every coordinate, their intervals overlap or the gap
between them is below the tolerance (a box inside the root is connected).
The code itself compares the new region's own width instead of the gap on
the upper side (:1038-1040); that slip is not reproduced.  As in the
reference, merging is greedy in output order, so a root can absorb boxes
that later connect two roots without merging those roots; `merge_roots`
then joins connected roots until none are left.
"""
from __future__ import annotations

import json
from typing import Dict, List, Optional, Tuple

import numpy as np


def connected(alo, ahi, blo, bhi, tol: float) -> bool:
    """Every coordinate overlaps or is within `tol` of touching."""
    gap = np.maximum(blo - ahi, alo - bhi)
    return bool(np.all(gap < tol))


def combine_roots(lo, hi, tol: float) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Greedy merge of boxes (rows of lo/hi) in order, as the reference's
    combine; returns root hulls (lo, hi) and the root index of each box."""
    lo = np.atleast_2d(np.asarray(lo, dtype=np.float64))
    hi = np.atleast_2d(np.asarray(hi, dtype=np.float64))
    rlo: List[np.ndarray] = []
    rhi: List[np.ndarray] = []
    owner = np.empty(len(lo), dtype=np.int64)
    for i in range(len(lo)):
        for r in range(len(rlo)):
            if connected(rlo[r], rhi[r], lo[i], hi[i], tol):
                rlo[r] = np.minimum(rlo[r], lo[i])
                rhi[r] = np.maximum(rhi[r], hi[i])
                owner[i] = r
                break
        else:
            rlo.append(lo[i].copy())
            rhi.append(hi[i].copy())
            owner[i] = len(rlo) - 1
    d = lo.shape[1] if lo.size else 0
    return (np.array(rlo).reshape(-1, d), np.array(rhi).reshape(-1, d), owner)


def merge_roots(rlo, rhi, owner, tol: float):
    """Join roots that are connected to each other until none are."""
    rlo, rhi = [r.copy() for r in rlo], [r.copy() for r in rhi]
    ids = list(range(len(rlo)))
    changed = True
    while changed:
        changed = False
        for a in range(len(rlo)):
            for b in range(a + 1, len(rlo)):
                if connected(rlo[a], rhi[a], rlo[b], rhi[b], tol):
                    rlo[a] = np.minimum(rlo[a], rlo[b])
                    rhi[a] = np.maximum(rhi[a], rhi[b])
                    owner = np.where(owner == ids[b], ids[a], owner)
                    del rlo[b], rhi[b], ids[b]
                    changed = True
                    break
            if changed:
                break
    remap = {old: new for new, old in enumerate(ids)}
    d = len(rlo[0]) if rlo else 0
    return (np.array(rlo).reshape(-1, d), np.array(rhi).reshape(-1, d),
            np.array([remap[o] for o in owner], dtype=np.int64))


def roots(lo, hi, tol: float):
    """Roots of a set of output boxes: greedy combine, then merge."""
    rlo, rhi, owner = combine_roots(lo, hi, tol)
    return merge_roots(rlo, rhi, owner, tol)


def record(res: Dict, boxes: Optional[Tuple] = None, tol: float = 0.0, meta: Optional[Dict] = None) -> Dict:
    """A JSON-ready result: the enclosure [lb, ub], its width and status,
    the counters, and (if boxes are given) the boxes with their lower
    bounds and the roots they merge into.  Replaces the reference's
    printed output (1-Ackley/ackley50.py:1335-1338, :1386-1393)."""
    keys = ("lb", "ub", "width", "status_name", "iterations", "boxes_bounded", "pruned_value", "pruned_kkt",
            "queue_size", "queue_peak", "n_out", "ms")
    out = {k: res[k] for k in keys if k in res}
    out["status"] = out.pop("status_name", res.get("status"))
    if boxes is not None:
        blo, bhi, blb = boxes
        out["boxes"] = {"lo": np.asarray(blo).tolist(), "hi": np.asarray(bhi).tolist(),
                        "lb": np.asarray(blb).tolist()}
        if len(blo) and tol > 0:
            rlo, rhi, owner = roots(blo, bhi, tol)
            out["roots"] = {"lo": rlo.tolist(), "hi": rhi.tolist(), "box_root": owner.tolist(), "tol": tol}
    if meta:
        out["meta"] = meta
    return out


def save(rec: Dict, path: str):
    """Write a record as JSON (.json) or NumPy arrays (.npz)."""
    if path.endswith(".npz"):
        arrays = {k: np.asarray(v) for k, v in rec.items() if isinstance(v, (int, float, str))}
        for name in ("boxes", "roots"):
            for k, v in rec.get(name, {}).items():
                if isinstance(v, list):
                    arrays[f"{name}_{k}"] = np.asarray(v)
        np.savez(path, **arrays)
    else:
        with open(path, "w") as fh:
            json.dump(rec, fh, indent=1)
