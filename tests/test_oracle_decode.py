"""Pins the oracle's subregion decode.

Reference: selectVarInSubRegion / completeRegion
(GPUGO_1.1/1-Ackley/ackley50.py:482-594, :820-948): coordinate
cycleIdx + k of the selected region takes base-`wid` digit k of the thread
index `pos`, least significant digit first, and the child bounds are
L + (U-L)/wid * digit and L + (U-L)/wid * (digit+1).  Checked by a
known-answer case written from that description, by exhaustive tiling of
the parent (every point of the parent lies in exactly the children the
digits predict, adjacent children share endpoints, the top child ends at
U) and by the bijection pos <-> digit tuple.
"""
import itertools

import numpy as np


def test_known_answer_lsb_first(oracle):
    # wid=4, window of 2 starting at coordinate 0: pos=6 -> digits (2, 1)
    plo, phi = np.array([0.0, 10.0, 5.0]), np.array([4.0, 14.0, 6.0])
    clo, chi = oracle.decode(plo, phi, win=0, wid=4, cycle=2, pos=[6])
    assert clo[0].tolist() == [2.0, 11.0, 5.0]
    assert chi[0].tolist() == [3.0, 12.0, 6.0]
    # window starting at coordinate 2 wraps to coordinate 0
    clo, chi = oracle.decode(plo, phi, win=2, wid=4, cycle=2, pos=[6])
    assert clo[0].tolist() == [1.0, 10.0, 5.5]
    assert chi[0].tolist() == [2.0, 14.0, 5.75]


def test_children_tile_parent_exhaustive(oracle):
    rng = np.random.default_rng(3)
    for wid, cycle, d, win in [(2, 3, 3, 0), (3, 2, 4, 3), (4, 3, 5, 1), (5, 2, 2, 1), (7, 1, 3, 2)]:
        plo = rng.uniform(-10, 0, d)
        phi = plo + rng.uniform(0.1, 7, d)
        c = min(cycle, d)
        n = wid ** c
        clo, chi = oracle.decode(plo, phi, win, wid, cycle, np.arange(n))
        window = [(win + k) % d for k in range(c)]
        seen = set()
        for pos in range(n):
            digits = tuple((pos // wid ** k) % wid for k in range(c))
            seen.add(digits)
            for j in range(d):
                if j not in window:
                    assert clo[pos, j] == plo[j] and chi[pos, j] == phi[j]
        assert len(seen) == n
        # per window coordinate: wid distinct pieces, contiguous, covering [L, U]
        for k, j in enumerate(window):
            pieces = sorted({(clo[p, j], chi[p, j]) for p in range(n)})
            assert len(pieces) == wid
            assert pieces[0][0] == plo[j] and pieces[-1][1] == phi[j]
            for (a0, a1), (b0, b1) in zip(pieces, pieces[1:]):
                assert a1 == b0 and a0 < a1
        # every grid point lands in the child the digits predict
        for _ in range(50):
            x = plo + rng.uniform(0, 1, d) * (phi - plo)
            inside = [p for p in range(n) if np.all(clo[p] <= x) and np.all(x <= chi[p])]
            assert 1 <= len(inside) <= 2 ** c
