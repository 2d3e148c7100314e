"""The oracle's RT-HDIST restatement (oracle/rthdist.py) on the CPU: the
launch loop with its growing radius gives the exact distance on ordinary
inputs (brute force), including A points beside, around and far from B,
and a hand-worked case."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle, rthdist


@pytest.mark.parametrize("nd", [2, 3])
def test_rt_hdist_restatement_is_exact_on_random_inputs(nd):
    rng = np.random.default_rng(40 + nd)
    for _ in range(25):
        b = rng.random((int(rng.integers(2, 60)), nd)).astype(np.float32)
        scale, shift = rng.choice([1.0, 2.0, 4.0]), rng.choice([0.0, 0.5, 1.5])
        a = (rng.random((int(rng.integers(1, 40)), nd)) * scale - shift).astype(np.float32)
        e = oracle.brute_directed(a, b)
        assert rthdist.rt_hdist(a, b) == (e.cmax2_bits, e.argmax)


def test_rt_hdist_restatement_by_hand():
    # B on the x axis from 0 to 1 (y widened to 1e-5): 128 cells along x,
    # one along y. A = (2, 0): centres 0.75 apart, first radius 68 cell
    # diagonals; A quantizes to x = 255, the nearest non-empty cell is 127,
    # reached 23 launches later; three launches on, only that cell is within
    # the radius, and its point (1, 0) gives cmax2 = 1.
    b = np.array([[0, 0], [1, 0]], np.float32)
    a = np.array([[2, 0]], np.float32)
    assert rthdist.rt_hdist(a, b) == (0x3F800000, 0)
