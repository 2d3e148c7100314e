"""Batched mode: every pair's (cmax2 bits, arg-max, NN) equals the CPU
oracle's on that pair, across a graph replay."""
from __future__ import annotations

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from oracle import oracle

pytestmark = pytest.mark.gpu


def _check(ctx, a, oa, b, ob, n_dims):
    import torch

    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    doa = torch.from_numpy(oa.astype(np.int64)).cuda()
    dob = torch.from_numpy(ob.astype(np.int64)).cuda()
    first = hd.batch_directed(ctx, da, doa, db, dob)
    second = hd.batch_directed(ctx, da, doa, db, dob)  # graph replay
    for p in range(len(oa) - 1):
        pa = a[oa[p]:oa[p + 1]]
        pb = b[ob[p]:ob[p + 1]]
        o = oracle.eb_directed(pa, pb, n_threads=4)
        want = (o.cmax2_bits, o.argmax, o.nn)
        assert (first[p].cmax2_bits, first[p].argmax, first[p].nn) == want, p
        assert (second[p].cmax2_bits, second[p].argmax, second[p].nn) == want, p
    assert first[0].time_ms > 0 and first[0].evals > 0


def test_batch_of_gaussian_blobs(ctx):
    bt = hd.gen.batch(9, 96)
    _check(ctx, bt.a, bt.offs_a, bt.b, bt.offs_b, 2)


def test_batch_ragged_3d_and_tiny_pairs(ctx):
    rng = np.random.default_rng(4)
    sizes_a = np.array([1, 2, 63, 64, 65, 1000, 3, 40000, 7], np.int64)
    sizes_b = np.array([1, 70, 1, 64, 129, 999, 5000, 31, 20000], np.int64)
    oa = np.concatenate([[0], np.cumsum(sizes_a)])
    ob = np.concatenate([[0], np.cumsum(sizes_b)])
    a = rng.random((oa[-1], 3)).astype(np.float32)
    b = rng.random((ob[-1], 3)).astype(np.float32)
    _check(ctx, a, oa, b, ob, 3)
