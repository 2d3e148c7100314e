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


def test_batch_offsets_rewritten_in_place(ctx):
    """The same device tensors holding another ragged split: the cached plan
    must not be reused (its pair boundaries would be stale)."""
    import torch

    rng = np.random.default_rng(11)
    a = rng.random((30_000, 2)).astype(np.float32)
    b = rng.random((24_000, 2)).astype(np.float32)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()

    def split(n_pts, n_pairs, seed):
        cuts = np.sort(np.random.default_rng(seed).choice(np.arange(1, n_pts), n_pairs - 1,
                                                          replace=False))
        return np.concatenate([[0], cuts, [n_pts]]).astype(np.int64)

    oa1, ob1 = split(len(a), 40, 1), split(len(b), 40, 2)
    oa2, ob2 = split(len(a) - 500, 40, 3), split(len(b) - 7, 40, 4)
    doa = torch.from_numpy(oa1).cuda()
    dob = torch.from_numpy(ob1).cuda()
    for oa, ob in ((oa1, ob1), (oa2, ob2), (oa1, ob1)):
        doa.copy_(torch.from_numpy(oa))
        dob.copy_(torch.from_numpy(ob))
        torch.cuda.synchronize()
        res = hd.batch_directed(ctx, da, doa, db, dob)
        for p in range(40):
            o = oracle.eb_directed(a[oa[p]:oa[p + 1]], b[ob[p]:ob[p + 1]], n_threads=4)
            assert (res[p].cmax2_bits, res[p].argmax, res[p].nn) == (
                o.cmax2_bits, o.argmax, o.nn), p


def test_batch_offsets_are_validated(ctx):
    import torch

    a = torch.rand((100, 2), device="cuda")
    b = torch.rand((80, 2), device="cuda")
    good_a = torch.tensor([0, 50, 100], dtype=torch.int64, device="cuda")
    good_b = torch.tensor([0, 40, 80], dtype=torch.int64, device="cuda")
    with pytest.raises(hd.HDError):  # past the end of B
        hd.batch_directed(ctx, a, good_a, b, torch.tensor([0, 40, 81], dtype=torch.int64,
                                                           device="cuda"))
    with pytest.raises(hd.HDError):  # wrong dtype
        hd.batch_directed(ctx, a, good_a.to(torch.int32), b, good_b)
    with pytest.raises(hd.HDError):  # host offsets
        hd.batch_directed(ctx, a, good_a.cpu(), b, good_b)
    with pytest.raises(hd.HDError):  # decreasing offsets (checked by the library)
        hd.batch_directed(ctx, a, torch.tensor([0, 60, 50], dtype=torch.int64, device="cuda"),
                          b, good_b)
    res = hd.batch_directed(ctx, a, good_a, b, good_b)
    assert len(res) == 2


def test_batch_b_order_on_spatially_sorted_input(ctx):
    """Pairs whose B is stored sorted along x (as polygon outlines or sorted
    files are): both visit orders give the oracle's bits; in input order a
    scan meets a point's neighbours only when it reaches their stretch of
    the array, the default rotation of B's indices spreads them over the
    scan."""
    import torch

    bt = hd.gen.batch(19, 48)
    b = bt.b.copy()
    for p in range(len(bt.offs_b) - 1):
        s, e = bt.offs_b[p], bt.offs_b[p + 1]
        b[s:e] = b[s:e][np.argsort(b[s:e, 0], kind="stable")]
    da = torch.from_numpy(np.ascontiguousarray(bt.a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    doa = torch.from_numpy(bt.offs_a.astype(np.int64)).cuda()
    dob = torch.from_numpy(bt.offs_b.astype(np.int64)).cuda()
    got = {}
    for shuffle in (True, False):
        r = hd.batch_directed(ctx, da, doa, db, dob, shuffle=shuffle)
        for p in range(len(bt.offs_a) - 1):
            o = oracle.eb_directed(bt.a[bt.offs_a[p]:bt.offs_a[p + 1]],
                                   b[bt.offs_b[p]:bt.offs_b[p + 1]], n_threads=4)
            assert (r[p].cmax2_bits, r[p].argmax, r[p].nn) == \
                (o.cmax2_bits, o.argmax, o.nn), (shuffle, p)
        got[shuffle] = r[0].evals
    print(f"[batch] evaluations, rotated / input order: {got[True]} / {got[False]}", flush=True)
    assert got[True] < 0.8 * got[False]


def _lattice_pair(n: int, n_dims: int, rng):
    """B: one point near the centre of each of n distinct lattice cells
    (side s, offsets below 0.1/s), in random order; A[i] = B[i] moved by less
    than 0.01/s per axis, so the nearest neighbour of A[i] is B[i] alone and
    a B point that the repack drops or duplicates changes h(A->B)."""
    side = int(np.ceil(n ** (1.0 / n_dims)))
    while side ** n_dims < n:
        side += 1
    cells = rng.permutation(side ** n_dims)[:n]
    coords = np.stack([(cells // side ** d) % side for d in range(n_dims)], 1)
    b = ((coords + 0.5 + rng.uniform(-0.1, 0.1, coords.shape)) / side).astype(np.float32)
    a = (b.astype(np.float64) + rng.uniform(-0.01, 0.01, coords.shape) / side).astype(np.float32)
    return a, b


def _lattice_expected(a: np.ndarray, b: np.ndarray):
    """h(A->B) of a lattice pair from its construction: FP32 d2[i] with the
    reference's unfused left-to-right sum; the arg-max is the smallest
    index of the largest d2, its nearest neighbour B at the same index."""
    d2 = np.zeros(len(a), np.float32)
    for k in range(a.shape[1]):
        t = a[:, k] - b[:, k]
        d2 = d2 + t * t
    i = int(np.flatnonzero(d2 == d2.max())[0])
    return int(d2.view(np.uint32)[i]), i, i


@pytest.mark.parametrize("n_dims", [2, 3])
def test_batch_repack_visits_every_b_point(ctx, n_dims):
    """The repack's visit orders (groups of 8 in a golden-ratio rotation,
    and input order) are permutations of each pair's B: sizes below one
    group, whole groups, remainders, group counts with small factors
    (1001 = 7 x 11 x 13, 1024) and a prime, each pair against the answer
    its construction fixes (brute force confirms it on the small pairs)."""
    import torch

    rng = np.random.default_rng(11 + n_dims)
    sizes = [1, 2, 7, 8, 9, 15, 16, 17, 64, 100, 257, 1000, 4097,
             8 * 1001, 8 * 1024, 8 * 3000 + 5, 30011]
    pairs = [_lattice_pair(n, n_dims, rng) for n in sizes]
    want = [_lattice_expected(a, b) for a, b in pairs]
    for (a, b), w in zip(pairs, want):
        if len(a) <= 1000:
            o = oracle.brute_directed(a, b)
            assert (o.cmax2_bits, o.argmax, o.nn) == w
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    da = torch.from_numpy(np.concatenate([p[0] for p in pairs])).cuda()
    db = torch.from_numpy(np.concatenate([p[1] for p in pairs])).cuda()
    doffs = torch.from_numpy(offs).cuda()
    print(f"[batch] lattice {n_dims}D, {len(sizes)} pairs", flush=True)
    for shuffle in (True, False):
        r = hd.batch_directed(ctx, da, doffs, db, doffs.clone(), shuffle=shuffle)
        for p, w in enumerate(want):
            assert (r[p].cmax2_bits, r[p].argmax, r[p].nn) == w, (shuffle, sizes[p])
