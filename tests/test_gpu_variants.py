"""GPU parity of the index-based variants (grid, LBVH, hybrid, the
nearest-neighbour-search baseline) and of the exact nearest-neighbour entry
point, through the C ABI, against the CPU
oracle: cmax2 bits, arg-max and NN index must be identical."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from oracle import oracle

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"

VARIANTS = [hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID, hd.Variant.NN]


def run(a, b, variant, symmetric=True, ctx=None, **kw):
    cfg = hd.RunConfig(variant=variant, n_dims=a.shape[1], repeat=1, symmetric=symmetric, **kw)
    return hd.run_hausdorff_distance(cfg, a, b, ctx=ctx)


def triple(r):
    return {"cmax2_bits": r.cmax2_bits, "argmax": r.argmax, "nn": r.nn}


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: v.name)
def test_golden_fixtures(ctx, variant):
    cases = np.load(GOLDEN / "cases.npz")
    expected = json.loads((GOLDEN / "expected.json").read_text())
    for name, exp in expected.items():
        a, b = cases[f"{name}__a"], cases[f"{name}__b"]
        print(f"[golden] {variant.name} {name} {a.shape} {b.shape}", flush=True)
        r = run(a, b, variant, ctx=ctx)
        assert triple(r.ab) == exp["ab"], name
        assert triple(r.ba) == exp["ba"], name


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: v.name)
@pytest.mark.parametrize("n_dims", [2, 3])
@pytest.mark.parametrize("na,nb", [(1, 1), (1, 500), (500, 1), (63, 65), (3000, 1000)])
def test_small_sizes(ctx, variant, n_dims, na, nb):
    a = hd.gen.uniform(300 + na, na, n_dims)
    b = hd.gen.uniform(400 + nb, nb, n_dims)
    print(f"[small] {variant.name} {n_dims}D {na}x{nb}", flush=True)
    r = run(a, b, variant, ctx=ctx)
    assert triple(r.ab) == triple(oracle.brute_directed(a, b))
    assert triple(r.ba) == triple(oracle.brute_directed(b, a))


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: v.name)
def test_gmm_200k_2d(ctx, variant):
    a, b = hd.gen.gmm_pair(200_000)
    print(f"[gmm] {variant.name}", flush=True)
    r = run(a, b, variant, ctx=ctx)
    o_ab, o_ba = oracle.eb_symmetric(a, b)
    assert triple(r.ab) == triple(o_ab)
    assert triple(r.ba) == triple(o_ba)


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: v.name)
def test_sphere_3d(ctx, variant):
    a = hd.gen.sphere(7, 80_000, 0.005, 0.0)
    b = hd.gen.sphere(8, 60_000, 0.005, 2.0)
    print(f"[sphere] {variant.name}", flush=True)
    r = run(a, b, variant, ctx=ctx)
    o_ab, o_ba = oracle.eb_symmetric(a, b)
    assert triple(r.ab) == triple(o_ab)
    assert triple(r.ba) == triple(o_ba)


@pytest.mark.parametrize("variant", VARIANTS, ids=lambda v: v.name)
def test_degenerate_and_far_apart(ctx, variant):
    # All B on a line (zero extent in y), A far away and spread out.
    b = np.stack([np.linspace(0, 1, 999, dtype=np.float32),
                  np.zeros(999, np.float32)], 1)
    a = hd.gen.uniform(77, 700, 2) * np.float32(50) + np.float32(10)
    r = run(a, b, variant, ctx=ctx)
    assert triple(r.ab) == triple(oracle.brute_directed(a, b))
    assert triple(r.ba) == triple(oracle.brute_directed(b, a))
    # Duplicated points everywhere (many exact ties).
    d = np.repeat(hd.gen.uniform(78, 300, 2), 4, axis=0)
    e = np.repeat(hd.gen.uniform(79, 200, 2), 3, axis=0)
    r = run(d, e, variant, ctx=ctx)
    assert triple(r.ab) == triple(oracle.brute_directed(d, e))
    assert triple(r.ba) == triple(oracle.brute_directed(e, d))


@pytest.mark.parametrize("index", [hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH],
                         ids=lambda v: v.name)
@pytest.mark.parametrize("n_dims", [2, 3])
def test_nearest_neighbour_entry_point(ctx, index, n_dims):
    import torch

    q = hd.gen.uniform(90, 20_000, n_dims)
    r = hd.gen.uniform(91, 15_000, n_dims)
    # Exact duplicates in the reference set create index ties.
    r[1000:1100] = r[2000:2100]
    q[:50] = r[2000:2050]
    want_d2, want_idx = oracle.nn_all(q, r)
    dq = torch.from_numpy(q).cuda()
    dr = torch.from_numpy(r).cuda()
    out_d2 = torch.empty(q.shape[0], dtype=torch.float32, device="cuda")
    out_idx = torch.empty(q.shape[0], dtype=torch.int32, device="cuda")
    hd.nn_device(ctx, index, dq, dr, out_d2, out_idx)
    got_d2 = out_d2.cpu().numpy()
    got_idx = out_idx.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got_d2.view(np.uint32), want_d2.view(np.uint32))
    np.testing.assert_array_equal(got_idx, want_idx)


def test_hybrid_auto_tune_keeps_bits(ctx):
    """auto_tune picks the strategy, n_points_cell and max_hit by trial on
    the input (the reference predicts them, hausdorff_distance_hybrid.h:
    722-778) and caches the choice under the input's content: the second
    call reuses it, other data at the same addresses is tuned anew, and the
    bits never change."""
    a, b = hd.gen.gmm_pair(300_000)
    o_ab, o_ba = oracle.eb_symmetric(a, b)
    for call in range(2):  # the first call tunes by trial, the second reuses
        r = run(a, b, hd.Variant.HYBRID, ctx=ctx, auto_tune=True, stats_n_points_cell=0)
        assert triple(r.ab) == triple(o_ab)
        assert triple(r.ba) == triple(o_ba)
        rep = r.stats["Running"]["Repeats"][0]
        ch = rep["HybridChoice"]["AB"]
        print(f"[tune] call {call}: {rep['HybridChoice']}", flush=True)
        assert ch["Tuned"] and ch["Variant"] in ("Uniform Grid", "LBVH", "Early Break")
        assert ch["NumPointsPerCell"] in (2, 4, 8, 16, 32)
        assert ch["MaxHit"] in (64, 256, 1024, 4096)
        assert rep["PredictedNPointsCell"] == ch["NumPointsPerCell"]
        assert (ch["Trials"] > 1) == (call == 0)
    # Other points of the same sizes, uploaded to the same context buffers:
    # a new signature, so a new trial.
    a2 = (a * np.float32(0.5)).astype(np.float32)
    r = run(a2, b, hd.Variant.HYBRID, ctx=ctx, auto_tune=True, stats_n_points_cell=0,
            symmetric=False)
    assert r.stats["Running"]["Repeats"][0]["HybridChoice"]["AB"]["Trials"] > 1
    assert triple(r.ab) == triple(oracle.eb_directed(a2, b))
    # Only n_points_cell, from a poor starting value (cells of ~4096
    # points): the grid path, the candidates 2..32 tried, and one of them
    # replaces the flag's value.
    r = run(a, b, hd.Variant.HYBRID, ctx=ctx, auto_tune_n_points_cell=True,
            n_points_cell=4096, stats_n_points_cell=0, symmetric=False)
    rep = r.stats["Running"]["Repeats"][0]
    ch = rep["HybridChoice"]["AB"]
    print(f"[tune] n_points_cell from 4096: {ch}", flush=True)
    assert ch["Variant"] == "Uniform Grid" and ch["MaxHit"] == 256 and ch["Trials"] == 6
    assert rep["PredictedNPointsCell"] in (2, 4, 8, 16, 32)
    assert triple(r.ab) == triple(o_ab)


def test_hybrid_escalates_far_queries(ctx):
    # B's points far from every A cluster force long ring searches; with a
    # small max_hit they are finished by the early-break scan.
    a, b = hd.gen.gmm_pair(60_000)
    o = oracle.eb_directed(b, a)
    r = run(b, a, hd.Variant.HYBRID, ctx=ctx, symmetric=False, max_hit=4)
    assert triple(r.ab) == triple(o)


RTH_CASES = {
    "uniform_2d": lambda: (hd.gen.uniform(101, 2000, 2), hd.gen.uniform(102, 1500, 2)),
    "uniform_3d": lambda: (hd.gen.uniform(103, 900, 3), hd.gen.uniform(104, 700, 3) * np.float32(0.5)),
    # A beside B: the first launches miss and the radius grows many times.
    "separated_2d": lambda: (hd.gen.uniform(105, 300, 2) + np.float32(2.5),
                             hd.gen.uniform(106, 400, 2)),
    # A around B on every side, some points left of and below it (negative
    # offsets truncate toward zero).
    "around_2d": lambda: (hd.gen.uniform(107, 600, 2) * np.float32(3) - np.float32(1),
                          hd.gen.uniform(108, 500, 2)),
    "flat_b_2d": lambda: (hd.gen.uniform(109, 500, 2),
                          np.c_[hd.gen.uniform(110, 400, 2)[:, :1], np.full(400, 0.5, np.float32)]),
    "two_b": lambda: (hd.gen.uniform(111, 50, 2), hd.gen.uniform(112, 2, 2)),
    "one_a": lambda: (hd.gen.uniform(113, 1, 3), hd.gen.uniform(114, 300, 3)),
}


@pytest.mark.parametrize("case", sorted(RTH_CASES))
def test_rt_hdist_matches_its_restatement(ctx, case):
    from oracle import rthdist

    a, b = (np.ascontiguousarray(x, np.float32) for x in RTH_CASES[case]())
    print(f"rt-hdist {case} {a.shape} {b.shape}", flush=True)
    r = run(a, b, hd.Variant.RT_HDIST, symmetric=False, ctx=ctx)
    bits, arg = rthdist.rt_hdist(a, b)
    assert (r.ab.cmax2_bits, r.ab.argmax) == (bits, arg)
    # On these inputs every A point's nearest B point lies within the scan
    # radius, so the baseline returns the exact distance (and the NN of the
    # arg-max is the exact one).
    exact = oracle.brute_directed(a, b)
    assert (bits, arg) == (exact.cmax2_bits, exact.argmax)
    assert r.ab.nn == exact.nn


def sample_counts(ctx, a, b, seed=0, rate=0.01):
    import ctypes as C
    import torch
    from hausdorffdistance_amd import _lib

    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    n = C.c_uint64()
    buf = (C.c_uint32 * (len(a) + 1))()
    rc = _lib.load().hd_hybrid_sample_iterations(
        ctx.handle, a.shape[1], C.c_void_p(da.data_ptr()), len(a), C.c_void_p(db.data_ptr()),
        len(b), seed, rate, buf, len(a) + 1, C.byref(n))
    _lib.check(rc)
    return np.array(buf[:n.value], np.uint32)


def test_hybrid_sampled_eb_only_decision(ctx):
    a = np.ascontiguousarray(hd.gen.uniform(121, 20000, 2))
    b = np.ascontiguousarray(hd.gen.uniform(122, 15000, 2))
    counts = sample_counts(ctx, a, b)
    full = -(-len(b) // 256)
    assert len(counts) == 200  # ceil(20000 * 0.01)
    # counts are the reference's: 32k + 1 at the first check that breaks,
    # or every iteration when none does
    assert all(c == full or (c % 32 == 1 and c < full) for c in counts), counts
    p80 = int(np.sort(counts)[int(len(counts) * 0.8)])
    exact = oracle.brute_directed(a, b)
    for thr, want in ((p80 + 1, hd.Variant.EB), (p80, hd.Variant.GRID),
                      (1, hd.Variant.GRID), (10 ** 6, hd.Variant.EB)):
        r = run(a, b, hd.Variant.HYBRID, symmetric=False, ctx=ctx, eb_only_threshold=thr)
        assert r.ab.variant_used == want, (thr, p80)
        assert (r.ab.cmax2_bits, r.ab.argmax, r.ab.nn) == (
            exact.cmax2_bits, exact.argmax, exact.nn)
    # B at one point: every sampled point is within cmax2 of the first B
    # point, so every count is 1.
    one = np.repeat(np.array([[0.5, 0.5]], np.float32), 3000, axis=0)
    assert set(sample_counts(ctx, a, one, rate=0.05).tolist()) == {1}


@pytest.mark.parametrize("variant", [hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID,
                                     hd.Variant.EB], ids=lambda v: v.name)
def test_flat_set_at_large_coordinates(ctx, variant):
    """B lies on the plane z = 1000 (3D) or the line y = 1000 (2D): the +1e-5
    widening of its zero extent rounds away in FP32, so the grid must size
    itself from the other axes instead of from a zero volume."""
    rng = np.random.default_rng(21)
    for nd in (3, 2):
        b = rng.random((200_000, nd)).astype(np.float32)
        b[:, nd - 1] = np.float32(1000.0)
        a = rng.random((50_000, nd)).astype(np.float32)
        a[:, nd - 1] = (np.float32(1000.0) + rng.normal(0, 0.05, 50_000)).astype(np.float32)
        print(f"[flat] {variant.name} {nd}D", flush=True)
        r = run(a, b, variant, ctx=ctx)
        o_ab, o_ba = oracle.eb_symmetric(a, b)
        assert triple(r.ab) == triple(o_ab)
        assert triple(r.ba) == triple(o_ba)
