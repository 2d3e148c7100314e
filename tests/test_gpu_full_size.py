"""Every benchmark workload at its full size against the oracle's known
answers (tests/golden/full_size.json, written by make_full_golden.py): the
reference checks each run against its CPU early break
(run_hausdorff_distance.cu:366-388); here every GPU variant, the learned
order and the batch must give the oracle's cmax2 bits, arg-max and NN.

The inputs are regenerated from their seeds and their SHA-256 is compared
first, so a generator change cannot pass for a kernel bug. These are the
sizes at which the size-gated paths run: the sampled first round (2^21
points and more), grouped 1024-step rounds, the long segmented tail and the
learned order with the 2000-step model on 1M labels."""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import hausdorffdistance_amd as hd

pytestmark = pytest.mark.gpu
FULL = json.loads((Path(__file__).resolve().parent / "golden" / "full_size.json").read_text())


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def triple(r):
    return {"cmax2_bits": r.cmax2_bits, "argmax": r.argmax, "nn": r.nn}


def device(*arrays):
    import torch

    out = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]
    torch.cuda.synchronize()
    return out


def run_directed(ctx, name, da, db, exp, **kw):
    for variant in (hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID):
        print(f"[full] {name} {variant.name} {kw}", flush=True)
        r = hd.directed_device(ctx, variant, da, db, **kw)
        assert triple(r) == exp, (name, variant.name)


def test_uniform_100k_2d(ctx):
    exp = FULL["uniform_100k_2d"]
    a, b = hd.gen.uniform_pair(100_000, 2)
    assert digest(a, b) == exp["digest"]
    da, db = device(a, b)
    run_directed(ctx, "uniform A->B", da, db, exp["ab"])
    run_directed(ctx, "uniform B->A", db, da, exp["ba"])
    for x, y, e in ((da, db, exp["ab"]), (db, da, exp["ba"])):
        assert triple(hd.directed_device(ctx, hd.Variant.NN, x, y)) == e
        assert triple(hd.directed_device(ctx, hd.Variant.EB, x, y, single_pass=True)) == e


def test_gmm_10m_2d(ctx):
    exp = FULL["gmm_10m_2d"]
    a, b = hd.gen.gmm_10m()
    assert digest(a, b) == exp["digest"]
    da, db = device(a, b)
    run_directed(ctx, "gmm10m A->B", da, db, exp["ab"])
    run_directed(ctx, "gmm10m B->A", db, da, exp["ba"])
    print("[full] gmm10m NN and single pass", flush=True)
    assert triple(hd.directed_device(ctx, hd.Variant.NN, da, db)) == exp["ab"]
    assert triple(hd.directed_device(ctx, hd.Variant.EB, da, db, single_pass=True)) == exp["ab"]
    print("[full] gmm10m learned order (2000 steps, 1M labels)", flush=True)
    model = hd.Model.train(ctx, da, db, steps=2000, batch=4096, n_samples=1_000_000)
    r = hd.directed_device(ctx, hd.Variant.EB, da, db, learned_order=True, model=model)
    assert triple(r) == exp["ab"]
    r = hd.directed_device(ctx, hd.Variant.EB, da, db, learned_order=True, model=model,
                           single_pass=True)
    assert triple(r) == exp["ab"]
    # Symmetric H through the run interface, with the host check.
    print("[full] gmm10m symmetric hybrid with -check", flush=True)
    cfg = hd.RunConfig(variant=hd.Variant.HYBRID, n_dims=2, repeat=1, symmetric=True,
                       check=True, stats_n_points_cell=0)
    res = hd.run_hausdorff_distance(cfg, da, db, ctx=ctx)
    assert triple(res.ab) == exp["ab"] and triple(res.ba) == exp["ba"]
    assert res.check_pass == 1
    win = 1 if exp["ba"]["cmax2_bits"] > exp["ab"]["cmax2_bits"] else 0
    assert res.direction == win


def test_sphere_20m_15m_3d(ctx):
    exp = FULL["sphere_20m_15m_3d"]
    a, b = hd.gen.sphere_pair()
    assert digest(a, b) == exp["digest"]
    da, db = device(a, b)
    run_directed(ctx, "sphere A->B", da, db, exp["ab"])
    run_directed(ctx, "sphere B->A", db, da, exp["ba"])


def test_batch_4096_2d(ctx):
    exp = FULL["batch_4096_2d"]
    bt = hd.gen.batch(9, 4096)
    assert digest(bt.a, bt.b, bt.offs_a, bt.offs_b) == exp["digest"]
    da, db, doa, dob = device(bt.a, bt.b, bt.offs_a.astype(np.int64),
                              bt.offs_b.astype(np.int64))
    for call in range(2):  # the first call tunes and captures, the second replays
        print(f"[full] batch call {call}", flush=True)
        res = hd.batch_directed(ctx, da, doa, db, dob)
        got = [[r.cmax2_bits, r.argmax, r.nn] for r in res]
        bad = [p for p in range(4096) if got[p] != exp["pairs"][p]]
        assert not bad, (len(bad), bad[:5])


def test_gmm_50m_256_2d(ctx):
    exp = FULL["gmm_50m_256_2d"]
    a, b = hd.gen.gmm_50m_256()
    assert digest(a, b) == exp["digest"]
    da, db = device(a, b)
    del a, b
    for variant in (hd.Variant.GRID, hd.Variant.HYBRID):
        print(f"[full] gmm50m {variant.name}", flush=True)
        assert triple(hd.directed_device(ctx, variant, da, db)) == exp["ab"]
    print("[full] gmm50m EB plain", flush=True)
    plain = hd.directed_device(ctx, hd.Variant.EB, da, db)
    assert triple(plain) == exp["ab"]
    print("[full] gmm50m learned order (2000 steps, 1M labels)", flush=True)
    model = hd.Model.train(ctx, da, db, steps=2000, batch=4096, n_samples=1_000_000)
    learned = hd.directed_device(ctx, hd.Variant.EB, da, db, learned_order=True, model=model)
    assert triple(learned) == exp["ab"]
    print(f"[full] gmm50m evals plain {plain.evals} learned {learned.evals}", flush=True)
