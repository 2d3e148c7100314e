"""Learned ordering: the GPU forward/backward matches an FP64 CPU evaluation
of the same BF16-rounded network, training lowers the loss, and the learned
visit order leaves every result bit-identical."""
from __future__ import annotations

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from oracle import oracle

pytestmark = pytest.mark.gpu


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32 to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("n_dims", [2, 3])
def test_gradient_matches_fp64_oracle(ctx, n_dims):
    lib = hd.load()
    n = lib.hd_mlp_n_params(n_dims)
    rng = np.random.default_rng(5 + n_dims)
    params = (rng.standard_normal(n) * 0.3).astype(np.float32)
    x = rng.random((512, n_dims)).astype(np.float32)
    y = rng.standard_normal(512).astype(np.float32)
    loss, grad = hd.mlp_loss_grad(ctx, params, x, y)
    ref_loss, ref_grad = oracle.mlp_loss_grad(bf16_round(params), n_dims, 64, x, y)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    scale = np.abs(ref_grad).max()
    np.testing.assert_allclose(grad, ref_grad, rtol=2e-3, atol=2e-4 * scale)


def oracle_forward(params: np.ndarray, n_dims: int, x: np.ndarray) -> np.ndarray:
    """FP64 network output per row of x: with target 0 and a batch of one,
    the oracle's gradient of b3 (last parameter) is 2 f(x)."""
    out = np.empty(x.shape[0])
    zero = np.zeros(1, np.float32)
    for i in range(x.shape[0]):
        _, g = oracle.mlp_loss_grad(params, n_dims, 64, x[i : i + 1], zero)
        out[i] = 0.5 * g[-1]
    return out


@pytest.mark.parametrize("n_dims", [2, 3])
def test_inference_within_bf16_activation_bound(ctx, n_dims):
    """Matrix-core inference rounds the sine activations to BF16 (unit
    roundoff 2^-9) before the 64x64 layer. ReLU is 1-Lipschitz, so the
    output error is at most 2^-9 sum_j |w3_j| sum_k |W2_jk| (plus FP32
    rounding of the sine argument and accumulations, covered by 2^-14)."""
    import torch

    a = hd.gen.uniform(30 + n_dims, 40_000, n_dims)
    b = (hd.gen.uniform(40 + n_dims, 3_000, n_dims) - np.float32(0.5)) * np.float32(0.05) \
        + np.float32(0.4)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    model = hd.Model.train(ctx, da, db, steps=200, batch=2048, n_samples=20_000)
    pts = a[:2048]
    pred = torch.empty(pts.shape[0], dtype=torch.float32, device="cuda")
    model.predict(ctx, torch.from_numpy(pts).cuda(), pred)
    got = pred.cpu().numpy().astype(np.float64)
    w = bf16_round(model.params())
    lo, inv = model.normalisation()
    x = ((pts - lo) * inv).astype(np.float32)
    ref = oracle_forward(w, n_dims, x)
    off = 64 * n_dims + 64
    w2 = np.abs(w[off : off + 64 * 64].astype(np.float64)).reshape(64, 64)
    w3 = np.abs(w[off + 64 * 64 + 64 : off + 64 * 64 + 128].astype(np.float64))
    bound = (2.0**-9 + 2.0**-14) * float(w3 @ w2.sum(axis=1))
    err = np.abs(got - ref)
    print(f"[mlp] {n_dims}D inference: max err {err.max():.3g}, bound {bound:.3g}")
    assert err.max() <= bound
    # The bound is a worst case; typical errors are far smaller, and a
    # transposed or misplaced operand block would not be.
    assert np.median(err) <= bound / 8
    model.close()


def test_training_fits_a_smooth_target(ctx):
    import torch

    # A uniform, B one tight cluster: the NN distance is (almost) the
    # distance to the cluster, a smooth function the network must fit.
    a = hd.gen.uniform(3, 60_000, 2)
    b = (hd.gen.uniform(4, 5_000, 2) - np.float32(0.5)) * np.float32(0.02) + np.float32(0.3)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    model = hd.Model.train(ctx, da, db, steps=600, batch=2048, n_samples=40_000)
    assert model.final_loss < 0.1
    model.close()


def test_training_lowers_loss_and_order_keeps_bits(ctx):
    import torch

    a, b = hd.gen.gmm_pair(200_000)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    model = hd.Model.train(ctx, da, db, steps=400, batch=2048, n_samples=50_000)
    # Targets are standardised, so predicting the mean scores 1.0.
    assert model.final_loss < 1.0
    pred = torch.empty(a.shape[0], dtype=torch.float32, device="cuda")
    model.predict(ctx, da, pred)
    assert torch.isfinite(pred).all()
    # What the network gains: the points it ranks first really are far from
    # B. Exact NN distances of all of A (grid index) against the ranking.
    nn_d2 = torch.empty(a.shape[0], dtype=torch.float32, device="cuda")
    nn_idx = torch.empty(a.shape[0], dtype=torch.int32, device="cuda")
    hd.nn_device(ctx, hd.Variant.GRID, da, db, nn_d2, nn_idx)
    d2 = nn_d2.cpu().numpy()
    top = np.argsort(-pred.cpu().numpy())[: a.shape[0] // 100]
    gain = float(np.median(d2[top]) / np.median(d2))
    print(f"[mlp] median NN d2 of the top 1% predicted / overall: {gain:.2f}")
    assert gain > 1.5
    o = oracle.eb_directed(a, b)
    for variant in (hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH):
        r = hd.directed_device(ctx, variant, da, db, learned_order=True, model=model)
        assert (r.cmax2_bits, r.argmax, r.nn) == (o.cmax2_bits, o.argmax, o.nn), variant
    # Without an explicit model the library trains one for these inputs.
    r = hd.directed_device(ctx, hd.Variant.EB, da, db, learned_order=True)
    assert (r.cmax2_bits, r.argmax) == (o.cmax2_bits, o.argmax)
    # A model trained on other inputs: its farthest labelled points index
    # another A (those past the end are skipped); the bits do not change.
    a2 = np.ascontiguousarray(a[:30_000][::-1])
    o2 = oracle.eb_directed(a2, b)
    r2 = hd.directed_device(ctx, hd.Variant.EB, torch.from_numpy(a2).cuda(), db,
                            learned_order=True, model=model)
    assert (r2.cmax2_bits, r2.argmax, r2.nn) == (o2.cmax2_bits, o2.argmax, o2.nn)
    model.close()


def test_learned_order_evaluates_fewer_pairs(ctx):
    """What the trained network gains over the method without it, on a
    clustered pair (the 256-component recipe of the 50M workload at 2M
    points): the same early break evaluates fewer pairs in the learned order
    than in the pseudo-random one, both as a single pass (the reference GPU
    kernel's schedule, where only the order differs) and through the full
    pipeline; the bits are the oracle's either way."""
    import torch

    a, b = hd.gen.gmm_pair(2_000_000, 256, 0.01, (0.002, 0.001), 0.01, (10, 11, 12, 13))
    o = oracle.eb_directed(a, b)
    want = (o.cmax2_bits, o.argmax, o.nn)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    model = hd.Model.train(ctx, da, db, steps=2000, batch=4096, n_samples=500_000)
    got = {}
    for name, kw in (("plain_single", dict(single_pass=True)),
                     ("learned_single", dict(single_pass=True, learned_order=True)),
                     ("plain", {}), ("learned", dict(learned_order=True))):
        r = hd.directed_device(ctx, hd.Variant.EB, da, db, model=model, **kw)
        assert (r.cmax2_bits, r.argmax, r.nn) == want, name
        got[name] = r.evals
    print(f"[mlp] evaluations: {got}", flush=True)
    # Measured on one MI355X: 6.5e10 against 1.02e11 as a single pass
    # (0.64x), 1.36e10 against 2.77e10 through the pipeline (0.49x).
    assert got["learned_single"] < 0.8 * got["plain_single"]
    assert got["learned"] < 0.7 * got["plain"]
    model.close()


def test_learned_order_on_a_3d_shell_keeps_bits(ctx):
    """A pair whose HD comes from radial noise of a thin spherical shell,
    which the network's inputs barely resolve: its predicted farthest points
    need not reach the final cmax2, and the rounds then complete their top
    entries (a round escalating more than an eighth of A). Bits, arg-max
    and NN stay the oracle's, with the learned order on the early break and
    on the index variants."""
    import torch

    a = hd.gen.sphere(31, 200_000, 0.005, 0.0)
    b = hd.gen.sphere(32, 150_000, 0.005, 2.0)
    o = oracle.eb_directed(a, b)
    want = (o.cmax2_bits, o.argmax, o.nn)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    model = hd.Model.train(ctx, da, db, steps=500, batch=4096, n_samples=100_000)
    for variant in (hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH):
        r = hd.directed_device(ctx, variant, da, db, learned_order=True, model=model)
        assert (r.cmax2_bits, r.argmax, r.nn) == want, variant
    r = hd.directed_device(ctx, hd.Variant.EB, da, db)
    assert (r.cmax2_bits, r.argmax, r.nn) == want
    model.close()
