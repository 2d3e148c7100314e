"""The allocation-free entry point (hd_directed_device_noalloc) against the
early break of hd_directed_device on the same inputs: cmax2 bits, arg-max and
NN must be identical, on two calls in a row (the second with a smaller
input, so stale scratch contents from the first would show), in 2D and 3D;
and inputs beyond its limit are refused with HD_EINVAL."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd import _lib

pytestmark = pytest.mark.gpu


def noalloc(ctx, da, db, flags=0, seed=0):
    out = _lib.DirectedResultC()
    rc = _lib.load().hd_directed_device_noalloc(
        ctx.handle, int(da.shape[1]), C.c_void_p(da.data_ptr()), int(da.shape[0]),
        C.c_void_p(db.data_ptr()), int(db.shape[0]), flags, seed, C.byref(out))
    return rc, hd.DirectedResult.from_c(out)


def triple(r):
    return (r.cmax2_bits, r.argmax, r.nn)


@pytest.mark.parametrize("n_dims", [2, 3])
def test_noalloc_matches_directed_device(ctx, n_dims):
    import torch

    big_a, big_b = hd.gen.gmm_pair(400_000) if n_dims == 2 else (
        hd.gen.sphere(7, 300_000), hd.gen.sphere(8, 1_000_000, rot_deg=2.0))
    small_a = hd.gen.uniform(41, 5_000, n_dims)
    small_b = hd.gen.uniform(42, 3_001, n_dims)
    for a, b in ((big_a, big_b), (small_a, small_b)):
        da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
        db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
        print(f"[noalloc] {n_dims}D {a.shape[0]} x {b.shape[0]}", flush=True)
        rc, got = noalloc(ctx, da, db)
        assert rc == 0, _lib.load().hd_last_error().decode()
        want = hd.directed_device(ctx, hd.Variant.EB, da, db, single_pass=True)
        assert triple(got) == triple(want)
        assert triple(got) == triple(hd.directed_device(ctx, hd.Variant.EB, da, db))
        assert got.evals > 0 and got.time_ms > 0
        # input order as well
        rc, got2 = noalloc(ctx, da, db, flags=_lib.HD_FLAG_NO_SHUFFLE)
        assert rc == 0 and triple(got2) == triple(want)


def test_noalloc_limits(ctx):
    import torch

    a = torch.rand((100, 2), device="cuda")
    b = torch.rand(((1 << 20) + 1, 2), device="cuda")
    rc, _ = noalloc(ctx, a, b)
    assert rc == _lib.HD_EINVAL
    assert "limit" in _lib.load().hd_last_error().decode()
    rc, _ = noalloc(ctx, a, b[: 1 << 20])
    assert rc == 0
    bad = a.clone()
    bad[3, 0] = float("nan")
    rc, _ = noalloc(ctx, bad, b[:10])
    assert rc == _lib.HD_EINVAL
    rc, _ = noalloc(ctx, a, b[:10], flags=_lib.HD_FLAG_LEARNED_ORDER)
    assert rc == _lib.HD_EUNSUPPORTED
