"""Error paths through the C ABI on the GPU: non-finite coordinates are an
HD_EINVAL error for every variant and entry point (the reference produces
garbage for them), and the library keeps working afterwards."""
from __future__ import annotations

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd import _lib

pytestmark = pytest.mark.gpu

VARIANTS = [hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID,
            hd.Variant.NN, hd.Variant.RT_HDIST]


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("which", ["a", "b"])
def test_non_finite_input_is_einval(ctx, bad, which):
    import torch

    a, b = hd.gen.uniform_pair(2_000, 2)
    (a if which == "a" else b)[123, 1] = bad
    for v in VARIANTS:
        cfg = hd.RunConfig(variant=v, n_dims=2, repeat=1)
        with pytest.raises(hd.HDError) as e:
            hd.run_hausdorff_distance(cfg, a, b, ctx=ctx)
        assert e.value.code == _lib.HD_EINVAL, v
        with pytest.raises(hd.HDError) as e:
            hd.directed_device(ctx, v, torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
        assert e.value.code == _lib.HD_EINVAL, v
    # the context still serves finite inputs
    a2, b2 = hd.gen.uniform_pair(2_000, 2)
    r = hd.directed_device(ctx, hd.Variant.EB, torch.from_numpy(a2).cuda(),
                           torch.from_numpy(b2).cuda())
    assert r.argmax < 2_000


MEM_VARIANTS = [hd.Variant.EB, hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID,
                hd.Variant.NN]


def test_memory_limit_fails_cleanly_and_context_recovers():
    """An allocation failure, forced by capping the context's device memory:
    every variant and the batch entry point return HD_ENOMEM (no abort), the
    buffers stay under the cap, and after the cap is lifted the same context
    returns the oracle's answer."""
    import torch

    from oracle import oracle

    a, b = hd.gen.uniform_pair(300_000, 2)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    offs = torch.tensor([0, 100_000, 300_000], dtype=torch.int64, device="cuda")
    want = oracle.eb_directed(a, b, n_threads=8)
    with hd.Context(0) as c:
        cap = 1 << 20
        c.set_memory_limit(cap)
        assert c.memory_held == 0
        for v in MEM_VARIANTS:
            with pytest.raises(hd.HDError) as e:
                hd.directed_device(c, v, da, db)
            assert e.value.code == _lib.HD_ENOMEM, v
            assert "memory limit" in str(e.value), v
            assert c.memory_held <= cap, v
        with pytest.raises(hd.HDError) as e:
            hd.batch_directed(c, da, offs, db, offs)
        assert e.value.code == _lib.HD_ENOMEM
        # a cap that fits a small problem serves it
        c.set_memory_limit(256 << 20)
        a2, b2 = hd.gen.uniform_pair(20_000, 2)
        r2 = hd.directed_device(c, hd.Variant.EB, torch.from_numpy(a2).cuda(),
                                torch.from_numpy(b2).cuda())
        o2 = oracle.eb_directed(a2, b2, n_threads=8)
        assert (r2.cmax2_bits, r2.argmax, r2.nn) == (o2.cmax2_bits, o2.argmax, o2.nn)
        assert 0 < c.memory_held <= 256 << 20
        c.set_memory_limit(0)
        for v in MEM_VARIANTS:
            r = hd.directed_device(c, v, da, db)
            assert (r.cmax2_bits, r.argmax, r.nn) == (want.cmax2_bits, want.argmax, want.nn), v
        got = hd.batch_directed(c, da, offs, db, offs)
        for p, (lo, hi) in enumerate([(0, 100_000), (100_000, 300_000)]):
            o = oracle.eb_directed(a[lo:hi], b[lo:hi], n_threads=8)
            assert (got[p].cmax2_bits, got[p].argmax, got[p].nn) == (o.cmax2_bits, o.argmax, o.nn)
