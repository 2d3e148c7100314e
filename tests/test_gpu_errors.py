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
