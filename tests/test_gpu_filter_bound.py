"""The matrix-core filter's error bound, measured. The early break drops a
point on the filter only when its approximate minimum is below
cmax2 (1 - 2^-19) - E, and escalates it with approx + E; both rest on
|D - d2| <= E for every pair the filter evaluates, with
E = 48 u rho^2 (2D) or 64 u rho^2 (3D), u = 2^-24 and rho the radius of the
joint box about its centre. The test hook hd_filter_probe_device computes D
with the filter's own tiles, operands and v_mfma_f32_16x16x4_f32; here
every pair is compared with the exact squared distance (FP64 from the FP32
coordinates) on inputs chosen to stress the bound: coordinates at the
corners of the box (largest products), pairs that nearly coincide far from
the centre (largest cancellation), sets far from the origin, and 3D."""
from __future__ import annotations

import ctypes as C
import zlib

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd import _lib

pytestmark = pytest.mark.gpu
U = 2.0 ** -24


def probe(ctx, a: np.ndarray, b: np.ndarray):
    import torch

    nd = a.shape[1]
    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    out = torch.empty((a.shape[0], b.shape[0]), dtype=torch.float32, device="cuda")
    cna = torch.empty(a.shape[0], dtype=torch.float32, device="cuda")
    params = np.zeros(5, np.float32)
    rc = _lib.load().hd_filter_probe_device(
        ctx.handle, nd, C.c_void_p(da.data_ptr()), a.shape[0], C.c_void_p(db.data_ptr()),
        b.shape[0], C.c_void_p(out.data_ptr()), C.c_void_p(cna.data_ptr()),
        params.ctypes.data_as(_lib._F))
    assert rc == 0, _lib.load().hd_last_error().decode()
    d = out.cpu().numpy().astype(np.float64)
    if nd == 3:
        d = d - cna.cpu().numpy().astype(np.float64)[:, None]
    return d, params


def exact_d2(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    return ((a64[:, None, :] - b64[None, :, :]) ** 2).sum(-1)


def rho2_of(a, b, centre):
    pts = np.concatenate([a, b]).astype(np.float64)
    lo, hi = pts.min(0), pts.max(0)
    c = centre[: a.shape[1]].astype(np.float64)
    r = np.maximum(np.abs(lo - c), np.abs(hi - c))
    return float((r * r).sum())


def corners(rng, n, nd, scale=1.0, shift=0.0):
    """Points at the box corners, with a little jitter so values differ."""
    s = rng.choice([-1.0, 1.0], size=(n, nd))
    return (shift + scale * (s * (1 - rng.random((n, nd)) * 1e-3))).astype(np.float32)


def near_pairs(rng, n, nd, shift=0.0):
    """Pairs that nearly coincide close to a corner of [-1, 1]^nd (the
    largest cancellation in |a'|^2 + |b'|^2 - 2 a'.b'), plus two points
    that make the joint box [-1, 1]^nd."""
    base = (0.999 - rng.random((n, nd)) * 1e-3) * rng.choice([-1.0, 1.0], size=(1, nd))
    a = base
    b = base + rng.normal(0, 1e-4, (n, nd))
    box = np.array([[-1.0] * nd, [1.0] * nd])
    return ((np.concatenate([a, box]) + shift).astype(np.float32),
            (np.concatenate([b, box]) + shift).astype(np.float32))


CASES = {
    "corners_2d": lambda rng: (corners(rng, 1000, 2), corners(rng, 900, 2)),
    "uniform_2d": lambda rng: (rng.uniform(-1, 1, (1500, 2)).astype(np.float32),
                               rng.uniform(-1, 1, (1300, 2)).astype(np.float32)),
    "cancellation_2d": lambda rng: near_pairs(rng, 1200, 2),
    "far_offset_2d": lambda rng: near_pairs(rng, 1200, 2, shift=1000.0),
    "corners_3d": lambda rng: (corners(rng, 800, 3), corners(rng, 700, 3)),
    "cancellation_3d": lambda rng: near_pairs(rng, 1000, 3),
    "far_offset_3d": lambda rng: near_pairs(rng, 1000, 3, shift=-500.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_filter_error_within_bound(ctx, name):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    a, b = CASES[name](rng)
    nd = a.shape[1]
    d, params = probe(ctx, a, b)
    exact = exact_d2(a, b)
    err = np.abs(d - exact).max()
    rho2 = rho2_of(a, b, params[:3])
    bound = float(params[3])
    model = (48.0 if nd == 2 else 64.0) * U * rho2
    print(f"[filter] {name}: max |D - d2| = {err / (U * rho2):.2f} u rho^2, "
          f"E = {bound / (U * rho2):.2f} u rho^2", flush=True)
    # the bound the library uses is the model's (with a 2^-20 margin)
    assert bound >= model
    assert err <= bound
