"""The library's radix sort (sort.hip) against numpy's stable argsort: the
same permutation for u16 keys, u32 keys on their low end_bit bits and f32
keys in descending order, at sizes around the 4096-key tile, with heavy
duplication, presorted and reversed inputs, and at 20M keys. Equal keys
must keep their input order (the grid's, the LBVH's and the learned order's
visit orders depend on it)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from hausdorffdistance_amd import _lib

pytestmark = pytest.mark.gpu

U16, U32, F32_DESC = 0, 1, 2


def device_sort(ctx, kind, keys: np.ndarray, vals: np.ndarray, end_bit: int, begin_bit: int = 0):
    import torch

    tk = torch.from_numpy(keys.view(np.int16 if keys.dtype == np.uint16 else np.int32)).cuda()
    tv = torch.from_numpy(vals.view(np.int32)).cuda()
    ok, ov = torch.empty_like(tk), torch.empty_like(tv)
    rc = _lib.load().hd_sort_pairs_probe_device(
        ctx.handle, kind, C.c_void_p(tk.data_ptr()), C.c_void_p(tv.data_ptr()), keys.size,
        begin_bit, end_bit, C.c_void_p(ok.data_ptr()), C.c_void_p(ov.data_ptr()))
    assert rc == 0, _lib.load().hd_last_error().decode()
    return ok.cpu().numpy().view(keys.dtype), ov.cpu().numpy().view(np.uint32)


def expected_order(kind, keys: np.ndarray, end_bit: int, begin_bit: int = 0) -> np.ndarray:
    if kind == U16:
        return np.argsort(keys, kind="stable")
    if kind == U32:
        k = keys.astype(np.uint64)
        mask = np.uint64((1 << (end_bit - begin_bit)) - 1)
        return np.argsort((k >> np.uint64(begin_bit)) & mask, kind="stable")
    # descending IEEE total order (+0.0 before -0.0)
    u = keys.view(np.uint32)
    o = np.where(u & np.uint32(0x80000000), ~u, u | np.uint32(0x80000000))
    return np.argsort(~o, kind="stable")


def check(ctx, kind, keys, end_bit=32, vals=None, begin_bit=0):
    n = keys.size
    if vals is None:
        vals = np.arange(n, dtype=np.uint32)
    ok, ov = device_sort(ctx, kind, keys, vals, end_bit, begin_bit)
    order = expected_order(kind, keys, end_bit, begin_bit)
    np.testing.assert_array_equal(ov, vals[order])
    np.testing.assert_array_equal(ok.view(np.uint32 if kind != U16 else np.uint16),
                                  keys[order].view(np.uint32 if kind != U16 else np.uint16))


SIZES = [1, 2, 63, 64, 65, 4095, 4096, 4097, 12_289, 100_003]


@pytest.mark.parametrize("n", SIZES)
def test_u32_full_width(ctx, n):
    rng = np.random.default_rng(n)
    check(ctx, U32, rng.integers(0, 2**32, n, dtype=np.uint32))


@pytest.mark.parametrize("n", SIZES)
def test_u32_duplicates_and_values(ctx, n):
    rng = np.random.default_rng(1000 + n)
    keys = rng.integers(0, 37, n, dtype=np.uint32) * np.uint32(65_537)
    vals = rng.integers(0, 2**32, n, dtype=np.uint32)
    check(ctx, U32, keys, 32, vals)


@pytest.mark.parametrize("end_bit", [0, 1, 7, 8, 9, 20, 22, 31])
def test_u32_low_bits_only(ctx, end_bit):
    """Bits at and above end_bit are carried, not sorted on."""
    rng = np.random.default_rng(end_bit)
    check(ctx, U32, rng.integers(0, 2**32, 50_000, dtype=np.uint32), end_bit)


@pytest.mark.parametrize("bits", [(16, 32), (8, 20), (31, 32), (5, 5)])
def test_u32_bit_range(ctx, bits):
    """Bits [begin_bit, end_bit) only (the query-order sort uses 16..32)."""
    begin, end = bits
    rng = np.random.default_rng(begin * 100 + end)
    check(ctx, U32, rng.integers(0, 2**32, 70_000, dtype=np.uint32), end, begin_bit=begin)


def test_u32_presorted_and_reversed(ctx):
    a = np.arange(70_000, dtype=np.uint32) * np.uint32(977)
    check(ctx, U32, a)
    check(ctx, U32, a[::-1].copy())
    check(ctx, U32, np.full(70_000, 12345, np.uint32))


@pytest.mark.parametrize("n", [1, 4097, 300_000])
def test_u16(ctx, n):
    rng = np.random.default_rng(n)
    check(ctx, U16, rng.integers(0, 2**16, n, dtype=np.uint16))
    check(ctx, U16, rng.integers(0, 5, n, dtype=np.uint16))


@pytest.mark.parametrize("n", [1, 4096, 200_001])
def test_f32_descending(ctx, n):
    rng = np.random.default_rng(n)
    x = rng.normal(0, 10, n).astype(np.float32)
    x[::7] = np.round(x[::7])          # duplicates
    x[::11] = 0.0
    x[5::13] = -0.0
    x[3::17] = np.float32(3.4e38)
    x[4::19] = np.float32(-1e-42)      # subnormal
    check(ctx, F32_DESC, x)


def test_large_u32_and_u16(ctx):
    rng = np.random.default_rng(7)
    check(ctx, U32, rng.integers(0, 2**26, 20_000_000, dtype=np.uint32), 26)
    check(ctx, U16, rng.integers(0, 2**16, 20_000_000, dtype=np.uint16))
