"""Point-set files of the reference's inputs (include/hd_io.h): WKT, OFF,
PLY and the .bin point cache, and the -normalize / -translate preprocessing
(src/loaders/loader.h, ply_loader.h, translate_points.h). Host-only: no GPU
is touched."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib

INPUT_TYPES = {"wkt": _lib.HD_INPUT_WKT, "off": _lib.HD_INPUT_OFF,
               "ply": _lib.HD_INPUT_PLY, "bin": _lib.HD_INPUT_BIN}


def load_points(path: str | Path, input_type: str, n_dims: int, limit: int | None = None,
                serialize: str | Path | None = None) -> np.ndarray:
    """Points of `path` as an (n, n_dims) float32 array. `input_type` is one
    of "wkt", "off", "ply", "bin" (the reference's -input_type); `limit`
    caps the count (-limit); `serialize` is the cache folder (-serialize)."""
    lib = _lib.load()
    t = INPUT_TYPES[input_type.lower()]
    out = C.POINTER(C.c_float)()
    n = C.c_uint64(0)
    lim = (1 << 63) if limit is None else int(limit)
    ser = str(serialize).encode() if serialize else None
    _lib.check(lib.hd_load_points(str(path).encode(), t, int(n_dims), lim, ser,
                                  C.byref(out), C.byref(n)))
    try:
        arr = np.ctypeslib.as_array(out, shape=(n.value * n_dims,)).copy()
    finally:
        lib.hd_free_points(out)
    return arr.reshape(n.value, n_dims)


def normalize_points(points: np.ndarray) -> np.ndarray:
    """(x - lower) / max_extent per coordinate, in place (float32)."""
    p = _f32(points)
    _lib.check(_lib.load().hd_normalize_points(
        p.ctypes.data_as(C.POINTER(C.c_float)), p.shape[0], p.shape[1]))
    return p


def translate_points(points: np.ndarray, ratio: float, dim: int = 0) -> np.ndarray:
    """Shift dimension `dim` by its extent times `ratio`, in place."""
    p = _f32(points)
    _lib.check(_lib.load().hd_translate_points(
        p.ctypes.data_as(C.POINTER(C.c_float)), p.shape[0], p.shape[1], int(dim),
        float(ratio)))
    return p


def _f32(points: np.ndarray) -> np.ndarray:
    if points.dtype != np.float32 or not points.flags.c_contiguous or points.ndim != 2:
        raise ValueError("points must be a C-contiguous (n, d) float32 array")
    return points
