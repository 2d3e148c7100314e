"""Deterministic synthetic workloads (wrappers of include/hd_gen.h).

Each workload is named by what it is. Seeds and recipes are fixed so the
benchmark, the CLI and the tests draw bit-identical inputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def uniform(seed: int, n: int, n_dims: int = 2) -> np.ndarray:
    out = np.empty((n, n_dims), dtype=np.float32)
    _lib.check(_lib.load().hd_gen_uniform(seed, n, n_dims, _fptr(out)))
    return out


def gmm_pair(
    n: int,
    n_clusters: int = 64,
    sigma: float = 0.01,
    shift=(0.002, 0.001),
    outlier_fraction: float = 0.01,
    seeds=(3, 4, 5, 6),
) -> tuple[np.ndarray, np.ndarray]:
    """A from a Gaussian mixture; B = A + shift with a fraction of uniform
    outliers (n // 100 points by default)."""
    a = np.empty((n, 2), dtype=np.float32)
    b = np.empty((n, 2), dtype=np.float32)
    n_out = int(round(n * outlier_fraction))
    _lib.check(
        _lib.load().hd_gen_gmm_pair(
            n, n_clusters, sigma, shift[0], shift[1], n_out, seeds[0], seeds[1], seeds[2],
            seeds[3], _fptr(a), _fptr(b)
        )
    )
    return a, b


def sphere(seed: int, n: int, noise: float = 0.005, rot_deg: float = 0.0) -> np.ndarray:
    out = np.empty((n, 3), dtype=np.float32)
    _lib.check(_lib.load().hd_gen_sphere(seed, n, noise, rot_deg, _fptr(out)))
    return out


@dataclass
class Batch:
    sizes_a: np.ndarray  # uint32 [n_pairs]
    sizes_b: np.ndarray
    offs_a: np.ndarray  # uint64 [n_pairs + 1], in points
    offs_b: np.ndarray
    a: np.ndarray  # float32 [sum sizes_a, 2]
    b: np.ndarray


def batch(seed: int = 9, n_pairs: int = 4096, sigma: float = 0.02,
          max_pairs_points: int | None = None) -> Batch:
    lib = _lib.load()
    sa = np.empty(n_pairs, dtype=np.uint32)
    sb = np.empty(n_pairs, dtype=np.uint32)
    _lib.check(lib.hd_gen_batch_sizes(
        seed, n_pairs, sa.ctypes.data_as(C.POINTER(C.c_uint32)),
        sb.ctypes.data_as(C.POINTER(C.c_uint32))))
    oa = np.zeros(n_pairs + 1, dtype=np.uint64)
    ob = np.zeros(n_pairs + 1, dtype=np.uint64)
    oa[1:] = np.cumsum(sa, dtype=np.uint64)
    ob[1:] = np.cumsum(sb, dtype=np.uint64)
    a = np.empty((int(oa[-1]), 2), dtype=np.float32)
    b = np.empty((int(ob[-1]), 2), dtype=np.float32)
    _lib.check(lib.hd_gen_batch_points(
        seed, n_pairs, sigma,
        sa.ctypes.data_as(C.POINTER(C.c_uint32)), sb.ctypes.data_as(C.POINTER(C.c_uint32)),
        oa.ctypes.data_as(C.POINTER(C.c_uint64)), ob.ctypes.data_as(C.POINTER(C.c_uint64)),
        _fptr(a), _fptr(b)))
    return Batch(sa, sb, oa, ob, a, b)


def subsample(seed: int, n: int, k: int) -> np.ndarray:
    out = np.empty(k, dtype=np.uint32)
    _lib.check(_lib.load().hd_gen_subsample(seed, n, k, out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


# Named workloads.
def uniform_pair(n: int = 100_000, n_dims: int = 2) -> tuple[np.ndarray, np.ndarray]:
    """A and B i.i.d. uniform in [0,1]^d, seeds 1 and 2."""
    return uniform(1, n, n_dims), uniform(2, n, n_dims)


def gmm_10m(n: int = 10_000_000) -> tuple[np.ndarray, np.ndarray]:
    """64-component mixture (sigma 0.01), B = A + (0.002, 0.001), 1% outliers;
    seeds 3-6."""
    return gmm_pair(n, 64, 0.01, (0.002, 0.001), 0.01, (3, 4, 5, 6))


def gmm_50m_256(n: int = 50_000_000) -> tuple[np.ndarray, np.ndarray]:
    """As gmm_10m with 256 components; seeds 10-13."""
    return gmm_pair(n, 256, 0.01, (0.002, 0.001), 0.01, (10, 11, 12, 13))


def sphere_pair(n_a: int = 20_000_000, n_b: int = 15_000_000) -> tuple[np.ndarray, np.ndarray]:
    """Noisy unit spheres (radial sigma 0.005), B rotated 2 degrees about z;
    seeds 7 and 8."""
    return sphere(7, n_a, 0.005, 0.0), sphere(8, n_b, 0.005, 2.0)
