"""ctypes wrapper of the CPU oracle (oracle/hd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libhd_oracle.so"


class OrcResult(C.Structure):
    _fields_ = [
        ("cmax2_bits", C.c_uint32),
        ("argmax", C.c_uint32),
        ("nn", C.c_uint32),
        ("n_threads", C.c_uint32),
        ("compared_pairs", C.c_uint64),
        ("compute_ms", C.c_double),
        ("shuffle_ms", C.c_double),
    ]


@dataclass
class Directed:
    cmax2_bits: int
    argmax: int
    nn: int
    n_threads: int
    compared_pairs: int
    compute_ms: float
    shuffle_ms: float

    @property
    def cmax2(self) -> float:
        return float(np.array([self.cmax2_bits], dtype=np.uint32).view(np.float32)[0])

    @property
    def hd(self) -> float:
        # sqrt of the FP32 value, rounded to FP32 (early_break.h:114).
        return float(np.sqrt(np.float32(self.cmax2)))


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    F = C.POINTER(C.c_float)
    U32 = C.POINTER(C.c_uint32)
    lib.orc_d2.restype = C.c_float
    lib.orc_d2.argtypes = [F, F, C.c_int]
    lib.orc_eb_directed.restype = C.c_int
    lib.orc_eb_directed.argtypes = [F, C.c_uint64, F, C.c_uint64, C.c_int, C.c_int,
                                    C.c_uint64, C.c_int, C.POINTER(OrcResult)]
    lib.orc_brute_directed.restype = C.c_int
    lib.orc_brute_directed.argtypes = [F, C.c_uint64, F, C.c_uint64, C.c_int, C.c_int,
                                       C.POINTER(OrcResult)]
    lib.orc_nn_all.restype = C.c_int
    lib.orc_nn_all.argtypes = [F, C.c_uint64, F, C.c_uint64, C.c_int, C.c_int, F, U32]
    lib.orc_first_at_d2.restype = C.c_uint32
    lib.orc_first_at_d2.argtypes = [F, F, C.c_uint64, C.c_int, C.c_uint32]
    lib.orc_hd_bounds2.restype = None
    lib.orc_hd_bounds2.argtypes = [F, F, C.c_int, F, F]
    lib.orc_mbr.restype = None
    lib.orc_mbr.argtypes = [F, C.c_uint64, C.c_int, F]
    lib.orc_mlp_loss_grad.restype = C.c_double
    lib.orc_mlp_loss_grad.argtypes = [F, C.c_int, C.c_int, F, F, C.c_int,
                                      C.POINTER(C.c_double)]
    _lib = lib
    return lib


def _f(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _pts(x) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    assert a.ndim == 2 and a.shape[1] in (2, 3), a.shape
    return a


def d2(p, q) -> np.float32:
    p = np.ascontiguousarray(np.asarray(p, dtype=np.float32))
    q = np.ascontiguousarray(np.asarray(q, dtype=np.float32))
    return np.float32(load().orc_d2(_f(p), _f(q), p.shape[0]))


def _directed(fn, a, b, *args) -> Directed:
    a, b = _pts(a), _pts(b)
    r = OrcResult()
    rc = fn(_f(a), a.shape[0], _f(b), b.shape[0], a.shape[1], *args, C.byref(r))
    if rc != 0:
        raise ValueError("oracle rejected its arguments")
    return Directed(r.cmax2_bits, r.argmax, r.nn, r.n_threads, r.compared_pairs,
                    r.compute_ms, r.shuffle_ms)


def eb_directed(a, b, n_threads: int = 0, seed: int = 0, shuffle: bool = True) -> Directed:
    """CPU early break (reference early_break.h:44-115)."""
    return _directed(load().orc_eb_directed, a, b, n_threads, seed, 1 if shuffle else 0)


def brute_directed(a, b, n_threads: int = 0) -> Directed:
    return _directed(load().orc_brute_directed, a, b, n_threads)


def eb_symmetric(a, b, n_threads: int = 0, seed: int = 0) -> tuple[Directed, Directed]:
    return eb_directed(a, b, n_threads, seed), eb_directed(b, a, n_threads, seed)


def nn_all(q, r, n_threads: int = 0) -> tuple[np.ndarray, np.ndarray]:
    q, r = _pts(q), _pts(r)
    d = np.empty(q.shape[0], dtype=np.float32)
    i = np.empty(q.shape[0], dtype=np.uint32)
    rc = load().orc_nn_all(_f(q), q.shape[0], _f(r), r.shape[0], q.shape[1], n_threads,
                           _f(d), i.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc == 0
    return d, i


def mbr(p) -> np.ndarray:
    p = _pts(p)
    out = np.empty(2 * p.shape[1], dtype=np.float32)
    load().orc_mbr(_f(p), p.shape[0], p.shape[1], _f(out))
    return out


def hd_bounds2(mbr_a, mbr_b, n_dims: int) -> tuple[np.float32, np.float32]:
    ma = np.ascontiguousarray(mbr_a, dtype=np.float32)
    mb = np.ascontiguousarray(mbr_b, dtype=np.float32)
    lb = np.zeros(1, np.float32)
    ub = np.zeros(1, np.float32)
    load().orc_hd_bounds2(_f(ma), _f(mb), n_dims, _f(lb), _f(ub))
    return lb[0], ub[0]


def mlp_loss_grad(params, d_in: int, hidden: int, x, y) -> tuple[float, np.ndarray]:
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.zeros(p.shape[0], dtype=np.float64)
    loss = load().orc_mlp_loss_grad(_f(p), d_in, hidden, _f(x), _f(y), y.shape[0],
                                    g.ctypes.data_as(C.POINTER(C.c_double)))
    return float(loss), g
