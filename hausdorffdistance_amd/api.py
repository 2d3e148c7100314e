"""Host-side mirror of the reference's run interface, over libhd.so.

Reference interface (pwrliang/HausdorffDistance):
  * RunConfig            src/run_config.h:20-61 (variant, execution, n_dims,
                         repeat, seed, parallelism, check, hybrid knobs)
  * RunHausdorffDistance src/run_hausdorff_distance.cuh:5, .cu:53-402
  * HausdorffDistance<T,N>::CalculateDistance / get_stats
                         src/hausdorff_distance.h:15-42
Same names (snake_case), same argument meaning; errors raise HDError with the
library's code instead of aborting the process.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Any, Optional

import numpy as np

from . import _lib
from ._lib import HDError


class Variant(IntEnum):
    EB = _lib.HD_VARIANT_EB
    GRID = _lib.HD_VARIANT_GRID
    LBVH = _lib.HD_VARIANT_LBVH
    HYBRID = _lib.HD_VARIANT_HYBRID
    NN = _lib.HD_VARIANT_NN
    RT_HDIST = _lib.HD_VARIANT_RT_HDIST

    @staticmethod
    def from_string(s: str) -> "Variant":
        """CLI strings of reference src/main.cpp:76-94; "rt" is the BVH query
        (a software LBVH here)."""
        key = s.lower()
        table = {"eb": Variant.EB, "rt": Variant.LBVH, "lbvh": Variant.LBVH,
                 "hybrid": Variant.HYBRID, "grid": Variant.GRID, "nn": Variant.NN,
                 "rt-hdist": Variant.RT_HDIST}
        if key not in table:
            raise HDError(_lib.HD_EINVAL, f"Unknown variant: {s}")
        return table[key]


class Execution(IntEnum):
    CPU = _lib.HD_EXEC_CPU
    GPU = _lib.HD_EXEC_GPU

    @staticmethod
    def from_string(s: str) -> "Execution":
        key = s.lower()
        if key == "cpu":
            return Execution.CPU
        if key == "gpu":
            return Execution.GPU
        raise HDError(_lib.HD_EINVAL, f"Unknown execution: {s}")


def _b_order_flags(b_order: str) -> int:
    table = {"auto": 0, "random": _lib.HD_FLAG_B_ORDER_RANDOM,
             "stratified": _lib.HD_FLAG_B_ORDER_STRATIFIED}
    if b_order not in table:
        raise HDError(_lib.HD_EINVAL, f"Unknown b_order: {b_order}")
    return table[b_order]


@dataclass
class RunConfig:
    """Reference RunConfig restricted to the hot path; defaults of
    src/flags.cc:3-57, except `check`: the reference's -check defaults to on
    (flags.cc:49) and re-runs its CPU early break after every run; here it is
    opt-in for library callers (the hd_exec CLI keeps the reference's
    default)."""

    variant: Variant = Variant.EB
    execution: Execution = Execution.GPU
    n_dims: int = 0
    repeat: int = 5
    seed: int = 0
    parallelism: int = -1
    check: bool = False
    symmetric: bool = False
    learned_order: bool = False
    shuffle: bool = True
    b_order: str = "auto"              # EB visit order of B: auto, random, stratified
    sample_rate: float = 0.01
    n_points_cell: int = 8
    max_hit: int = 256
    eb_only_threshold: int = 0
    auto_tune: bool = False            # strategy and parameters by trial
    auto_tune_n_points_cell: bool = False
    auto_tune_max_hit: bool = False
    stats_n_points_cell: int = 10  # > 0: input statistics in the stats JSON

    def to_c(self) -> _lib.RunConfigC:
        c = _lib.RunConfigC()
        _lib.load().hd_run_config_init(C.byref(c))
        c.variant = int(self.variant)
        c.execution = int(self.execution)
        c.n_dims = int(self.n_dims)
        c.repeat = int(self.repeat)
        c.seed = int(self.seed)
        c.parallelism = int(self.parallelism)
        flags = 0
        if self.check:
            flags |= _lib.HD_FLAG_CHECK
        if self.symmetric:
            flags |= _lib.HD_FLAG_SYMMETRIC
        if self.learned_order:
            flags |= _lib.HD_FLAG_LEARNED_ORDER
        if not self.shuffle:
            flags |= _lib.HD_FLAG_NO_SHUFFLE
        flags |= _b_order_flags(self.b_order)
        c.flags = flags
        c.sample_rate = float(self.sample_rate)
        c.n_points_cell = int(self.n_points_cell)
        c.max_hit = int(self.max_hit)
        c.eb_only_threshold = int(self.eb_only_threshold)
        c.auto_tune = ((_lib.HD_AUTO_TUNE_ALL if self.auto_tune else 0)
                       | (_lib.HD_AUTO_TUNE_N_POINTS_CELL if self.auto_tune_n_points_cell else 0)
                       | (_lib.HD_AUTO_TUNE_MAX_HIT if self.auto_tune_max_hit else 0))
        c.stats_n_points_cell = int(self.stats_n_points_cell)
        return c


@dataclass
class DirectedResult:
    cmax2_bits: int
    argmax: int
    nn: int
    variant_used: int
    hd: float
    evals: int
    time_ms: float

    @property
    def cmax2(self) -> float:
        return float(np.array([self.cmax2_bits], dtype=np.uint32).view(np.float32)[0])

    @staticmethod
    def from_c(r: _lib.DirectedResultC) -> "DirectedResult":
        return DirectedResult(r.cmax2_bits, r.argmax, r.nn, r.variant_used, r.hd, r.evals,
                              r.time_ms)


class BatchResults:
    """Per-pair results of a batch (a sequence of DirectedResult, built on
    access) over one array of the C results; the fields are also available
    as numpy arrays (cmax2_bits, argmax, nn, hd) for all pairs at once."""

    _dtype = None

    def __init__(self, out, n_pairs: int):
        if BatchResults._dtype is None:
            BatchResults._dtype = np.ctypeslib.as_array((_lib.DirectedResultC * 1)()).dtype
        self._out = out  # keeps the ctypes buffer alive
        self._rows = np.frombuffer(out, dtype=BatchResults._dtype, count=n_pairs)

    def __len__(self) -> int:
        return int(self._rows.shape[0])

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        r = self._rows[i]
        return DirectedResult(int(r["cmax2_bits"]), int(r["argmax"]), int(r["nn"]),
                              int(r["variant_used"]), float(r["hd"]), int(r["evals"]),
                              float(r["time_ms"]))

    def __iter__(self):
        for i in range(len(self)):
            yield self[i]

    @property
    def cmax2_bits(self) -> np.ndarray:
        return self._rows["cmax2_bits"].copy()

    @property
    def argmax(self) -> np.ndarray:
        return self._rows["argmax"].copy()

    @property
    def nn(self) -> np.ndarray:
        return self._rows["nn"].copy()

    @property
    def hd(self) -> np.ndarray:
        return self._rows["hd"].copy()


@dataclass
class Result:
    hd: float
    cmax2_bits: int
    argmax: int
    nn: int
    direction: int
    check_pass: int
    ab: DirectedResult
    ba: DirectedResult
    avg_time_ms: float
    median_time_ms: float
    min_time_ms: float
    stats: dict = field(default_factory=dict)


class Context:
    """One HIP stream plus cached device buffers on one device."""

    def __init__(self, device: int = 0):
        lib = _lib.load()
        h = C.c_void_p()
        _lib.check(lib.hd_context_create(device, C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise HDError(_lib.HD_EINVAL, "context is closed")
        return self._h

    @property
    def stream(self) -> int:
        return int(_lib.load().hd_context_stream(self.handle) or 0)

    def set_memory_limit(self, nbytes: int) -> None:
        """Caps the device bytes this context's work buffers may hold (0: no
        cap) and releases the buffers held now; a call that needs more
        raises HDError(HD_ENOMEM) and leaves the context usable."""
        _lib.check(_lib.load().hd_context_set_memory_limit(self.handle, int(nbytes)))

    @property
    def memory_held(self) -> int:
        """Device bytes the context's work buffers hold now."""
        return int(_lib.load().hd_context_memory_held(self.handle))

    def set_profiling(self, enable: bool) -> None:
        """Per-stage device timing on this context's stream."""
        _lib.check(_lib.load().hd_context_set_profiling(self.handle, 1 if enable else 0))

    def stage_times(self, reset: bool = True) -> dict:
        """{stage: {"ms", "calls", "evals", "points"}} since the last reset."""
        buf = C.create_string_buffer(1 << 16)
        _lib.check(_lib.load().hd_context_stage_times(self.handle, buf, 1 << 16,
                                                      1 if reset else 0))
        return json.loads(buf.value.decode()) if buf.value else {}

    def close(self) -> None:
        if self._h is not None:
            _lib.load().hd_context_destroy(self._h)
            self._h = None

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def _is_torch_cuda(x: Any) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _points(x: Any, n_dims: int):
    """Returns (pointer, n, on_device, keepalive) for an (n, d) float32 set."""
    if _is_torch_cuda(x):
        if x.dtype.__repr__() != "torch.float32" or x.dim() != 2 or x.shape[1] != n_dims:
            raise HDError(_lib.HD_EINVAL, f"expected a float32 (n, {n_dims}) CUDA tensor")
        if not x.is_contiguous():
            x = x.contiguous()
        return C.c_void_p(x.data_ptr()), int(x.shape[0]), True, x
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if a.ndim != 2 or a.shape[1] != n_dims:
        raise HDError(_lib.HD_EINVAL, f"expected an (n, {n_dims}) array, got {a.shape}")
    return C.c_void_p(a.ctypes.data), int(a.shape[0]), False, a


def run_hausdorff_distance(config: RunConfig, points_a, points_b,
                           ctx: Optional[Context] = None,
                           stats_cap: int = 1 << 20) -> Result:
    """RunHausdorffDistance (reference src/run_hausdorff_distance.cu:53-402)
    for in-memory inputs: `repeat` timed runs, optional check, stats JSON."""
    lib = _lib.load()
    if config.n_dims not in (2, 3):
        raise HDError(_lib.HD_EINVAL,
                      "Wrong number of dimensions, which can only be 2 or 3")
    pa, na, dev_a, ka = _points(points_a, config.n_dims)
    pb, nb, dev_b, kb = _points(points_b, config.n_dims)
    if dev_a != dev_b:
        raise HDError(_lib.HD_EINVAL, "both sets must be on the host or both on the device")
    c = config.to_c()
    if dev_a:
        c.flags |= _lib.HD_FLAG_INPUT_ON_DEVICE
    out = _lib.ResultC()
    buf = C.create_string_buffer(stats_cap)
    # CPU execution touches no GPU: no context is made for it.
    on_cpu = config.execution == Execution.CPU
    own = ctx is None and not on_cpu
    if own:
        ctx = Context(0)
    try:
        rc = lib.hd_run_hausdorff_distance(ctx.handle if ctx is not None else None, C.byref(c),
                                           pa, na, pb, nb, C.byref(out), buf, stats_cap)
        _lib.check(rc)
    finally:
        if own:
            ctx.close()
    stats = json.loads(buf.value.decode()) if buf.value else {}
    return Result(out.hd, out.cmax2_bits, out.argmax, out.nn, out.direction, out.check_pass,
                  DirectedResult.from_c(out.ab), DirectedResult.from_c(out.ba),
                  out.avg_time_ms, out.median_time_ms, out.min_time_ms, stats)


def directed_device(ctx: Context, variant: Variant, points_a, points_b, *,
                    n_dims: Optional[int] = None, seed: int = 0, shuffle: bool = True,
                    learned_order: bool = False, model=None,
                    single_pass: bool = False,
                    b_order: str = "auto") -> DirectedResult:
    """h(A -> B) on device-resident (torch CUDA) inputs. single_pass (early
    break only): one pass over A in the visit order with no seed, rounds or
    filter, the reference GPU kernel's schedule (a comparison baseline).
    b_order (early break): the visit order of B, "random" (pseudo-random,
    as the reference shuffles it), "stratified" (Morton ranks spread evenly
    over the sequence) or "auto" (stratified with learned_order unless
    single_pass, else random); shuffle=False (input order) overrides it.
    The result is the same in every order."""
    nd = n_dims or int(points_a.shape[1])
    pa, na, da, _ = _points(points_a, nd)
    pb, nb, db, _ = _points(points_b, nd)
    if not (da and db):
        raise HDError(_lib.HD_EINVAL, "directed_device needs CUDA tensors")
    flags = 0
    if not shuffle:
        flags |= _lib.HD_FLAG_NO_SHUFFLE
    if learned_order:
        flags |= _lib.HD_FLAG_LEARNED_ORDER
    if single_pass:
        flags |= _lib.HD_FLAG_SINGLE_PASS
    flags |= _b_order_flags(b_order)
    out = _lib.DirectedResultC()
    mh = getattr(model, "handle", None)
    _lib.check(_lib.load().hd_directed_device(ctx.handle, int(variant), nd, pa, na, pb, nb,
                                              flags, seed, mh, C.byref(out)))
    return DirectedResult.from_c(out)


def directed_host(points_a, points_b, *, n_dims: Optional[int] = None, seed: int = 0,
                  shuffle: bool = True, parallelism: int = -1) -> DirectedResult:
    """h(A -> B) on the host CPU (the reference's CPU overload,
    early_break.h:44-115): host arrays, `parallelism` threads (<= 0: all
    hardware threads). Same bits, arg-max and NN as the GPU variants."""
    a = np.asarray(points_a)
    nd = n_dims or int(a.shape[1])
    pa, na, da, ka = _points(points_a, nd)
    pb, nb, db, kb = _points(points_b, nd)
    if da or db:
        raise HDError(_lib.HD_EINVAL, "directed_host takes host arrays")
    flags = 0 if shuffle else _lib.HD_FLAG_NO_SHUFFLE
    out = _lib.DirectedResultC()
    _lib.check(_lib.load().hd_directed_host(nd, pa, na, pb, nb, flags, seed, parallelism,
                                            C.byref(out)))
    return DirectedResult.from_c(out)


def nn_device(ctx: Context, index: Variant, queries, refs, out_d2, out_idx) -> None:
    """Exact NN of every query (CUDA tensors; out_d2 float32, out_idx int32)."""
    nd = int(queries.shape[1])
    pq, nq, _, _ = _points(queries, nd)
    pr, nr, _, _ = _points(refs, nd)
    _lib.check(_lib.load().hd_nn_device(ctx.handle, int(index), nd, pq, nq, pr, nr,
                                        C.c_void_p(out_d2.data_ptr()),
                                        C.c_void_p(out_idx.data_ptr())))


class HausdorffDistance:
    """Mirror of hd::HausdorffDistance<COORD_T, N_DIMS>
    (src/hausdorff_distance.h:15-42): calculate_distance returns sqrt(cmax2)
    as a float and fills get_stats() with Algorithm / Execution /
    ReportedTime. The two overloads of the reference are told apart by where
    the inputs live: torch CUDA tensors run on the GPU (:27-32), host arrays
    on the CPU (:21-26; only the early break has a CPU path, the other
    variants raise HD_EUNSUPPORTED as the reference aborts). Unlike the
    reference the inputs are not shuffled in place."""

    variant = Variant.EB
    algorithm = "Early Break"

    def __init__(self, ctx: Optional[Context] = None, seed: int = 0, n_threads: int = -1,
                 **_: Any):
        self._ctx = ctx
        self.seed = seed
        self.n_threads = n_threads
        self._stats: dict = {}
        self.last: Optional[DirectedResult] = None

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(0)
        return self._ctx

    def calculate_distance(self, points_a, points_b) -> float:
        if _is_torch_cuda(points_a) or _is_torch_cuda(points_b):
            r = directed_device(self.ctx, self.variant, points_a, points_b, seed=self.seed)
            execution = "GPU"
        elif self.variant == Variant.EB:
            r = directed_host(points_a, points_b, seed=self.seed, parallelism=self.n_threads)
            execution = "CPU"
        else:
            raise HDError(_lib.HD_EUNSUPPORTED,
                          f"CalculateDistance on the CPU is not implemented for {self.algorithm}")
        self.last = r
        self._stats = {"Algorithm": self.algorithm, "Execution": execution,
                       "ReportedTime": r.time_ms, "ComparedPairs": r.evals,
                       "CMax2Bits": r.cmax2_bits, "ArgMax": r.argmax, "NN": r.nn}
        return r.hd

    def get_stats(self) -> dict:
        for k in ("Algorithm", "Execution", "ReportedTime"):
            if k not in self._stats:
                raise HDError(_lib.HD_EINVAL, f"stats lack {k}: call calculate_distance first")
        return self._stats


class HausdorffDistanceEarlyBreak(HausdorffDistance):
    variant = Variant.EB
    algorithm = "Early Break"


class HausdorffDistanceGrid(HausdorffDistance):
    variant = Variant.GRID
    algorithm = "Uniform Grid"


class HausdorffDistanceLbvh(HausdorffDistance):
    variant = Variant.LBVH
    algorithm = "LBVH"


class HausdorffDistanceHybrid(HausdorffDistance):
    variant = Variant.HYBRID
    algorithm = "Hybrid"


class HausdorffDistanceRTHDist(HausdorffDistance):
    """The RT-HDIST baseline (hausdorff_distance_rt_hdist.h) on a quantized
    grid; its result is at least the exact HD."""
    variant = Variant.RT_HDIST
    algorithm = "RT-HDIST"


class HausdorffDistanceNearestNeighborSearch(HausdorffDistance):
    """The baseline of hausdorff_distance_nearest_neighbor_search.h: every A
    point's exact NN, no early break."""
    variant = Variant.NN
    algorithm = "Nearest Neighbor Search"


def batch_directed(ctx: Context, points_a, offs_a, points_b, offs_b, *,
                   shuffle: bool = True) -> "BatchResults":
    """Directed h(A_p -> B_p) for every pair p of a batch (CUDA tensors:
    points (N, d) float32, offsets (n_pairs + 1,) int64 in points). One
    persistent-kernel launch replayed from a cached HIP graph. shuffle=False
    scans each pair's B in input order (else its groups of 8 consecutive
    points in a golden-ratio rotation). Returns a BatchResults: result p is
    a DirectedResult, the
    fields of all pairs are numpy arrays."""
    nd = int(points_a.shape[1])
    pa, n_pts_a, da, _ = _points(points_a, nd)
    pb, n_pts_b, db, _ = _points(points_b, nd)
    if not (da and db):
        raise HDError(_lib.HD_EINVAL, "batch_directed needs CUDA tensors")
    for name, offs, n_pts in (("offs_a", offs_a, n_pts_a), ("offs_b", offs_b, n_pts_b)):
        if not _is_torch_cuda(offs) or str(offs.dtype) != "torch.int64" or offs.dim() != 1:
            raise HDError(_lib.HD_EINVAL, f"{name} must be a 1-D int64 CUDA tensor")
        if not offs.is_contiguous():
            raise HDError(_lib.HD_EINVAL, f"{name} must be contiguous")
        if offs.shape[0] < 2:
            raise HDError(_lib.HD_EINVAL, f"{name} needs n_pairs + 1 >= 2 entries")
        first, last = (int(v) for v in offs[[0, -1]].tolist())
        if first < 0 or last > n_pts:
            raise HDError(_lib.HD_EINVAL,
                          f"{name} spans [{first}, {last}) outside the {n_pts} points")
    if offs_a.shape[0] != offs_b.shape[0]:
        raise HDError(_lib.HD_EINVAL, "offs_a and offs_b must have the same length")
    n_pairs = int(offs_a.shape[0]) - 1
    out = (_lib.DirectedResultC * max(n_pairs, 1))()
    _lib.check(_lib.load().hd_batch_directed_device(
        ctx.handle, nd, pa, C.c_void_p(offs_a.data_ptr()), pb, C.c_void_p(offs_b.data_ptr()),
        n_pairs, 0 if shuffle else _lib.HD_FLAG_NO_SHUFFLE, out))
    return BatchResults(out, n_pairs)


class Model:
    """Ordering MLP (hd_mlp): d -> 64 -> 64 -> 1, BF16 weights, trained on
    the GPU on exact NN distances of a subsample of A."""

    def __init__(self, handle: C.c_void_p, n_dims: int, final_loss: float):
        self._h = handle
        self.n_dims = n_dims
        self.final_loss = final_loss

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @staticmethod
    def train(ctx: "Context", points_a, points_b, *, steps: int = 2000, batch: int = 4096,
              n_samples: int = 1_000_000, seed: int = 14, lr: float = 3e-3,
              label_index: Variant = Variant.GRID) -> "Model":
        lib = _lib.load()
        nd = int(points_a.shape[1])
        pa, na, da, _ = _points(points_a, nd)
        pb, nb, db, _ = _points(points_b, nd)
        if not (da and db):
            raise HDError(_lib.HD_EINVAL, "Model.train needs CUDA tensors")
        cfg = _lib.MlpTrainConfigC()
        lib.hd_mlp_train_config_init(C.byref(cfg))
        cfg.steps, cfg.batch, cfg.n_samples, cfg.seed = steps, batch, n_samples, seed
        cfg.lr = lr
        cfg.label_index = int(label_index)
        h = C.c_void_p()
        loss = C.c_float(0)
        _lib.check(lib.hd_mlp_train_device(ctx.handle, nd, pa, na, pb, nb, C.byref(cfg),
                                           C.byref(h), C.byref(loss)))
        return Model(h, nd, float(loss.value))

    def predict(self, ctx: "Context", points, out) -> None:
        p, n, _, _ = _points(points, self.n_dims)
        _lib.check(_lib.load().hd_mlp_predict_device(ctx.handle, self._h, p, n,
                                                     C.c_void_p(out.data_ptr())))

    def params(self) -> np.ndarray:
        lib = _lib.load()
        out = np.empty(lib.hd_mlp_n_params(self.n_dims), np.float32)
        _lib.check(lib.hd_mlp_get_params(self._h, out.ctypes.data_as(C.POINTER(C.c_float)),
                                         None, None))
        return out

    def normalisation(self) -> tuple[np.ndarray, np.ndarray]:
        """Input scaling of the network: x = (point - lower) * inv_extent."""
        lib = _lib.load()
        lo = np.zeros(3, np.float32)
        inv = np.zeros(3, np.float32)
        _lib.check(lib.hd_mlp_get_params(self._h, None,
                                         lo.ctypes.data_as(C.POINTER(C.c_float)),
                                         inv.ctypes.data_as(C.POINTER(C.c_float))))
        return lo[: self.n_dims], inv[: self.n_dims]

    def close(self) -> None:
        if self._h is not None:
            _lib.load().hd_mlp_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def mlp_loss_grad(ctx: "Context", params: np.ndarray, x: np.ndarray, y: np.ndarray):
    """One GPU forward+backward (weights rounded to BF16): (loss, grad)."""
    lib = _lib.load()
    p = np.ascontiguousarray(params, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    nd = int(x.shape[1])
    g = np.empty_like(p)
    loss = C.c_float(0)
    F = C.POINTER(C.c_float)
    _lib.check(lib.hd_mlp_loss_grad_device(ctx.handle, nd, p.ctypes.data_as(F),
                                           x.ctypes.data_as(F), y.ctypes.data_as(F),
                                           int(y.shape[0]), C.byref(loss), g.ctypes.data_as(F)))
    return float(loss.value), g


def device_count() -> int:
    return int(_lib.load().hd_device_count())


def device_info(device: int = 0) -> dict:
    buf = C.create_string_buffer(4096)
    _lib.check(_lib.load().hd_device_info(device, buf, 4096))
    return json.loads(buf.value.decode())
