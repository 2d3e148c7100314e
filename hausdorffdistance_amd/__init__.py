"""MI355X-native exact Hausdorff distance (directed and symmetric).

The compute lives in libhd.so (hand-written HIP for gfx950, C ABI in
include/hd.h); this package is its Python host mirror of the reference's run
interface (RunConfig / RunHausdorffDistance / HausdorffDistance).
"""
from ._lib import HDError, LIB_PATH, load  # noqa: F401
from .api import (  # noqa: F401
    Context,
    DirectedResult,
    Execution,
    HausdorffDistance,
    HausdorffDistanceEarlyBreak,
    HausdorffDistanceGrid,
    HausdorffDistanceHybrid,
    HausdorffDistanceLbvh,
    HausdorffDistanceNearestNeighborSearch,
    HausdorffDistanceRTHDist,
    Model,
    Result,
    RunConfig,
    Variant,
    batch_directed,
    device_count,
    device_info,
    directed_device,
    directed_host,
    mlp_loss_grad,
    nn_device,
    run_hausdorff_distance,
)
from . import features, gen  # noqa: F401
from .io import load_points, normalize_points, translate_points  # noqa: F401
