"""Selector feature vector from a run's stats JSON (src/models/features.h).

The reference's hybrid turns the "Input" statistics and the repeat's MBR
bounds into 63 doubles that its generated decision trees read
(hausdorff_distance_hybrid.h:713-769). This module builds the same vector
from the stats JSON of `run_hausdorff_distance`, so tooling that trains or
evaluates such a selector can use runs of this library. Layout, in order:

  per file (2 x 23): Density, NumPoints, MBR_Range[3], GiniIndex,
      P99/P95/P75/P50/P25/P10 Value, the same six Counts, GridSize[3],
      NonEmptyCells, TotalCells
  union histogram (15): the six Values, the six Counts, GridSize[3]
  runtime (2): HDLowerBound / half the MBR diagonal, HDUpperBound / same

Percentile picks follow features.h:95-118: walking the histogram from its
last row down, a row sets the first still-unset percentile of the chain
P99, P95, P75, P50, P25, P10 whose threshold it is below (one per row).
A percentile no row reaches keeps the reference's -1 value; entries the
reference leaves uninitialised (its count, the third axis in 2D) are 0
here.
"""
from __future__ import annotations

import math
from typing import Sequence

N_FEATURES = 63
_LEVELS = (0.99, 0.95, 0.75, 0.50, 0.25, 0.10)


def _histogram_block(grid: dict, n_dims: int) -> list[float]:
    values = [-1.0] * 6
    counts = [0.0] * 6
    rows = grid["Histogram"]
    for row in reversed(rows):
        pct = float(row["percentile"])
        # The else-if chain: the first level that is unset and whose
        # threshold the row is below takes the row.
        for k, level in enumerate(_LEVELS):
            if values[k] == -1.0 and pct < level:
                values[k] = float(row["value"])
                counts[k] = float(row["count"])
                break
    size = [0.0, 0.0, 0.0]
    for d in range(n_dims):
        size[d] = float(grid["GridSize"][d])
    return values + counts + size


def static_features(file_stats: dict, n_dims: int) -> list[float]:
    rng = [0.0, 0.0, 0.0]
    for d in range(n_dims):
        m = file_stats["MBR"][d]
        rng[d] = float(m["Upper"]) - float(m["Lower"])
    g = file_stats["Grid"]
    return ([float(file_stats["Density"]), float(file_stats["NumPoints"])] + rng +
            [float(g["GiniIndex"])] + _histogram_block(g, n_dims) +
            [float(g["NonEmptyCells"]), float(g["TotalCells"])])


def runtime_features(files: Sequence[dict], repeat: dict, n_dims: int) -> list[float]:
    diag2 = 0.0
    for f in files:
        for d in range(n_dims):
            r = float(f["MBR"][d]["Upper"]) - float(f["MBR"][d]["Lower"])
            diag2 += r * r
    half = math.sqrt(diag2) / 2
    return [float(repeat["HDLowerBound"]) / half, float(repeat["HDUpperBound"]) / half]


def feature_vector(stats: dict, repeat: int = 0) -> list[float]:
    """The 63 features of a stats JSON (needs stats_n_points_cell > 0)."""
    inp = stats["Input"]
    nd = int(inp["NumDims"])
    files = inp["Files"]
    out = static_features(files[0], nd) + static_features(files[1], nd)
    out += _histogram_block(inp["Grid"], nd)
    out += runtime_features(files, stats["Running"]["Repeats"][repeat], nd)
    assert len(out) == N_FEATURES
    return out
