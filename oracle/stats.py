"""CPU restatement of the reference's input statistics
(src/run_hausdorff_distance.cu:151-204).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Pieces and the reference lines each follows:
  * nonempty_mbr      -- mbr.h:288-296 (zero extent -> FP32(lower + 1e-5))
  * grid_dims         -- uniform_grid.h:222-244 (FP32 extents, FP64 side)
  * cell_counts       -- uniform_grid.h:138-159 cell position in FP32,
                         EncodeCellPos (:37-48, x fastest)
  * gini              -- running_stats.h:32-57 over the sorted non-empty
                         counts (here in FP64)
  * hdr_percentiles   -- uniform_grid.h:675-692 and running_stats.h:106-140:
                         hdr_init(1, n_nonempty, 3), hdr_record_value per
                         cell, hdr_percentiles_print(..., 3, 1.0, CLASSIC)
                         parsed back. HdrHistogram_c is a system library the
                         reference finds with cmake/FindHdrHistogram.cmake
                         (no version pinned, not vendored); this follows its
                         published hdr_histogram.c: the counts array of
                         (bucket_count + 1) * sub_bucket_half_count entries,
                         value_at_index / highest_equivalent_value, and the
                         percentile iterator's level rule.
  * input_stats       -- one "Files" entry / the "Input" object.
"""
from __future__ import annotations

import math

import numpy as np


def nonempty_mbr(points: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    pts = np.asarray(points, np.float32)
    lo = pts.min(axis=0).astype(np.float32)
    hi = pts.max(axis=0).astype(np.float32)
    for d in range(pts.shape[1]):
        if np.float32(hi[d] - lo[d]) == 0:
            hi[d] = np.float32(float(lo[d]) + 1e-5)
    return lo, hi


def mbr_volume32(lo: np.ndarray, hi: np.ndarray) -> np.float32:
    v = np.float32(1)
    for d in range(len(lo)):
        v = np.float32(v * np.float32(hi[d] - lo[d]))
    return v


def grid_dims(lo: np.ndarray, hi: np.ndarray, n: int, npc: int) -> list[int]:
    nd = len(lo)
    ext = [float(np.float32(hi[d] - lo[d])) for d in range(nd)]
    volume = 1.0
    for e in ext:
        volume *= e
    q = volume * npc / n
    s = math.sqrt(q) if nd == 2 else float(np.cbrt(np.float64(q)))
    return [int(math.ceil(e / s)) for e in ext]


def cell_counts(sets: list[np.ndarray], lo, hi, dims) -> np.ndarray:
    """Points per cell of the dense table, all sets counted together."""
    nd = len(dims)
    counts = np.zeros(int(np.prod(dims, dtype=np.uint64)), np.uint64)
    for pts in sets:
        pts = np.asarray(pts, np.float32)
        cell = np.zeros(len(pts), np.uint64)
        mul = 1
        for d in range(nd):
            norm = (pts[:, d] - np.float32(lo[d])) / np.float32(np.float32(hi[d]) - np.float32(lo[d]))
            gs = np.float32(dims[d])
            x = np.minimum(np.maximum(norm.astype(np.float32) * gs, np.float32(0)),
                           np.float32(gs - np.float32(1)))
            cell += x.astype(np.uint32).astype(np.uint64) * np.uint64(mul)
            mul *= dims[d]
        np.add.at(counts, cell, 1)
    return counts


def gini(values) -> float:
    """sum_i (2i - n + 1) x_i / (n sum x) over the values sorted ascending."""
    x = np.sort(np.asarray(values, np.float64))
    n = len(x)
    if n == 0 or x.sum() == 0:
        return 0.0
    i = np.arange(n, dtype=np.float64)
    return float(((2 * i - n + 1) * x).sum() / (n * x.sum()))


class Hdr:
    """HdrHistogram_c's layout for lowest value 1 and 3 significant figures."""

    def __init__(self, highest: int):
        assert highest >= 2, "hdr_init needs highest >= 2 * lowest"
        largest_single = 2 * 10 ** 3
        self.sub_bucket_count_mag = math.ceil(math.log2(largest_single))  # 11
        self.sub_bucket_half_count_mag = self.sub_bucket_count_mag - 1
        self.sub_bucket_count = 1 << self.sub_bucket_count_mag
        self.sub_bucket_half_count = self.sub_bucket_count // 2
        self.sub_bucket_mask = self.sub_bucket_count - 1
        smallest_untrackable = self.sub_bucket_count
        buckets = 1
        while smallest_untrackable <= highest:
            smallest_untrackable <<= 1
            buckets += 1
        self.bucket_count = buckets
        self.counts = [0] * ((buckets + 1) * self.sub_bucket_half_count)
        self.total = 0

    def counts_index(self, v: int) -> int:
        b = (v | self.sub_bucket_mask).bit_length() - self.sub_bucket_count_mag
        sb = v >> b
        return ((b + 1) << self.sub_bucket_half_count_mag) + (sb - self.sub_bucket_half_count)

    def record(self, v: int) -> bool:
        i = self.counts_index(v)
        if i < 0 or i >= len(self.counts):
            return False
        self.counts[i] += 1
        self.total += 1
        return True

    def value_at_index(self, i: int) -> int:
        b = (i >> self.sub_bucket_half_count_mag) - 1
        sb = (i & (self.sub_bucket_half_count - 1)) + self.sub_bucket_half_count
        if b < 0:
            sb -= self.sub_bucket_half_count
            b = 0
        return sb << b

    def highest_equivalent(self, v: int) -> int:
        b = (v | self.sub_bucket_mask).bit_length() - self.sub_bucket_count_mag
        sb = v >> b
        adjusted = b + 1 if sb >= self.sub_bucket_count else b
        lowest = sb << b
        return lowest + (1 << adjusted) - 1

    def percentiles(self, ticks_per_half_distance: int = 3) -> list[tuple[float, float, int]]:
        """Rows (value, percentile / 100 as printed with %f, cumulative count)
        of the percentile iterator, one per call of its next()."""
        it = {"idx": -1, "cum": 0, "count": 0, "heq": 0, "to": 0.0, "pct": 0.0,
              "seen_last": False}

        def has_next() -> bool:
            return it["cum"] < self.total

        def basic_next() -> bool:
            if not has_next():
                return False
            it["idx"] += 1
            if it["idx"] >= len(self.counts):
                return False
            it["count"] = self.counts[it["idx"]]
            it["cum"] += it["count"]
            it["heq"] = self.highest_equivalent(self.value_at_index(it["idx"]))
            return True

        def percentile_next() -> bool:
            if not has_next():
                if it["seen_last"]:
                    return False
                it["seen_last"] = True
                it["pct"] = 100.0
                return True
            if it["idx"] == -1 and not basic_next():
                return False
            while True:
                current = (100.0 * it["cum"]) / self.total
                if it["count"] != 0 and it["to"] <= current:
                    it["pct"] = it["to"]
                    temp = int(math.log(100 / (100.0 - it["to"])) / math.log(2)) + 1
                    half = int(math.pow(2, temp))
                    it["to"] += 100.0 / (ticks_per_half_distance * half)
                    return True
                if not basic_next():
                    return True

        rows = []
        while percentile_next():
            rows.append((float(it["heq"]), float("%f" % (it["pct"] / 100.0)), it["cum"]))
        return rows


def hdr_percentiles(values, highest: int) -> list[tuple[float, float, int]]:
    """The reference's "Histogram" list for the given per-cell counts."""
    h = Hdr(max(int(highest), 2))
    for v in values:
        h.record(int(v))
    return h.percentiles()


def input_stats(sets: list[np.ndarray], npc: int) -> dict:
    """One stats object: NumPoints, MBR, Density, Grid, StatsGridNumPointsPerCell."""
    allp = np.concatenate([np.asarray(s, np.float32) for s in sets])
    lo, hi = nonempty_mbr(allp)
    n = len(allp)
    out = {"NumPoints": n,
           "MBR": [{"Lower": float(lo[d]), "Upper": float(hi[d])} for d in range(len(lo))],
           "Density": float(np.float32(np.float32(n) / mbr_volume32(lo, hi)))}
    if npc <= 0:
        return out
    dims = grid_dims(lo, hi, n, npc)
    counts = cell_counts(sets, lo, hi, dims)
    nz = counts[counts > 0]
    out["Grid"] = {"TotalCells": int(np.prod(dims)), "NonEmptyCells": int(len(nz)),
                   "MaxPoints": int(nz.max()), "GiniIndex": gini(nz), "GridSize": dims,
                   "Histogram": [{"value": v, "percentile": p, "count": c}
                                 for v, p, c in hdr_percentiles(nz, len(nz))]}
    out["StatsGridNumPointsPerCell"] = npc
    return out
