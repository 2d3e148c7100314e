"""CPU restatement of the RT-HDIST baseline (reference
src/hd_impl/hausdorff_distance_rt_hdist.h:45-222, the OptiX programs of
rt/shaders/shaders_nn_quantized_grid_2d.cu:1-136 and the grid of
src/index/quantized_grid.h:91-103,142-158,290-296), launch by launch.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Differences from the reference kept on purpose (the library does the same):
the centres are summed in FP64 (the reference sums FP32 in thrust's order),
the MBR of B is made non-empty, and B points are clamped into the grid.
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32


def _d2(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """FP32 squared distances, left to right, unfused (distance.h:17-24)."""
    acc = F(0) + (p[..., 0] - q[..., 0]) * (p[..., 0] - q[..., 0])
    for d in range(1, p.shape[-1]):
        t = (p[..., d] - q[..., d]).astype(F)
        acc = (acc + (t * t).astype(F)).astype(F)
    return acc.astype(F)


def rt_hdist(a: np.ndarray, b: np.ndarray) -> tuple[int, int]:
    """(cmax2 bits, arg-max A index) of the launch loop."""
    a = np.asarray(a, F)
    b = np.asarray(b, F)
    nd = a.shape[1]
    ca = (a.astype(np.float64).sum(0) / len(a)).astype(F)
    cb = (b.astype(np.float64).sum(0) / len(b)).astype(F)
    cd2 = F(0)
    for d in range(nd):
        t = F(ca[d] - cb[d])
        cd2 = F(cd2 + F(t * t))
    half = F(np.sqrt(cd2) / F(2))
    lo = b.min(0)
    hi = b.max(0)
    for d in range(nd):
        if F(hi[d] - lo[d]) == 0:
            hi[d] = F(float(lo[d]) + 1e-5)
    longest = max(F(hi[d] - lo[d]) for d in range(nd))
    cell = np.nextafter(F(longest / F(128)), F(longest))
    if not cell > 0:
        cell = F(1)
    res = [max(1, int(math.ceil(F(F(hi[d] - lo[d]) / cell)))) for d in range(nd)]
    diag_q = F(math.sqrt(nd))
    diag = F(math.sqrt(nd) * float(cell))
    radius = F(F(math.ceil(F(half / diag))) * diag_q)

    def quantize(p):
        t = ((p - lo).astype(F) / cell).astype(F)
        return np.clip(t, F(-2 ** 30), F(2 ** 30)).astype(np.int64)

    qb = np.clip(quantize(b), 0, np.array(res) - 1)
    cells = {}
    for j, c in enumerate(map(tuple, qb)):
        cells.setdefault(c, []).append(j)
    cell_pos = np.array(list(cells.keys()), np.int64)
    members = list(cells.values())
    qa = quantize(a)

    queue = [(i, 0) for i in range(len(a))]
    cmin_final = np.full(len(a), np.inf, np.float64)
    while queue:
        hits, misses = [], []
        r2 = F(radius * radius)
        for i, it in queue:
            dist2 = ((cell_pos - qa[i]) ** 2).sum(1).astype(F)  # exact ints, then FP32
            within = np.nonzero(dist2 <= r2)[0]
            if len(within) == 0:
                misses.append(i)
                continue
            if it < 3:
                hits.append((i, it + 1))
                continue
            idx = np.concatenate([members[k] for k in within])
            cmin_final[i] = float(_d2(a[i][None, :], b[idx]).min())
        queue = hits + [(i, 0) for i in misses]
        if queue:
            radius = F(radius + diag_q)
    done = np.isfinite(cmin_final)
    cmax2 = F(cmin_final[done].max())
    arg = int(np.nonzero(cmin_final.astype(F) == cmax2)[0][0])
    return int(np.array(cmax2, F).view(np.uint32)), arg
