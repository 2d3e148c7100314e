"""How much can a visit order of A save the early break? (analysis script of
the oracle's early break; test infrastructure, not shipped).

The early break (reference early_break.h:44-115) scans B in a fixed
pseudo-random order for each A point and stops at the first B point whose
d2 is strictly below the current cmax2. A learned order of A only changes
*when* cmax2 reaches its final value; it cannot change how far a point must
scan once cmax2 is known. So the evaluations of the early break with the
final cmax2 known from the start are a lower bound for any order of A (per
point: the position of the first B point below the final cmax2, or |B| if
none), and the evaluations with a smaller cmax2 show what reaching the final
value late costs.

For a sample of A against all of B (the same shuffled order for every
point), this prints the mean evaluations per A point with cmax2 fixed at
fractions of the final value.

Run: python oracle/eb_bound.py [--n 10000000] [--sample 4000]
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path[:] = [p for p in sys.path if Path(p or ".").resolve() != HERE]
sys.path.insert(0, str(HERE.parent))

from oracle import oracle  # noqa: E402

import hausdorffdistance_amd.gen as gen  # noqa: E402


def first_below(px: np.float32, py: np.float32, bx: np.ndarray, by: np.ndarray,
                cmax2: np.float32, chunk: int = 1 << 16) -> int:
    """Evaluations until the first B point (in array order) with FP32 d2
    strictly below cmax2, with the reference's unfused formula; |B| if
    none."""
    n = bx.shape[0]
    for s in range(0, n, chunk):
        dx = px - bx[s:s + chunk]
        dy = py - by[s:s + chunk]
        d2 = dx * dx + dy * dy  # FP32: two products, one sum, each rounded
        hit = np.flatnonzero(d2 < cmax2)
        if hit.size:
            return s + int(hit[0]) + 1
    return n


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--sample", type=int, default=4000)
    args = ap.parse_args()
    a, b = gen.gmm_pair(args.n, 64, 0.01, (0.002, 0.001), 0.01, (3, 4, 5, 6))
    final = oracle.eb_directed(a, b).cmax2
    rng = np.random.default_rng(0)
    perm = rng.permutation(b.shape[0])
    bx = np.ascontiguousarray(b[perm, 0])
    by = np.ascontiguousarray(b[perm, 1])
    idx = rng.choice(a.shape[0], args.sample, replace=False)
    print(f"|A| = |B| = {args.n}, final cmax2 = {final:.6g}, sample of {args.sample} A points")
    for frac in (1.0, 0.9, 0.7, 0.5, 0.25):
        c = np.float32(final * frac)
        ev = np.array([first_below(a[i, 0], a[i, 1], bx, by, c) for i in idx], dtype=np.float64)
        print(f"cmax2 = {frac:4.2f} x final: {ev.mean():10.1f} evaluations per A point "
              f"(median {np.median(ev):8.0f}; {np.mean(ev >= b.shape[0]) * 100:.3f}% scan all of B)")


if __name__ == "__main__":
    main()
