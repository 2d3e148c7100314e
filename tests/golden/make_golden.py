"""Writes the golden fixtures in tests/golden/ (cases.npz + expected.json).

The reference ships no test vectors for this path (it has no tests; its only
check re-runs the CPU early break, run_hausdorff_distance.cu:366-388), so
the expected outputs here come from two independent sources:
  * closed forms for the constructed cases (lattices shifted by exactly
    representable offsets, single points, identical sets, overflow and
    denormal pairs), asserted below before anything is written;
  * the exhaustive max-min of the CPU oracle (no early break) for the
    random cases, cross-checked against the oracle's early-break path.
Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402

import hausdorffdistance_amd.gen as gen  # noqa: E402


def bits(x: float) -> int:
    return int(np.array([x], dtype=np.float32).view(np.uint32)[0])


def lattice(k: int, step: float) -> np.ndarray:
    """k x k lattice, x fastest (index = iy * k + ix)."""
    ys, xs = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    return (np.stack([xs.ravel(), ys.ravel()], 1) * step).astype(np.float32)


def cases() -> dict[str, tuple[np.ndarray, np.ndarray, dict]]:
    out: dict[str, tuple[np.ndarray, np.ndarray, dict]] = {}
    step = 1.0 / 16
    # Every point's NN is its own translate at exactly 1/64: all tie.
    a = lattice(16, step)
    b = a + np.float32(1.0 / 64)
    b[:, 1] = a[:, 1]
    out["lattice_shift_2d"] = (a, b, {"ab": {"cmax2_bits": bits(2.0 ** -12), "argmax": 0,
                                             "nn": 0}})
    # Translates of a few interior points removed: those points' NN moves to
    # the translate of their left neighbour at 3/64 -> d2 = 9/4096, a tie
    # among them; the smallest index wins.
    removed = [37, 21, 150]
    keep = np.ones(256, bool)
    keep[removed] = False
    b2 = b[keep]
    nn_of_21 = int(np.flatnonzero(keep)[:20].size)  # B index of the translate of a[20]
    out["lattice_ties_2d"] = (a, b2, {"ab": {"cmax2_bits": bits(9.0 / 4096), "argmax": 21,
                                             "nn": nn_of_21}})
    # Single A point: h = its NN distance.
    u = gen.uniform(11, 500, 2)
    p = np.array([[0.25, 0.75]], np.float32)
    d2 = ((u - p) ** 2).sum(1)
    out["single_point_a_2d"] = (p, u, {})
    out["single_point_b_2d"] = (u[:300], u[300:301], {})
    # Identical sets: every min is 0, smallest index wins, NN is itself.
    v = gen.uniform(12, 500, 3)
    out["identical_3d"] = (v, v.copy(), {"ab": {"cmax2_bits": 0, "argmax": 0, "nn": 0}})
    # Exact duplicates of the farthest point at indices 3 and 7.
    w = gen.uniform(13, 64, 2) * np.float32(0.1)
    far = np.array([0.9, 0.9], np.float32)
    w[3] = far
    w[7] = far
    out["duplicate_max_2d"] = (w, gen.uniform(14, 64, 2) * np.float32(0.1), {})
    # Overflowing pairs never count (d2 >= FLT_MAX); a point with no counting
    # pair is skipped, as the reference's FLT_MAX-initialised min does.
    ov_a = np.array([[0, 0], [3e19, 0], [1, 1]], np.float32)
    ov_b = np.array([[0, 0], [0.5, 0.5]], np.float32)
    out["overflow_2d"] = (ov_a, ov_b, {"ab": {"cmax2_bits": bits(np.float32(0.5)), "argmax": 2,
                                              "nn": 1}})
    # Denormal squared distances must not be flushed to zero.
    dn_a = np.array([[0, 0], [1e-20, 0]], np.float32)
    dn_b = np.array([[0, 0]], np.float32)
    den = np.float32(np.float32(1e-20) * np.float32(1e-20))
    assert 0 < den < np.finfo(np.float32).tiny
    out["denormal_2d"] = (dn_a, dn_b, {"ab": {"cmax2_bits": bits(den), "argmax": 1, "nn": 0}})
    # Random cases.
    out["uniform_2000_2d"] = (gen.uniform(1, 2000, 2), gen.uniform(2, 2000, 2), {})
    out["uniform_1500x900_3d"] = (gen.uniform(1, 1500, 3), gen.uniform(2, 900, 3), {})
    out["ragged_65x129_3d"] = (gen.uniform(5, 65, 3), gen.uniform(6, 129, 3), {})
    ga, gb = gen.gmm_pair(10_000, 64, 0.01, (0.002, 0.001), 0.01, (3, 4, 5, 6))
    out["gmm_10k_2d"] = (ga, gb, {})
    sa = gen.sphere(7, 3000, 0.005, 0.0)
    sb = gen.sphere(8, 2500, 0.005, 2.0)
    out["sphere_3000x2500_3d"] = (sa, sb, {})
    return out


def expected_for(a: np.ndarray, b: np.ndarray) -> dict:
    r = oracle.brute_directed(a, b)
    e = oracle.eb_directed(a, b, n_threads=4, seed=7)
    assert (r.cmax2_bits, r.argmax, r.nn) == (e.cmax2_bits, e.argmax, e.nn), "oracle paths disagree"
    return {"cmax2_bits": r.cmax2_bits, "argmax": r.argmax, "nn": r.nn}


def main() -> None:
    arrays = {}
    expected = {}
    for name, (a, b, known) in cases().items():
        ab = expected_for(a, b)
        ba = expected_for(b, a)
        for k, v in known.get("ab", {}).items():
            assert ab[k] == v, f"{name}: closed form {k}={v}, brute force {ab[k]}"
        arrays[f"{name}__a"] = a
        arrays[f"{name}__b"] = b
        expected[name] = {"n_dims": int(a.shape[1]), "ab": ab, "ba": ba}
    np.savez_compressed(HERE / "cases.npz", **arrays)
    (HERE / "expected.json").write_text(json.dumps(expected, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(expected)} cases")


if __name__ == "__main__":
    main()
