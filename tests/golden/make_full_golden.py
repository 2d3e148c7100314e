"""Writes tests/golden/full_size.json: the CPU oracle's results on the five
benchmark workloads at their full sizes.

The reference checks every run against its CPU early break
(run_hausdorff_distance.cu:366-388); these are the same known answers,
computed once here (the oracle at full size takes minutes on a few cores,
too long for a GPU test) and compared by tests/test_gpu_full_size.py with
every GPU variant on the same generated inputs.

Workloads (hausdorffdistance_amd.gen, fixed seeds):
  uniform_100k_2d     A, B 100,000 uniform points (seeds 1, 2); both ways
  gmm_10m_2d          10M x 10M 64-component mixture, B = A + (0.002, 0.001)
                      with 1% uniform outliers (seeds 3-6); both ways
  sphere_20m_15m_3d   20M / 15M noisy unit-sphere points, B rotated 2 degrees
                      about z (seeds 7, 8); both ways
  batch_4096_2d       4096 pairs, sizes log-uniform in [32, 65536) (seed 9);
                      every pair, A -> B
  gmm_50m_256_2d      50M x 50M 256-component mixture (seeds 10-13); A -> B

Each directed entry is (cmax2_bits, argmax, nn). The oracle's early break
gives them; on the 100k pair the exhaustive max-min confirms them.
Run: python tests/golden/make_full_golden.py [--only NAME ...] [--threads N]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402

import hausdorffdistance_amd.gen as gen  # noqa: E402

OUT = HERE / "full_size.json"


def digest(*arrays: np.ndarray) -> str:
    """SHA-256 of the generated inputs, so a test can tell a generator change
    from a kernel bug."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def entry(r) -> dict:
    return {"cmax2_bits": int(r.cmax2_bits), "argmax": int(r.argmax), "nn": int(r.nn)}


def directed(a, b, threads: int, name: str) -> dict:
    t0 = time.time()
    r = oracle.eb_directed(a, b, n_threads=threads, seed=0)
    print(f"[golden] {name}: bits {r.cmax2_bits} argmax {r.argmax} nn {r.nn} "
          f"({r.compared_pairs} evals, {time.time() - t0:.1f} s)", flush=True)
    return entry(r)


def uniform_100k(threads: int) -> dict:
    a, b = gen.uniform_pair(100_000, 2)
    ab = directed(a, b, threads, "uniform_100k_2d A->B")
    ba = directed(b, a, threads, "uniform_100k_2d B->A")
    assert entry(oracle.brute_directed(a, b, threads)) == ab
    assert entry(oracle.brute_directed(b, a, threads)) == ba
    return {"n_a": 100_000, "n_b": 100_000, "digest": digest(a, b), "ab": ab, "ba": ba}


def gmm_10m(threads: int) -> dict:
    a, b = gen.gmm_10m()
    return {"n_a": len(a), "n_b": len(b), "digest": digest(a, b),
            "ab": directed(a, b, threads, "gmm_10m_2d A->B"),
            "ba": directed(b, a, threads, "gmm_10m_2d B->A")}


def sphere(threads: int) -> dict:
    a, b = gen.sphere_pair()
    return {"n_a": len(a), "n_b": len(b), "digest": digest(a, b),
            "ab": directed(a, b, threads, "sphere_20m_15m_3d A->B"),
            "ba": directed(b, a, threads, "sphere_20m_15m_3d B->A")}


def batch(threads: int) -> dict:
    bt = gen.batch(9, 4096)
    pairs = []
    t0 = time.time()
    for p in range(4096):
        a = bt.a[bt.offs_a[p]:bt.offs_a[p + 1]]
        b = bt.b[bt.offs_b[p]:bt.offs_b[p + 1]]
        r = oracle.eb_directed(a, b, n_threads=threads, seed=0)
        pairs.append([int(r.cmax2_bits), int(r.argmax), int(r.nn)])
    print(f"[golden] batch_4096_2d: 4096 pairs in {time.time() - t0:.1f} s", flush=True)
    return {"n_pairs": 4096, "total_a": int(bt.offs_a[-1]), "total_b": int(bt.offs_b[-1]),
            "digest": digest(bt.a, bt.b, bt.offs_a, bt.offs_b),
            "pairs": pairs}


def gmm_50m(threads: int) -> dict:
    a, b = gen.gmm_50m_256()
    return {"n_a": len(a), "n_b": len(b), "digest": digest(a, b),
            "ab": directed(a, b, threads, "gmm_50m_256_2d A->B")}


WORKLOADS = {
    "uniform_100k_2d": uniform_100k,
    "gmm_10m_2d": gmm_10m,
    "sphere_20m_15m_3d": sphere,
    "batch_4096_2d": batch,
    "gmm_50m_256_2d": gmm_50m,
}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    for name in args.only:
        data[name] = WORKLOADS[name](args.threads)
        OUT.write_text(json.dumps(data, sort_keys=True, separators=(",", ":")) + "\n")
    print(f"[golden] wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
