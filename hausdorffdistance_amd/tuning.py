"""Training data for a parameter selector from `hd_exec -vary_params` logs.

The reference sweeps the hybrid's parameters per input pair
(src/autotune_hausdorff_distance.cu:181-296, driven by
expr/for_the_paper/gen_train.sh), keeps the fastest combination per pair
(expr/for_the_paper/train.py:229-233: the pair is identified by both file
paths and the translation) and fits decision trees on the statistics of the
inputs. This module does the first two steps on the JSON files the sweep
writes and exports one row per pair: the selector's feature vector
(`features.feature_vector`, the layout the hybrid reads at run time) and
the winning parameters. Fitting the trees is left to the caller (the
reference uses xgboost, which is not a dependency here).

    python -m hausdorffdistance_amd.tuning LOG_DIR OUT.csv [PARAM_DIR_NAME]
"""
from __future__ import annotations

import csv
import json
import os
import sys
from typing import Iterable, Optional

from .features import N_FEATURES, feature_vector

LABELS = ("SampleRate", "EBOnlyThreshold", "MaxHit", "NumPointsPerCell")


def load_runs(directory: str, param_name: Optional[str] = None) -> list[dict]:
    """Every non-empty .json under `directory`; with `param_name`, only those
    whose folder has that name (train.py:20-45)."""
    runs = []
    for root, _, files in os.walk(directory):
        if param_name is not None and os.path.basename(root) != param_name:
            continue
        for name in sorted(files):
            path = os.path.join(root, name)
            if name.lower().endswith(".json") and os.path.getsize(path) > 0:
                with open(path) as f:
                    runs.append(json.load(f))
    return runs


def run_id(run: dict) -> str:
    files = run["Input"]["Files"]
    return f'{files[0].get("Path", "")}-{files[1].get("Path", "")}-{run["Input"].get("Translate", 0)}'


def best_params(runs: Iterable[dict]) -> list[dict]:
    """The run with the smallest Running.AvgTime per input pair (first one
    on ties), in order of first appearance."""
    best: dict[str, dict] = {}
    for r in runs:
        key = run_id(r)
        if key not in best or r["Running"]["AvgTime"] < best[key]["Running"]["AvgTime"]:
            best[key] = r
    return list(best.values())


def training_rows(runs: Iterable[dict]) -> list[dict]:
    rows = []
    for r in best_params(runs):
        row = {"RunID": run_id(r), "AvgTime": r["Running"]["AvgTime"]}
        for k in LABELS:
            row[k] = r["Running"][k]
        for i, v in enumerate(feature_vector(r)):
            row[f"f{i}"] = v
        rows.append(row)
    return rows


def export_csv(rows: list[dict], path: str) -> None:
    cols = ["RunID", "AvgTime", *LABELS] + [f"f{i}" for i in range(N_FEATURES)]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for row in rows:
            w.writerow(row)


def main(argv: list[str]) -> int:
    if len(argv) not in (2, 3):
        print(__doc__.strip().splitlines()[-1].strip(), file=sys.stderr)
        return 1
    rows = training_rows(load_runs(argv[0], argv[2] if len(argv) == 3 else None))
    export_csv(rows, argv[1])
    print(f"{len(rows)} input pairs -> {argv[1]}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
