"""Input statistics of the run's stats JSON on the GPU against the oracle
(oracle/stats.py): per-file and union MBR, density, stats-grid size, cell
counts (total, non-empty, largest), Gini index and the percentile list
(src/run_hausdorff_distance.cu:151-204). Counts, sizes, MBRs, densities and
the list are exact; the Gini index is within 1e-12 relative (both FP64)."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd.build import EXE
from oracle import stats as S

pytestmark = pytest.mark.gpu


def histogram_rows(obj: dict) -> list:
    return [(r["value"], r["percentile"], r["count"]) for r in obj["Grid"]["Histogram"]]


def check_object(got: dict, want: dict) -> None:
    assert got["NumPoints"] == want["NumPoints"]
    assert got["MBR"] == want["MBR"]
    assert got["StatsGridNumPointsPerCell"] == want["StatsGridNumPointsPerCell"]
    g, w = got["Grid"], want["Grid"]
    for k in ("TotalCells", "NonEmptyCells", "MaxPoints", "GridSize"):
        assert g[k] == w[k], k
    assert g["GiniIndex"] == pytest.approx(w["GiniIndex"], rel=1e-12, abs=1e-15)
    assert histogram_rows(got) == [(r["value"], r["percentile"], r["count"])
                                   for r in w["Histogram"]]


def run(a: np.ndarray, b: np.ndarray, ctx, npc: int = 10) -> dict:
    cfg = hd.RunConfig(n_dims=a.shape[1], repeat=1, stats_n_points_cell=npc)
    print(f"stats run n_a={len(a)} n_b={len(b)} nd={a.shape[1]} npc={npc}", flush=True)
    return hd.run_hausdorff_distance(cfg, a, b, ctx=ctx).stats["Input"]


CASES = {
    "uniform_2d": lambda: (hd.gen.uniform(1, 30000, 2), hd.gen.uniform(2, 20000, 2)),
    "gmm_2d": lambda: hd.gen.gmm_pair(40000, n_clusters=16, seeds=(3, 4, 5, 6)),
    "skewed_2d": lambda: (hd.gen.uniform(4, 5000, 2) ** 4, hd.gen.uniform(5, 7000, 2) ** 3),
    "uniform_3d": lambda: (hd.gen.uniform(6, 25000, 3), hd.gen.uniform(7, 12000, 3) * 0.5),
    # y is constant: the non-empty MBR rule gives that axis extent 1e-5.
    "flat_axis": lambda: (np.c_[hd.gen.uniform(8, 3000, 2)[:, :1], np.full(3000, 2.0)],
                          np.c_[hd.gen.uniform(9, 2000, 2)[:, :1], np.full(2000, 2.0)]),
    "tiny": lambda: (hd.gen.uniform(10, 5, 2), hd.gen.uniform(11, 3, 2)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_input_stats_match_oracle(ctx, case):
    a, b = (np.ascontiguousarray(x, np.float32) for x in CASES[case]())
    got = run(a, b, ctx)
    check_object(got["Files"][0], S.input_stats([a], 10))
    check_object(got["Files"][1], S.input_stats([b], 10))
    both = S.input_stats([a, b], 10)
    check_object(got, both)
    # The top-level density is over the union of the two non-empty MBRs.
    la, ha = S.nonempty_mbr(a)
    lb, hb = S.nonempty_mbr(b)
    vol = S.mbr_volume32(np.minimum(la, lb), np.maximum(ha, hb))
    assert got["Density"] == float(np.float32(np.float32(len(a) + len(b)) / vol))
    for f, p in zip(got["Files"], (a, b)):
        assert f["Density"] == S.input_stats([p], 0)["Density"]


def test_input_stats_off_and_other_cell_sizes(ctx):
    a = np.ascontiguousarray(hd.gen.uniform(12, 20000, 2), np.float32)
    b = np.ascontiguousarray(hd.gen.uniform(13, 9000, 2), np.float32)
    off = run(a, b, ctx, npc=0)
    assert "Grid" not in off and "Grid" not in off["Files"][0]
    assert off["Files"][0]["MBR"] == S.input_stats([a], 0)["MBR"]
    got = run(a, b, ctx, npc=3)
    check_object(got, S.input_stats([a, b], 3))


def test_cli_json_has_input_options(tmp_path):
    out = tmp_path / "s.json"
    r = subprocess.run([str(EXE), "-input1", "uniform_100k_2d", "-input_type", "synthetic",
                        "-n_dims", "2", "-repeat", "1", "-limit", "12000",
                        "-json", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    js = json.loads(out.read_text())
    inp = js["Input"]
    assert inp["Normalize"] is False and inp["Translate"] == 0 and inp["Limit"] == 12000
    assert inp["SerializationPrefix"] == ""
    assert inp["Grid"]["NonEmptyCells"] > 0 and len(inp["Files"]) == 2


def test_repeat_bounds_and_feature_vector(ctx):
    a = np.ascontiguousarray(hd.gen.uniform(14, 20000, 2) * np.float32(0.6), np.float32)
    b = np.ascontiguousarray(hd.gen.uniform(15, 15000, 2) * np.float32(0.6) + np.float32(0.3),
                             np.float32)
    cfg = hd.RunConfig(n_dims=2, repeat=2, stats_n_points_cell=10)
    res = hd.run_hausdorff_distance(cfg, a, b, ctx=ctx)
    for rep in res.stats["Running"]["Repeats"]:
        assert rep["HDLowerBound"] <= res.hd * (1 + 1e-6)
        assert res.hd <= rep["HDUpperBound"] * (1 + 1e-6)
    v = hd.features.feature_vector(res.stats, repeat=1)
    assert len(v) == hd.features.N_FEATURES and all(np.isfinite(v))
    assert v[1] == 20000 and v[24] == 15000  # NumPoints of each file


def test_cli_vary_params_sweep(tmp_path):
    from hausdorffdistance_amd import tuning

    log = tmp_path / "train" / "n_points_cell" / "pair"
    log.parent.mkdir(parents=True)
    args = [str(EXE), "-input1", "uniform_100k_2d", "-input_type", "synthetic", "-n_dims", "2",
            "-limit", "30000", "-repeat", "2", "-vary_params", "-n_points_cell_list", "4,8",
            "-max_hit_list", "64,256", "-json", str(log)]
    print("sweep", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    files = sorted(log.parent.glob("pair_sample_rate_*.json"))
    assert [p.name for p in files] == sorted(
        f"pair_sample_rate_0.010000_eb_only_threshold_0_max_hit_{h}_n_points_cell_{c}.json"
        for h in (64, 256) for c in (4, 8))
    runs = [json.loads(p.read_text()) for p in files]
    hds = {run["HDResult"] for run in runs}
    assert len(hds) == 1  # every combination gives the same distance
    for run in runs:
        assert run["Input"]["Files"][0]["Path"] == "uniform_100k_2d"
        assert len(run["Running"]["Repeats"]) == 2 and "DateTime" in run
    # A second sweep skips what exists.
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.count("Skip ") == 4
    rows = tuning.training_rows(tuning.load_runs(str(tmp_path / "train"), "n_points_cell"))
    assert len(rows) == 1
    assert rows[0]["AvgTime"] == min(run["Running"]["AvgTime"] for run in runs)
