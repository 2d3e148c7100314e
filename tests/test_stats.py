"""Input statistics without a GPU: the Gini index and the HdrHistogram
percentile list of the library's host code (hd_stats_gini,
hd_stats_percentiles) and of the oracle (oracle/stats.py), pinned by closed
forms, hand-derived lists and a brute-force identity, then compared with each
other on random histograms; and the oracle's stats grid on inputs whose cells
are known by hand (uniform_grid.h:138-159, :222-244; mbr.h:288-296)."""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pytest

from hausdorffdistance_amd import _lib
from oracle import stats as S


def lib_gini(hist) -> float:
    h = np.ascontiguousarray(hist, np.uint64)
    out = C.c_double()
    rc = _lib.load().hd_stats_gini(h.ctypes.data_as(_lib._U64), len(h), C.byref(out))
    assert rc == 0
    return out.value


def lib_percentiles(hist) -> list[tuple[float, float, int]]:
    h = np.ascontiguousarray(hist, np.uint64)
    buf = C.create_string_buffer(1 << 20)
    rc = _lib.load().hd_stats_percentiles(h.ctypes.data_as(_lib._U64), len(h), buf, len(buf))
    assert rc == 0
    return [(r["value"], r["percentile"], r["count"]) for r in json.loads(buf.value.decode())]


def hist_of(values) -> np.ndarray:
    v = np.asarray(values, np.int64)
    return np.bincount(v, minlength=int(v.max()) + 1).astype(np.uint64)


def brute_gini(values) -> float:
    # Mean absolute difference over twice the mean: sum_ij |x_i - x_j| / (2 n sum x).
    x = np.asarray(values, np.float64)
    return float(np.abs(x[:, None] - x[None, :]).sum() / (2 * len(x) * x.sum()))


def test_gini_closed_forms():
    for values, want in (([1, 3], 0.25), ([5, 5, 5, 5], 0.0), ([7], 0.0),
                         ([1, 1, 1, 5], 0.375)):
        assert S.gini(values) == pytest.approx(want, abs=1e-15)
        assert lib_gini(hist_of(values)) == pytest.approx(want, abs=1e-15)
    assert lib_gini(np.zeros(4, np.uint64)) == 0.0


def test_gini_matches_mean_absolute_difference():
    rng = np.random.default_rng(5)
    for _ in range(5):
        values = rng.integers(1, 60, size=rng.integers(2, 400))
        want = brute_gini(values)
        assert S.gini(values) == pytest.approx(want, rel=1e-12)
        assert lib_gini(hist_of(values)) == pytest.approx(want, rel=1e-12)


# Seven cells holding 1..7 points; levels 0, 1/6, 1/3, 1/2, 7/12, 2/3, 3/4,
# 19/24, 5/6, 7/8 (three per halving of the distance to 1), derived by hand.
SEVEN = [(1.0, 0.0, 1), (2.0, 0.166667, 2), (3.0, 0.333333, 3), (4.0, 0.5, 4),
         (5.0, 0.583333, 5), (5.0, 0.666667, 5), (6.0, 0.75, 6), (6.0, 0.791667, 6),
         (6.0, 0.833333, 6), (7.0, 0.875, 7), (7.0, 1.0, 7)]


def test_percentiles_hand_derived():
    values = [1, 2, 3, 4, 5, 6, 7]
    assert S.hdr_percentiles(values, len(values)) == SEVEN
    assert lib_percentiles(hist_of(values)) == SEVEN
    # Four equal cells: the first level, then 100%.
    four = [(4.0, 0.0, 4), (4.0, 1.0, 4)]
    assert S.hdr_percentiles([4, 4, 4, 4], 4) == four
    assert lib_percentiles(hist_of([4, 4, 4, 4])) == four


def test_percentiles_drop_and_merge():
    # 3001 cells: the histogram tracks values below 4096, so 5000 is dropped.
    values = [1] * 3000 + [5000]
    want = [(1.0, 0.0, 3000), (1.0, 1.0, 3000)]
    assert S.hdr_percentiles(values, len(values)) == want
    assert lib_percentiles(hist_of(values)) == want
    # 5001 cells: values below 8192 are tracked; 5000 shares a counter with
    # 5000..5003 and is reported as 5003.
    values = [1] * 5000 + [5000]
    for rows in (S.hdr_percentiles(values, len(values)), lib_percentiles(hist_of(values))):
        assert rows[-2:] == [(5003.0, 0.999837, 5001), (5003.0, 1.0, 5001)]
        assert all(r[0] == 1.0 for r in rows[:-2])


@pytest.mark.parametrize("seed", range(6))
def test_percentiles_library_matches_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    n_cells = int(rng.integers(1, 20000))
    scale = [5, 40, 3000, 20000][seed % 4]
    values = np.maximum(1, rng.geometric(1.0 / scale, size=n_cells))
    assert lib_percentiles(hist_of(values)) == S.hdr_percentiles(values, n_cells)


def test_oracle_grid_lattice_by_hand():
    # 4 x 4 lattice 0..3: extent 3, 16 points, 4 per cell -> side 1.5, 2 x 2
    # cells; x in {0, 1} -> cell 0, {2, 3} -> cell 1 (3 clamps to the last).
    g = np.stack(np.meshgrid(np.arange(4), np.arange(4)), -1).reshape(-1, 2).astype(np.float32)
    st = S.input_stats([g], 4)
    assert st["Grid"]["GridSize"] == [2, 2]
    assert st["Grid"]["TotalCells"] == 4 and st["Grid"]["NonEmptyCells"] == 4
    assert st["Grid"]["MaxPoints"] == 4 and st["Grid"]["GiniIndex"] == 0.0
    assert st["Density"] == pytest.approx(16 / 9, rel=1e-7)
    # Split across two sets: the same table.
    st2 = S.input_stats([g[:5], g[5:]], 4)
    assert st2["Grid"] == st["Grid"]


def test_oracle_nonempty_mbr_rule():
    p = np.array([[0, 2], [1, 2], [3, 2]], np.float32)
    lo, hi = S.nonempty_mbr(p)
    assert hi[1] == np.float32(2.00001) and lo[1] == 2
    assert hi[0] == 3


def lib_bounds(mbr_a, mbr_b, nd: int) -> tuple[float, float]:
    ma = np.ascontiguousarray(mbr_a, np.float32)
    mb = np.ascontiguousarray(mbr_b, np.float32)
    lo, hi = np.zeros(1, np.float32), np.zeros(1, np.float32)
    f = _lib._F
    rc = _lib.load().hd_mbr_hd_bounds(ma.ctypes.data_as(f), mb.ctypes.data_as(f), nd,
                                      lo.ctypes.data_as(f), hi.ctypes.data_as(f))
    assert rc == 0
    return lo[0], hi[0]


def test_hd_bounds_library_matches_oracle():
    from oracle import oracle
    # Boxes side by side: the farthest face of A (x = 0) is 3 away from B.
    assert lib_bounds([0, 0, 1, 1], [3, 0, 4, 1], 2)[0] == np.float32(3)
    # and its upper bound, sqrt(17) (worked by hand in test_oracle.py)
    assert lib_bounds([0, 0, 1, 1], [3, 0, 4, 1], 2)[1] == np.sqrt(np.float32(17))
    rng = np.random.default_rng(8)
    for nd in (2, 3):
        for _ in range(200):
            a = rng.normal(size=(rng.integers(1, 50), nd)).astype(np.float32)
            b = (rng.normal(size=(rng.integers(1, 50), nd)) * rng.uniform(0.1, 3)
                 + rng.uniform(-2, 2)).astype(np.float32)
            if rng.random() < 0.2:
                b[:, 0] = b[0, 0]  # a zero extent: widened by 1e-5
            raw_a = np.r_[a.min(0), a.max(0)]
            raw_b = np.r_[b.min(0), b.max(0)]
            lb2, ub2 = oracle.hd_bounds2(oracle.mbr(a), oracle.mbr(b), nd)
            lo, hi = lib_bounds(raw_a, raw_b, nd)
            assert lo == np.sqrt(np.float32(lb2)) and hi == np.sqrt(np.float32(ub2))
            assert lo <= hi


def test_feature_vector_by_hand():
    from hausdorffdistance_amd.features import N_FEATURES, feature_vector

    rows = [{"value": v, "percentile": p, "count": c} for v, p, c in SEVEN]

    def file(density, n, mbr, gini, size, nonempty, total):
        return {"Density": density, "NumPoints": n,
                "MBR": [{"Lower": lo, "Upper": hi} for lo, hi in mbr],
                "Grid": {"GiniIndex": gini, "Histogram": rows, "GridSize": size,
                         "NonEmptyCells": nonempty, "TotalCells": total}}

    stats = {"Input": {"NumDims": 2,
                       "Files": [file(2.5, 100, [(0, 2), (1, 5)], 0.25, [3, 4], 7, 12),
                                 file(9.0, 9, [(0, 1), (0, 1)], 0.0, [1, 1], 1, 1)],
                       "Grid": {"Histogram": rows, "GridSize": [5, 6]}},
             "Running": {"Repeats": [{"HDLowerBound": 1.0, "HDUpperBound": 4.0}]}}
    # Walking the rows from the end: 7/0.875 -> P99, 6/0.833 -> P95,
    # 5/0.667 -> P75, 3/0.333 -> P50, 2/0.167 -> P25, 1/0 -> P10.
    histo = [7, 6, 5, 3, 2, 1, 7, 6, 5, 3, 2, 1]
    want = ([2.5, 100, 2, 4, 0, 0.25] + histo + [3, 4, 0, 7, 12] +
            [9.0, 9, 1, 1, 0, 0.0] + histo + [1, 1, 0, 1, 1] +
            histo + [5, 6, 0])
    half = np.sqrt(4 + 16 + 1 + 1) / 2
    want += [1.0 / half, 4.0 / half]
    got = feature_vector(stats)
    assert len(got) == N_FEATURES == len(want)
    np.testing.assert_allclose(got, want, rtol=1e-15)
    # A percentile no row reaches stays -1.
    stats["Input"]["Grid"]["Histogram"] = [{"value": 4, "percentile": 0.0, "count": 4},
                                           {"value": 4, "percentile": 1.0, "count": 4}]
    got = feature_vector(stats)
    assert got[46:52] == [4.0, -1, -1, -1, -1, -1]


def sweep_run(path_a, path_b, translate, avg, npc, max_hit):
    rows = [{"value": v, "percentile": p, "count": c} for v, p, c in SEVEN]
    f = {"Density": 1.0, "NumPoints": 7, "MBR": [{"Lower": 0, "Upper": 1}] * 2,
         "Grid": {"GiniIndex": 0.1, "Histogram": rows, "GridSize": [2, 2],
                  "NonEmptyCells": 4, "TotalCells": 4}}
    return {"Input": {"NumDims": 2, "Translate": translate,
                      "Files": [dict(f, Path=path_a), dict(f, Path=path_b)],
                      "Grid": {"Histogram": rows, "GridSize": [2, 2]}},
            "Running": {"AvgTime": avg, "SampleRate": 0.01, "EBOnlyThreshold": 0,
                        "MaxHit": max_hit, "NumPointsPerCell": npc,
                        "Repeats": [{"HDLowerBound": 0.5, "HDUpperBound": 1.0}]}}


def test_tuning_export_picks_the_fastest_per_pair(tmp_path):
    from hausdorffdistance_amd import tuning

    d = tmp_path / "train" / "n_points_cell"
    d.mkdir(parents=True)
    runs = [sweep_run("a", "b", 0, 5.0, 4, 256), sweep_run("a", "b", 0, 3.0, 8, 256),
            sweep_run("a", "b", 0, 4.0, 16, 256), sweep_run("a", "b", 0.1, 2.0, 16, 64),
            sweep_run("a", "c", 0, 7.0, 2, 256)]
    for i, r in enumerate(runs):
        (d / f"x_{i}.json").write_text(json.dumps(r))
    (d / "empty.json").write_text("")
    (tmp_path / "train" / "other").mkdir()
    (tmp_path / "train" / "other" / "y.json").write_text(json.dumps(runs[0]))
    loaded = tuning.load_runs(str(tmp_path / "train"), "n_points_cell")
    assert len(loaded) == 5
    rows = {r["RunID"]: r for r in tuning.training_rows(loaded)}
    assert set(rows) == {"a-b-0", "a-b-0.1", "a-c-0"}
    assert rows["a-b-0"]["NumPointsPerCell"] == 8 and rows["a-b-0"]["AvgTime"] == 3.0
    assert rows["a-b-0.1"]["MaxHit"] == 64
    out = tmp_path / "t.csv"
    tuning.export_csv(list(rows.values()), str(out))
    lines = out.read_text().splitlines()
    assert len(lines) == 4 and lines[0].split(",")[:6] == [
        "RunID", "AvgTime", "SampleRate", "EBOnlyThreshold", "MaxHit", "NumPointsPerCell"]
    assert len(lines[0].split(",")) == 6 + 63
