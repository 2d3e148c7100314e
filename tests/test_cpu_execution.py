"""execution = CPU: the library's host early break (the reference's CPU
overload, early_break.h:44-115, and its -check answer,
run_hausdorff_distance.cu:366-388) against the oracle. Host code only: these
run without a GPU."""
from __future__ import annotations

import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd.build import EXE
from oracle import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def triple(r):
    return {"cmax2_bits": r.cmax2_bits, "argmax": r.argmax, "nn": r.nn}


def otriple(o):
    return {"cmax2_bits": o.cmax2_bits, "argmax": o.argmax, "nn": o.nn}


def test_uniform_100k_both_ways_equals_full_size_golden():
    full = json.loads((GOLDEN / "full_size.json").read_text())["uniform_100k_2d"]
    a, b = hd.gen.uniform_pair(100_000, 2)
    assert triple(hd.directed_host(a, b)) == full["ab"]
    assert triple(hd.directed_host(b, a)) == full["ba"]


def test_golden_fixtures_on_host():
    cases = np.load(GOLDEN / "cases.npz")
    expected = json.loads((GOLDEN / "expected.json").read_text())
    for name, exp in expected.items():
        a, b = cases[f"{name}__a"], cases[f"{name}__b"]
        assert triple(hd.directed_host(a, b)) == exp["ab"], name
        assert triple(hd.directed_host(b, a)) == exp["ba"], name


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("shuffle", [True, False])
def test_thread_count_and_order_do_not_change_bits(threads, shuffle):
    a, b = hd.gen.gmm_pair(20_000)
    o = oracle.eb_directed(a, b)
    r = hd.directed_host(a, b, parallelism=threads, shuffle=shuffle, seed=5)
    assert triple(r) == otriple(o)
    # every evaluation counted is a real pair: at most |A| |B|
    assert 0 < r.evals <= a.shape[0] * b.shape[0]


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 777), (777, 1), (31, 33), (64, 64), (1000, 3000)])
def test_small_and_ragged_sizes_3d(na, nb):
    a = hd.gen.uniform(31, na, 3)
    b = hd.gen.uniform(32, nb, 3)
    o = oracle.brute_directed(a, b)
    assert triple(hd.directed_host(a, b)) == otriple(o)


def test_single_thread_without_shuffle_never_breaks_on_first_point():
    # the first A point of the input order scans all of B: its evaluations
    # alone are |B| (padding blocks are not counted)
    a = np.array([[0.5, 0.5]], np.float32)
    b = hd.gen.uniform(7, 1000, 2)
    r = hd.directed_host(a, b, parallelism=1, shuffle=False)
    assert r.evals == 1000


def test_run_config_cpu_execution_stats_and_check():
    a, b = hd.gen.gmm_pair(30_000)
    cfg = hd.RunConfig(variant=hd.Variant.EB, execution=hd.Execution.CPU, n_dims=2, repeat=3,
                       symmetric=True, check=True, parallelism=4)
    res = hd.run_hausdorff_distance(cfg, a, b)
    o_ab, o_ba = oracle.eb_symmetric(a, b)
    assert triple(res.ab) == otriple(o_ab)
    assert triple(res.ba) == otriple(o_ba)
    assert res.check_pass == 1
    s = res.stats
    assert s["Check"]["Pass"] is True and s["Check"]["Diff"] == 0
    reps = s["Running"]["Repeats"]
    assert len(reps) == 3
    for rep in reps:
        assert rep["Algorithm"] == "Early Break" and rep["Execution"] == "CPU"
        assert rep["Threads"] == 4 and rep["ReportedTime"] > 0
    assert s["Input"]["NumPoints"] == 60_000
    want = max(o_ab.cmax2_bits, o_ba.cmax2_bits)
    assert res.cmax2_bits == want


def test_cpu_execution_errors():
    a, b = hd.gen.uniform_pair(100, 2)
    for v in (hd.Variant.GRID, hd.Variant.LBVH, hd.Variant.HYBRID, hd.Variant.NN):
        cfg = hd.RunConfig(variant=v, execution=hd.Execution.CPU, n_dims=2, repeat=1)
        with pytest.raises(hd.HDError) as e:
            hd.run_hausdorff_distance(cfg, a, b)
        assert e.value.code == hd._lib.HD_EUNSUPPORTED
        assert "not implemented" in str(e.value)
    cfg = hd.RunConfig(execution=hd.Execution.CPU, n_dims=2, repeat=1, learned_order=True)
    with pytest.raises(hd.HDError) as e:
        hd.run_hausdorff_distance(cfg, a, b)
    assert e.value.code == hd._lib.HD_EUNSUPPORTED
    bad = a.copy()
    bad[17, 1] = np.nan
    with pytest.raises(hd.HDError) as e:
        hd.directed_host(bad, b)
    assert e.value.code == hd._lib.HD_EINVAL
    bad[17, 1] = np.inf
    with pytest.raises(hd.HDError) as e:
        hd.directed_host(a, bad)
    assert e.value.code == hd._lib.HD_EINVAL
    with pytest.raises(hd.HDError):
        hd.directed_host(a[:0], b)


def test_calculate_distance_host_overload():
    a, b = hd.gen.uniform_pair(5_000, 2)
    eb = hd.HausdorffDistanceEarlyBreak(n_threads=2)
    h = eb.calculate_distance(a, b)
    o = oracle.eb_directed(a, b)
    assert np.float32(h) == np.float32(o.hd)
    st = eb.get_stats()
    assert st["Execution"] == "CPU" and st["Algorithm"] == "Early Break"
    with pytest.raises(hd.HDError):
        hd.HausdorffDistanceHybrid().calculate_distance(a, b)


def test_cli_cpu_execution_checks_by_default():
    out = subprocess.run([str(EXE), "-input1", "uniform_100k_2d", "-input_type", "synthetic",
                          "-n_dims", "2", "-execution", "cpu", "-repeat", "2", "-symmetric"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    assert "HausdorffDistance is checked" in out.stdout
    full = json.loads((GOLDEN / "full_size.json").read_text())["uniform_100k_2d"]
    m = re.search(r"cmax2 bits 0x([0-9a-f]+), argmax (\d+), nn (\d+), direction (\d)", out.stdout)
    win = max((full["ab"], 0), (full["ba"], 1), key=lambda t: t[0]["cmax2_bits"])
    assert (int(m.group(1), 16), int(m.group(2)), int(m.group(3)), int(m.group(4))) == (
        win[0]["cmax2_bits"], win[0]["argmax"], win[0]["nn"], win[1])


def test_b_order_flags_exclude_each_other():
    """HD_FLAG_B_ORDER_RANDOM with HD_FLAG_B_ORDER_STRATIFIED is refused at
    the boundary (checked before the execution is chosen, so on the CPU
    too); one of them alone runs, and the Python interface refuses an
    unknown order name."""
    import ctypes as C

    from hausdorffdistance_amd import _lib

    a, b = hd.gen.uniform_pair(200, 2)
    lib = _lib.load()
    pa = a.ctypes.data_as(C.POINTER(C.c_float))
    pb = b.ctypes.data_as(C.POINTER(C.c_float))
    for extra, ok in ((_lib.HD_FLAG_B_ORDER_RANDOM, True),
                      (_lib.HD_FLAG_B_ORDER_STRATIFIED, True),
                      (_lib.HD_FLAG_B_ORDER_RANDOM | _lib.HD_FLAG_B_ORDER_STRATIFIED, False)):
        c = hd.RunConfig(execution=hd.Execution.CPU, n_dims=2, repeat=1).to_c()
        c.flags |= extra
        out = _lib.ResultC()
        buf = C.create_string_buffer(1 << 16)
        rc = lib.hd_run_hausdorff_distance(None, C.byref(c), pa, len(a), pb, len(b),
                                           C.byref(out), buf, 1 << 16)
        assert (rc == 0) == ok, (extra, rc, lib.hd_last_error())
        if not ok:
            assert rc == _lib.HD_EINVAL and b"exclude" in lib.hd_last_error()
    with pytest.raises(hd.HDError):
        hd.RunConfig(b_order="sorted").to_c()
