"""CPU-only checks of the C ABI: the library loads, exports every symbol the
headers declare, fills the reference defaults, and rejects bad requests with
error codes (never an abort) before touching a device."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd import _lib

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions() -> set[str]:
    names = set()
    for h in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(hd_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    unbound = sorted(names - set(_lib.SIGNATURES))
    assert not unbound, unbound


def test_config_defaults_follow_reference_flags():
    c = _lib.RunConfigC()
    _lib.load().hd_run_config_init(C.byref(c))
    assert c.abi_version == _lib.HD_ABI_VERSION
    assert c.variant == _lib.HD_VARIANT_EB and c.execution == _lib.HD_EXEC_GPU
    assert c.repeat == 5 and c.seed == 0 and c.parallelism == -1
    assert abs(c.sample_rate - 0.01) < 1e-9
    assert (c.n_points_cell, c.max_hit, c.eb_only_threshold) == (8, 256, 0)


def _run(cfg: _lib.RunConfigC, a, b):
    lib = _lib.load()
    out = _lib.ResultC()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    rc = lib.hd_run_hausdorff_distance(None, C.byref(cfg), C.c_void_p(a.ctypes.data),
                                       a.shape[0], C.c_void_p(b.ctypes.data), b.shape[0],
                                       C.byref(out), None, 0)
    return rc, lib.hd_last_error().decode()


def _cfg(**kw) -> _lib.RunConfigC:
    c = _lib.RunConfigC()
    _lib.load().hd_run_config_init(C.byref(c))
    c.n_dims = 2
    for k, v in kw.items():
        setattr(c, k, v)
    return c


PTS = np.zeros((4, 2), np.float32)


@pytest.mark.parametrize("kw,code,needle", [
    ({"n_dims": 4}, _lib.HD_EINVAL, "dimensions"),
    ({"n_dims": 1}, _lib.HD_EINVAL, "dimensions"),
    ({"variant": 9}, _lib.HD_EINVAL, "variant"),
    ({"execution": 5}, _lib.HD_EINVAL, "execution"),
    # the CPU path serves the early break only (hausdorff_distance.h:21-26)
    ({"execution": _lib.HD_EXEC_CPU, "variant": _lib.HD_VARIANT_HYBRID},
     _lib.HD_EUNSUPPORTED, "CPU"),
    ({"repeat": 0}, _lib.HD_EINVAL, "repeat"),
    ({"abi_version": 99}, _lib.HD_EINVAL, "abi_version"),
])
def test_bad_config_is_an_error_code(kw, code, needle):
    rc, msg = _run(_cfg(**kw), PTS, PTS)
    assert rc == code and needle in msg


def test_empty_set_is_an_error_code():
    rc, msg = _run(_cfg(), np.zeros((0, 2), np.float32), PTS)
    assert rc == _lib.HD_EINVAL and "empty" in msg
    rc, msg = _run(_cfg(), PTS, np.zeros((0, 2), np.float32))
    assert rc == _lib.HD_EINVAL and "empty" in msg


def test_python_mirror_rejects_bad_requests():
    with pytest.raises(hd.HDError):
        hd.Variant.from_string("bnb")
    with pytest.raises(hd.HDError):
        hd.Execution.from_string("tpu")
    assert hd.Variant.from_string("rt") == hd.Variant.LBVH
    with pytest.raises(hd.HDError):
        hd.run_hausdorff_distance(hd.RunConfig(n_dims=5), PTS, PTS)


def test_directed_device_validates_before_device():
    lib = _lib.load()
    out = _lib.DirectedResultC()
    rc = lib.hd_directed_device(None, 0, 7, None, 1, None, 1, 0, 0, None, C.byref(out))
    assert rc == _lib.HD_EINVAL
    rc = lib.hd_directed_device(None, 0, 2, None, 0, None, 1, 0, 0, None, C.byref(out))
    assert rc == _lib.HD_EINVAL
    rc = lib.hd_nn_device(None, 7, 2, None, 1, None, 1, None, None)
    assert rc == _lib.HD_EINVAL


def test_generators_are_deterministic_and_follow_their_recipes():
    a1 = hd.gen.uniform(1, 1000, 2)
    a2 = hd.gen.uniform(1, 1000, 2)
    np.testing.assert_array_equal(a1, a2)
    assert a1.min() >= 0 and a1.max() < 1
    # 24-bit uniforms are exact multiples of 2^-24.
    assert np.all((a1 * 2.0 ** 24) == np.floor(a1 * 2.0 ** 24))
    a, b = hd.gen.gmm_pair(20_000)
    moved = np.any(b != a + np.array([0.002, 0.001], np.float32), axis=1)
    assert moved.sum() == 200  # exactly 1% outliers
    s = hd.gen.sphere(7, 5000, 0.005, 0.0)
    r = np.linalg.norm(s.astype(np.float64), axis=1)
    assert abs(r.mean() - 1) < 1e-3 and 0.003 < r.std() < 0.007
    bt = hd.gen.batch(9, 16)
    assert bt.sizes_a.min() >= 32 and bt.sizes_a.max() < 65536
    assert bt.a.shape[0] == int(bt.offs_a[-1])
    # Parallel generation equals a sequential SplitMix64 stream.
    lib = _lib.load()
    x = lib.hd_splitmix64_at(1, 0)
    assert np.float32((x >> 40) * 2.0 ** -24) == a1[0, 0]
    sub = hd.gen.subsample(14, 1000, 50)
    assert len(set(sub.tolist())) == 50 and sub.max() < 1000


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors have the size and field offsets the C compiler
    gives the header's structs."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"hd_run_config": _lib.RunConfigC, "hd_directed_result": _lib.DirectedResultC,
               "hd_result": _lib.ResultC, "hd_mlp_train_config": _lib.MlpTrainConfigC}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "hd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", f"-I{INCLUDE}", str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = []
    for cname, py in structs.items():
        want.append(f"{cname} size {C.sizeof(py)}")
        for fname, _ in py._fields_:
            want.append(f"{cname} {fname} {getattr(py, fname).offset}")
    assert [g for g in got if g] == want
