"""hd_exec end to end on files: WKT and PLY inputs (with -normalize and
-translate, -serialize) give the oracle's bits on the same points."""
from __future__ import annotations

import re
import subprocess

import numpy as np
import pytest

import hausdorffdistance_amd as hd
from hausdorffdistance_amd.build import EXE
from oracle import oracle

pytestmark = pytest.mark.gpu


def run_cli(*args: str) -> tuple[int, int, int]:
    out = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr + out.stdout
    m = re.search(r"cmax2 bits 0x([0-9a-f]+), argmax (\d+), nn (\d+)", out.stdout)
    assert m, out.stdout
    return int(m.group(1), 16), int(m.group(2)), int(m.group(3))


def wkt_file(path, pts: np.ndarray) -> str:
    lines = [f"LINESTRING ({', '.join(f'{float(x)!r} {float(y)!r}' for x, y in pts[i:i + 50])})"
             for i in range(0, len(pts), 50)]
    path.write_text("\n".join(lines) + "\n")
    return str(path)


def test_cli_wkt_with_cache(tmp_path):
    a = hd.gen.uniform(21, 3000, 2).astype(np.float64)
    b = hd.gen.uniform(22, 2500, 2).astype(np.float64)
    fa, fb = wkt_file(tmp_path / "a.wkt", a), wkt_file(tmp_path / "b.wkt", b)
    pa = hd.load_points(fa, "wkt", 2)
    pb = hd.load_points(fb, "wkt", 2)
    o = oracle.eb_directed(pa, pb)
    # rt-hdist is exact on these uniform inputs (its scan radius covers
    # every nearest point), like the other variants by construction.
    for variant in ("eb", "hybrid", "rt", "grid", "nn", "rt-hdist"):
        got = run_cli("-input1", fa, "-input2", fb, "-input_type", "wkt", "-n_dims", "2",
                      "-variant", variant, "-repeat", "1", "-check",
                      "-serialize", str(tmp_path / "cache"))
        assert got == (o.cmax2_bits, o.argmax, o.nn), variant
    assert len(list((tmp_path / "cache").glob("points_*_limit_*.bin"))) == 2


def test_cli_ply_normalize_translate(tmp_path):
    rng = np.random.default_rng(3)
    for name, n in (("a", 4000), ("b", 3000)):
        v = rng.normal(size=(n, 3)) * np.array([2.0, 1.0, 0.5]) + 5.0
        head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {n}\n"
                "property float x\nproperty float y\nproperty float z\nend_header\n")
        (tmp_path / f"{name}.ply").write_bytes(head.encode() + v.astype("<f4").tobytes())
    fa, fb = str(tmp_path / "a.ply"), str(tmp_path / "b.ply")
    pa = hd.normalize_points(hd.load_points(fa, "ply", 3))
    pb = hd.translate_points(hd.normalize_points(hd.load_points(fb, "ply", 3)), 0.1)
    o = oracle.eb_directed(pa, pb)
    got = run_cli("-input1", fa, "-input2", fb, "-input_type", "ply", "-n_dims", "3",
                  "-variant", "eb", "-repeat", "1", "-normalize", "-translate", "0.1")
    assert got == (o.cmax2_bits, o.argmax, o.nn)


def test_cli_b_order(tmp_path):
    """-b_order random|stratified|auto selects the early break's visit order
    of B; every order gives the oracle's bits, an unknown one is refused."""
    a = hd.gen.uniform(23, 5000, 2).astype(np.float64)
    b = hd.gen.uniform(24, 4000, 2).astype(np.float64)
    fa, fb = wkt_file(tmp_path / "a.wkt", a), wkt_file(tmp_path / "b.wkt", b)
    o = oracle.eb_directed(hd.load_points(fa, "wkt", 2), hd.load_points(fb, "wkt", 2))
    for order in ("random", "stratified", "auto"):
        got = run_cli("-input1", fa, "-input2", fb, "-input_type", "wkt", "-n_dims", "2",
                      "-variant", "eb", "-repeat", "1", "-nocheck", "-b_order", order)
        assert got == (o.cmax2_bits, o.argmax, o.nn), order
    bad = subprocess.run([str(EXE), "-input1", fa, "-input2", fb, "-input_type", "wkt",
                          "-n_dims", "2", "-b_order", "sorted"],
                         capture_output=True, text=True, timeout=120)
    assert bad.returncode != 0 and "b_order" in bad.stderr
