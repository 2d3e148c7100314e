"""Builds libhd.so (HIP for gfx950) and the hd_exec CLI in-tree.

Every .hip / .cpp under csrc/ is compiled with hipcc into _build/ and linked
into hausdorffdistance_amd/lib/libhd.so. Device code is compiled with FP
contraction off (the reference builds its kernels with --fmad=false "for
consistency to the CPU results", CMakeLists.txt:122) and with packed FP32
instructions disabled.

Usage: python -m hausdorffdistance_amd.build [-j N] [--force]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build" / "hd"
LIBDIR = PKG / "lib"
LIB = LIBDIR / "libhd.so"
EXE = LIBDIR / "hd_exec"

ARCH = os.environ.get("HD_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-fno-gpu-flush-denormals-to-zero",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


# Per-file device flags. eb.hip: matrix-core results go to VGPRs (the
# filter consumes every result with vector integer minima; the AGPR form
# would add one copy per result).
FILE_DEVICE = {
    "eb.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def _flags_for(src: Path) -> list[str]:
    if src.suffix == ".hip":
        return COMMON + DEVICE + FILE_DEVICE.get(src.name, []) + ["-x", "hip"]
    return COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++"]


def _obj_for(src: Path) -> Path:
    return BUILD / (src.name + ".o")


def _stamp(src: Path, flags: list[str]) -> str:
    h = hashlib.sha256()
    h.update(" ".join(flags).encode())
    h.update(src.read_bytes())
    for hdr in sorted(list(CSRC.glob("*.h")) + list(INCLUDE.glob("*.h"))):
        h.update(hdr.read_bytes())
    return h.hexdigest()


def compile_one(src: Path, force: bool) -> tuple[Path, float]:
    import time

    flags = _flags_for(src)
    obj = _obj_for(src)
    stamp_file = obj.with_suffix(".stamp")
    stamp = _stamp(src, flags)
    if not force and obj.exists() and stamp_file.exists() and stamp_file.read_text() == stamp:
        return obj, 0.0
    t0 = time.time()
    cmd = [hipcc()] + flags + ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        # hipcc's host pass warns about the device-only target feature; keep
        # other diagnostics visible.
        lines = [
            ln
            for ln in res.stderr.splitlines()
            if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
        ]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    stamp_file.write_text(stamp)
    return obj, time.time() - t0


def build(jobs: int | None = None, force: bool = False, verbose: bool = True) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted(
        p for p in list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")) if p.name != "hd_exec.cpp"
    )
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(compile_one, s, force): s for s in srcs}
        for fut in cf.as_completed(futs):
            obj, dt = fut.result()
            objs.append(obj)
            if verbose and dt > 0:
                print(f"[build] {futs[fut].name}: {dt:.1f}s", flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [
        str(o) for o in sorted(objs)
    ] + ["-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    exe_src = CSRC / "hd_exec.cpp"
    if exe_src.exists():
        cmd = [hipcc()] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                                   str(exe_src), "-o", str(EXE), f"-L{LIBDIR}", "-lhd",
                                   "-Wl,-rpath,$ORIGIN", "-lpthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hd_exec build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"[build] linked {LIB}", flush=True)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    args = ap.parse_args()
    build(args.jobs, args.force)


if __name__ == "__main__":
    main()
