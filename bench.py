"""Benchmark of the directed Hausdorff-distance hot path on one MI355X.

Workload (default): 10M x 10M 2D FP32 Gaussian mixture (64 components,
sigma 0.01), B = A + (0.002, 0.001) with 1% uniform outliers. One step is one
directed h(A -> B) over device-resident inputs through the C ABI
(hd_directed_device), early-break variant; the plain (pseudo-random) visit
order and the learned order are both timed, `value` is the learned order's
A-points/s when the learned path is available, else the plain one's.

Prints ONE JSON line. The roofline object is for the dominant kernel (the
early-break kernel with the most device time: the matrix-core filter of the
budgeted rounds, or the exact scan), its duration measured with HIP events on
the library's stream; the CPU baseline is the CPU oracle (a C restatement of
the reference's CPU early break) on a bounded sample of the same workload.
The other variants (grid, LBVH, hybrid, nearest-neighbour search) and, under
config.next_rows, the input-statistics pipeline, the RT-HDIST baseline (on
1M x 1M uniform 2D) and the loaders' parse rates are timed beside it; none
of them enters `value`.

Secondary workloads (--workload): uniform_100k_2d, sphere_20m_15m_3d,
batch_4096_2d, and gmm_50m_256_2d, which also times the learned order
against the plain one.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# VALU lane-op ceiling: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz.
VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9
LANE_OPS_PER_EVAL = {2: 6, 3: 9}
# Dense FP32 matrix-core peak (MI355X).
MFMA_F32_TFLOPS = 157.3
# The early-break kernels and the pipeline stages that launch them: the
# exact scan (seed, unfiltered rounds, top completions, tail) and the
# matrix-core filter of the budgeted rounds (16x16x4 FP32 MFMA: 4 MACs =
# 8 FLOP per A/B pair).
EB_KERNELS = {
    "EbScanKernel": {"full": "EbScanKernel<2,8,4>",
                     "stages": ("eb_seed", "eb_scan_r1", "eb_scan_r2", "eb_scan_rest",
                                "eb_scan_r3", "eb_scan_r4", "eb_top", "eb_top_exact",
                                "eb_tail", "eb_tail_exact")},
    "EbMfmaKernel": {"full": "EbMfmaKernel<2, false>",
                     "stages": ("eb_mfma_r1", "eb_mfma_r2", "eb_mfma_rest", "eb_mfma_top",
                                "eb_mfma_tail")},
    # The 1024-step round, split over four 256-step segments per point.
    "EbMfmaKernelGrouped": {"full": "EbMfmaKernel<2, true>",
                            "stages": ("eb_mfma_r3", "eb_mfma_r4")},
}
MFMA_FLOP_PER_PAIR = 8


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_times(lib, ctx, reset: bool) -> dict:
    buf = C.create_string_buffer(1 << 16)
    rc = lib.hd_context_stage_times(ctx.handle, buf, 1 << 16, 1 if reset else 0)
    if rc != 0:
        raise RuntimeError(lib.hd_last_error().decode())
    return json.loads(buf.value.decode())


def time_steps(hd, ctx, variant, da, db, steps, warmup, learned, model=None,
               single_pass=False):
    """W untimed warm-up steps, then K timed steps bracketed by device
    synchronisation; returns (ms per step over the K, results, stage times)."""
    import torch

    lib = hd.load()
    kw = dict(seed=0, learned_order=learned, model=model, single_pass=single_pass)
    for _ in range(warmup):
        hd.directed_device(ctx, variant, da, db, **kw)
    torch.cuda.synchronize()
    lib.hd_context_set_profiling(ctx.handle, 1)
    stage_times(lib, ctx, reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for _ in range(steps):
        results.append(hd.directed_device(ctx, variant, da, db, **kw))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = stage_times(lib, ctx, reset=True)
    lib.hd_context_set_profiling(ctx.handle, 0)
    return (t1 - t0) * 1e3 / steps, results, stages


def spread(results) -> dict:
    """Median, minimum and maximum of the per-step device times (HIP events
    around each call)."""
    t = sorted(r.time_ms for r in results)
    return {"median_ms": t[len(t) // 2], "min_ms": t[0], "max_ms": t[-1], "n": len(t)}


def next_rows(hd, ctx, da, db, n_a: int, n_b: int) -> dict:
    """The widened rows around the hot path, each timed once on the same
    workload: the input statistics of the stats JSON (three stats grids:
    A, B and both, each one counting pass over its points) and the
    loaders' parse rates (host code, on files written here)."""
    import tempfile

    import torch

    lib = hd.load()
    out = {}
    cfg = hd.RunConfig(variant=hd.Variant.HYBRID, n_dims=2, repeat=1, stats_n_points_cell=10)
    hd.run_hausdorff_distance(cfg, da, db, ctx=ctx)  # warm-up
    torch.cuda.synchronize()
    lib.hd_context_set_profiling(ctx.handle, 1)
    stage_times(lib, ctx, reset=True)
    res = hd.run_hausdorff_distance(cfg, da, db, ctx=ctx)
    st = stage_times(lib, ctx, reset=True)
    lib.hd_context_set_profiling(ctx.handle, 0)
    s = st.get("input_stats")
    if s:
        pts = 2 * (n_a + n_b)
        grid = res.stats["Input"]["Grid"]
        out["input_stats"] = {
            "ms": s["ms"], "points_counted": pts,
            "points_per_s": pts / (s["ms"] / 1e3),
            # point reads (8 B) and the count tables' clear, count and
            # histogram passes (4 B per cell each)
            "gb_per_s": (pts * 8 + 3 * 4 * 2 * grid["TotalCells"]) / (s["ms"] / 1e3) / 1e9,
            "union_grid": {k: grid[k] for k in ("GridSize", "NonEmptyCells", "MaxPoints")}}
    # RT-HDIST on 1M x 1M uniform 2D (on the clustered 10M pair its scan of
    # every cell within three diagonals is millions of evaluations per point).
    ua = torch.from_numpy(hd.gen.uniform(1, 1_000_000, 2)).cuda()
    ub = torch.from_numpy(hd.gen.uniform(2, 1_000_000, 2)).cuda()
    ms_r, res_r, st_r = time_steps(hd, ctx, hd.Variant.RT_HDIST, ua, ub, 3, 1, learned=False)
    ms_e, res_e, _ = time_steps(hd, ctx, hd.Variant.HYBRID, ua, ub, 3, 1, learned=False)
    rr, re_ = res_r[-1], res_e[-1]
    out["rt_hdist_uniform_1m_2d"] = {
        "ms_per_step": ms_r, "a_points_per_s": 1e6 / (ms_r / 1e3), "evals": rr.evals,
        "exact": (rr.cmax2_bits, rr.argmax) == (re_.cmax2_bits, re_.argmax),
        "hybrid_ms_per_step": ms_e,
        "stages_ms": {k: x["ms"] / max(1, x["calls"]) for k, x in st_r.items()}}
    rng = np.random.default_rng(1)
    with tempfile.TemporaryDirectory() as tmp:
        pts = rng.random((1_000_000, 2))
        wkt = os.path.join(tmp, "a.wkt")
        with open(wkt, "w") as f:
            for i in range(0, len(pts), 100):
                f.write("LINESTRING (" + ", ".join(f"{x:.9f} {y:.9f}" for x, y in pts[i:i + 100])
                        + ")\n")
        ply = os.path.join(tmp, "a.ply")
        v = rng.random((2_000_000, 3)).astype("<f4")
        with open(ply, "wb") as f:
            f.write((f"ply\nformat binary_little_endian 1.0\nelement vertex {len(v)}\n"
                     "property float x\nproperty float y\nproperty float z\nend_header\n").encode())
            f.write(v.tobytes())
        for name, path, kind, nd, n in (("wkt", wkt, "wkt", 2, len(pts)),
                                        ("ply_binary", ply, "ply", 3, len(v))):
            t0 = time.perf_counter()
            p = hd.load_points(path, kind, nd)
            dt = time.perf_counter() - t0
            assert p.shape == (n, nd)
            out[f"load_{name}"] = {"points": n, "s": dt, "points_per_s": n / dt,
                                   "mb_per_s": os.path.getsize(path) / dt / 1e6}
    return out


def eb_roofline(stages: dict, steps: int) -> dict:
    """Roofline of the dominant early-break kernel over the timed steps: the
    kernel whose stages take the most device time (HIP events on the
    library's stream around each stage), its average dispatch duration
    (stage time over the kernel dispatches the stages made), and its
    algorithmic work per dispatch (the kernels' own evaluation counters).
    `traffic` is null here: HBM bytes per dispatch come from rocprofv3
    counter runs of the same command (FETCH_SIZE / WRITE_SIZE), kept with
    the command in profiles/."""
    per = {}
    for name, k in EB_KERNELS.items():
        ms = sum(stages.get(st, {}).get("ms", 0.0) for st in k["stages"])
        disp = sum(stages.get(st, {}).get("dispatches", 0) for st in k["stages"])
        evals = sum(stages.get(st, {}).get("evals", 0) for st in k["stages"])
        points = sum(stages.get(st, {}).get("points", 0) for st in k["stages"])
        per[name] = (ms, disp, evals, points)
    name = max(per, key=lambda n: per[n][0])
    ms, disp, evals, points = per[name]
    launches = max(1, disp)
    t_kernel = ms / launches / 1e3
    other = {n: {"ms_per_step": v[0] / max(1, steps), "dispatches_per_step": v[1] / max(1, steps),
                 "evals_per_step": v[2] / max(1, steps),
                 "ms_per_dispatch": v[0] / max(1, v[1])} for n, v in per.items()}
    if name.startswith("EbMfmaKernel"):
        flop = MFMA_FLOP_PER_PAIR * evals / launches
        achieved = flop / t_kernel / 1e12 if t_kernel > 0 else None
        return {
            "bound": "mfma", "kernel": EB_KERNELS[name]["full"], "achieved": achieved,
            "peak": MFMA_F32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / MFMA_F32_TFLOPS if achieved else None,
            "traffic": None, "algorithmic_flop": flop, "kernel_ms": t_kernel * 1e3,
            "dispatches_per_step": disp / max(1, steps),
            "kernels": other,
        }
    bytes_alg = 8.0 * (points + evals) / launches
    achieved = bytes_alg / t_kernel / 1e9 if t_kernel > 0 else None
    return {
        "bound": "hbm", "kernel": EB_KERNELS[name]["full"], "achieved": achieved,
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS if achieved else None,
        "traffic": None, "algorithmic_bytes": bytes_alg, "kernel_ms": t_kernel * 1e3,
        "dispatches_per_step": disp / max(1, steps),
        "valu_frac": (evals / launches * LANE_OPS_PER_EVAL[2] / t_kernel / VALU_LANE_OPS)
        if t_kernel > 0 else None,
        "kernels": other,
    }


# Index queries: algorithmic bytes per dispatch = P per query point + the
# index records read (8 B per grid cell lookup: start and end; 64 B per LBVH
# node record: both child boxes and ids) + P per candidate B point.
INDEX_KERNELS = {
    "grid": {"kernel": "GridQueryKernel<2>", "stages": ("grid_seed", "grid_query"),
             "record_bytes": 8},
    "lbvh": {"kernel": "LbvhQueryKernel<2>", "stages": ("lbvh_seed", "lbvh_query"),
             "record_bytes": 64},
}


def index_roofline(variant: str, stages: dict, n_dims: int = 2) -> dict | None:
    k = INDEX_KERNELS.get(variant)
    if k is None:
        return None
    p = 4 * n_dims
    ms = sum(stages.get(st, {}).get("ms", 0.0) for st in k["stages"])
    disp = sum(stages.get(st, {}).get("dispatches", 0) for st in k["stages"])
    evals = sum(stages.get(st, {}).get("evals", 0) for st in k["stages"])
    points = sum(stages.get(st, {}).get("points", 0) for st in k["stages"])
    lookups = sum(stages.get(st, {}).get("lookups", 0) for st in k["stages"])
    if disp == 0 or ms <= 0:
        return None
    bytes_alg = (p * points + k["record_bytes"] * lookups + p * evals) / disp
    t = ms / disp / 1e3
    gbs = bytes_alg / t / 1e9
    return {"bound": "hbm", "kernel": k["kernel"], "achieved": gbs, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": bytes_alg,
            "kernel_ms": t * 1e3, "dispatches": disp,
            "evals_per_point": evals / max(1, points),
            "records_per_point": lookups / max(1, points)}


# The host-CPU share of one GPU on the benchmark machines: worker pools are
# sized to it, whatever the machine reports as hardware concurrency.
CPU_SHARE = 16


def cpu_info() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": model, "hardware_concurrency": os.cpu_count(), "affinity": aff,
            "threads_used": max(1, min(aff or 1, CPU_SHARE))}


def cpu_baseline(a: np.ndarray, b: np.ndarray, target_s: float = 12.0) -> dict:
    """CPU oracle early break on a random sample of A against all of B, on
    the host threads this process may use (at most the GPU's CPU share)."""
    from oracle import oracle

    info = cpu_info()
    cores = info["threads_used"]
    rng = np.random.default_rng(0)
    k = min(100_000, a.shape[0])
    idx = rng.choice(a.shape[0], k, replace=False)
    r = oracle.eb_directed(a[idx], b, n_threads=cores, seed=0)
    t = r.compute_ms / 1e3
    if t < target_s / 4 and k < a.shape[0]:
        k2 = int(min(a.shape[0], k * target_s / max(t, 1e-3)))
        idx = rng.choice(a.shape[0], k2, replace=False)
        r = oracle.eb_directed(a[idx], b, n_threads=cores, seed=0)
        k = k2
        t = r.compute_ms / 1e3
    return {
        "value": k / t,
        "unit": "A-points/s",
        "cores": int(r.n_threads),
        "kind": "port",
        "sample": f"{k} random A points (seed 0) vs all {b.shape[0]} B points, "
                  f"directed A->B, {r.compared_pairs} d2 evaluations in {t:.2f} s; "
                  f"host {info['model']}, {info['hardware_concurrency']} hardware threads, "
                  f"{info['affinity']} in this process's affinity, {cores} used "
                  f"(the GPU's host-CPU share is {CPU_SHARE})",
    }


def host_early_break(hd, a: np.ndarray, b: np.ndarray, want: tuple) -> dict:
    """The library's own CPU execution path (hd_directed_host: the early
    break on the host, AVX2) on the whole workload, same bits required."""
    info = cpu_info()
    r = hd.directed_host(a, b, parallelism=info["threads_used"])
    if (r.cmax2_bits, r.argmax, r.nn) != want:
        raise RuntimeError("host early break differs from the GPU result")
    return {"a_points_per_s": a.shape[0] / (r.time_ms / 1e3), "ms": r.time_ms,
            "threads": info["threads_used"], "evals": r.evals,
            "note": "execution = CPU of the library (not the oracle), whole workload"}


def copy_probe(mib: int) -> None:
    """A plain device-to-device copy of known size, to calibrate the HBM
    byte counters (FETCH_SIZE / WRITE_SIZE) against."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "repeats": 3}), flush=True)


def secondary(args) -> None:
    """The other workloads: each variant timed, same bits required; one JSON
    line with the fastest variant's throughput."""
    import torch

    import hausdorffdistance_amd as hd

    ctx = hd.Context(0)
    if args.workload == "batch_4096_2d":
        log("generating 4096 pairs")
        bt = hd.gen.batch(9, 4096)
        da = torch.from_numpy(bt.a).cuda()
        db = torch.from_numpy(bt.b).cuda()
        doa = torch.from_numpy(bt.offs_a.astype(np.int64)).cuda()
        dob = torch.from_numpy(bt.offs_b.astype(np.int64)).cuda()
        for _ in range(max(1, args.warmup)):
            hd.batch_directed(ctx, da, doa, db, dob)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = hd.batch_directed(ctx, da, doa, db, dob)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        n_a = int(bt.offs_a[-1])
        line = {"metric": "batched directed HD, A-points/s over 4096 pairs", "value": n_a / (ms / 1e3),
                "unit": "A-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic",
                "config": {"workload": "batch_4096_2d: 4096 pairs, sizes log-uniform in [32, 65536), "
                                       "4-component blobs sigma 0.02, one persistent launch in a HIP graph",
                           "pairs_per_s": 4096 / (ms / 1e3), "total_a": n_a,
                           "total_b": int(bt.offs_b[-1]), "evals": res[0].evals,
                           "graph_ms": res[0].time_ms}}
        print(json.dumps(line), flush=True)
        ctx.close()
        return
    if args.workload == "uniform_100k_2d":
        a, b = hd.gen.uniform_pair(100_000, 2)
        desc = "uniform_100k_2d: A, B 100,000 i.i.d. uniform points in [0,1]^2 (seeds 1, 2)"
        variants = ["eb", "grid", "lbvh", "hybrid"]
    elif args.workload == "sphere_20m_15m_3d":
        log("generating spheres")
        a, b = hd.gen.sphere_pair()
        desc = ("sphere_20m_15m_3d: |A| 20M, |B| 15M points on the unit sphere, radial sigma "
                "0.005, B rotated 2 degrees about z")
        variants = ["lbvh", "grid", "hybrid"]
    else:
        log("generating gmm 50M")
        a, b = hd.gen.gmm_50m_256()
        desc = "gmm_50m_256_2d: as gmm_10m_2d with 50M points and 256 components"
        variants = ["hybrid", "grid", "eb"]
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    results = {}
    ref = None
    for name in variants:
        v = hd.Variant.from_string(name)
        ms_v, res_v, st_v = time_steps(hd, ctx, v, da, db, args.steps, args.warmup, learned=False)
        r = res_v[-1]
        if ref is None:
            ref = (r.cmax2_bits, r.argmax, r.nn)
        elif (r.cmax2_bits, r.argmax, r.nn) != ref:
            raise RuntimeError(f"{name} differs: {(r.cmax2_bits, r.argmax, r.nn)} vs {ref}")
        log(f"{name}: {ms_v:.3f} ms/step, evals {r.evals}")
        results[name] = {"ms_per_step": ms_v, "a_points_per_s": a.shape[0] / (ms_v / 1e3),
                         "evals": r.evals,
                         "stages_ms": {k: x["ms"] / max(1, x["calls"]) for k, x in st_v.items()}}
    if "eb" in results and args.workload == "gmm_50m_256_2d" and not args.no_learned:
        # The learned order on the larger mixture: the model trained on a
        # 1M-point subsample of A (exact NN labels), then the same EB run.
        t0 = time.perf_counter()
        model = hd.Model.train(ctx, da, db, steps=2000, batch=4096, n_samples=1_000_000)
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t0
        ms_l, res_l, st_l = time_steps(hd, ctx, hd.Variant.EB, da, db, args.steps, args.warmup,
                                       learned=True, model=model)
        r = res_l[-1]
        if (r.cmax2_bits, r.argmax, r.nn) != ref:
            raise RuntimeError("learned order changed the result")
        log(f"eb learned: {ms_l:.3f} ms/step, evals {r.evals}, train {train_s:.2f} s")
        results["eb_learned"] = {
            "ms_per_step": ms_l, "a_points_per_s": a.shape[0] / (ms_l / 1e3), "evals": r.evals,
            "model_train_s": train_s,
            "speedup_vs_plain_eb": results["eb"]["ms_per_step"] / ms_l,
            "stages_ms": {k: x["ms"] / max(1, x["calls"]) for k, x in st_l.items()}}
    best = min(results, key=lambda k: results[k]["ms_per_step"])
    ms = results[best]["ms_per_step"]
    line = {"metric": f"A-points/sec, {args.workload} directed HD", "value": a.shape[0] / (ms / 1e3),
            "unit": "A-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32", "data": "synthetic",
            "config": {"workload": desc, "best_variant": best, "variants": results,
                       "cmax2_bits": ref[0], "argmax": ref[1], "nn": ref[2]}}
    if args.workload == "uniform_100k_2d" and not args.no_cpu:
        from oracle import oracle
        cores = min(os.cpu_count() or 1, 16)
        o = oracle.eb_directed(a, b, n_threads=cores)
        if (o.cmax2_bits, o.argmax, o.nn) != ref:
            raise RuntimeError("GPU result differs from the CPU oracle")
        line["cpu_baseline"] = {"value": a.shape[0] / (o.compute_ms / 1e3), "unit": "A-points/s",
                                "cores": int(o.n_threads), "kind": "port",
                                "sample": f"whole workload, {o.compared_pairs} d2 evaluations"}
    print(json.dumps(line), flush=True)
    ctx.close()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-next", action="store_true",
                    help="skip the input-statistics and loader timings")
    ap.add_argument("--variants", default="grid,lbvh,hybrid,nn",
                    help="other variants timed on the same workload")
    ap.add_argument("--no-symmetric", action="store_true",
                    help="skip the symmetric hybrid measurement")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 3 times (counter calibration)")
    ap.add_argument("--workload", default="gmm_10m_2d",
                    choices=["gmm_10m_2d", "uniform_100k_2d", "sphere_20m_15m_3d",
                             "batch_4096_2d", "gmm_50m_256_2d"])
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch

    import hausdorffdistance_amd as hd

    if args.workload != "gmm_10m_2d":
        secondary(args)
        return

    log(f"generating gmm {args.n} x {args.n}")
    a, b = hd.gen.gmm_10m(args.n)
    da = torch.from_numpy(a).cuda()
    db = torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    ctx = hd.Context(0)
    info = hd.device_info(0)
    log(f"device {info.get('name')} {info.get('gcnArchName')} CUs={info.get('multiProcessorCount')}")

    log("plain order")
    ms_plain, res_plain, st_plain = time_steps(hd, ctx, hd.Variant.EB, da, db, args.steps,
                                               args.warmup, learned=False)
    r0 = res_plain[-1]
    log(f"plain: {ms_plain:.3f} ms/step, cmax2 bits {r0.cmax2_bits} argmax {r0.argmax} "
        f"evals {r0.evals}")
    log(f"plain stages: {json.dumps(st_plain)}")
    # The reference GPU kernel's schedule (early_break.h:117-219): one pass
    # over the pseudo-random order, cmax2 from 0, no ranking or filter.
    log("plain order, single pass")
    ms_single, res_single, _ = time_steps(hd, ctx, hd.Variant.EB, da, db, args.steps,
                                          args.warmup, learned=False, single_pass=True)
    rs = res_single[-1]
    if (rs.cmax2_bits, rs.argmax, rs.nn) != (r0.cmax2_bits, r0.argmax, r0.nn):
        raise RuntimeError("single pass differs from the pipeline")
    log(f"single pass: {ms_single:.3f} ms/step, evals {rs.evals}")
    ms_learned = None
    st_learned = None
    res_learned = None
    r1 = None
    train_s = None
    model = None
    final_loss = None
    if not args.no_learned:
        try:
            log("training the ordering model")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            model = hd.Model.train(ctx, da, db, steps=2000, batch=4096,
                                   n_samples=min(1_000_000, a.shape[0]))
            torch.cuda.synchronize()
            train_s = time.perf_counter() - t0
            final_loss = model.final_loss
            log(f"trained in {train_s:.2f} s, final loss {final_loss:.4f}")
            ms_learned, res_learned, st_learned = time_steps(
                hd, ctx, hd.Variant.EB, da, db, args.steps, args.warmup, learned=True,
                model=model)
            r1 = res_learned[-1]
            log(f"learned: {ms_learned:.3f} ms/step, evals {r1.evals}")
            log(f"learned stages: {json.dumps(st_learned)}")
            if (r1.cmax2_bits, r1.argmax) != (r0.cmax2_bits, r0.argmax):
                raise RuntimeError("learned order changed the result")
        except hd.HDError as e:
            log(f"learned order unavailable: {e}")
            ms_learned = None

    # Other variants on the same workload (same bits required).
    other = {}
    for name in args.variants.split(","):
        name = name.strip()
        if not name or name in ("eb", "none"):
            continue
        v = hd.Variant.from_string(name)
        try:
            ms_v, res_v, st_v = time_steps(hd, ctx, v, da, db, args.steps, args.warmup,
                                           learned=False)
        except hd.HDError as e:
            log(f"{name} unavailable: {e}")
            continue
        rv = res_v[-1]
        if (rv.cmax2_bits, rv.argmax, rv.nn) != (r0.cmax2_bits, r0.argmax, r0.nn):
            raise RuntimeError(f"{name} result differs from early break")
        log(f"{name}: {ms_v:.3f} ms/step, evals {rv.evals}, stages {json.dumps(st_v)}")
        other[name] = {"ms_per_step": ms_v, "a_points_per_s": a.shape[0] / (ms_v / 1e3),
                       "evals": rv.evals, "spread": spread(res_v),
                       "query_roofline": index_roofline(name, st_v),
                       "stages_ms": {k: x["ms"] / max(1, x["calls"]) for k, x in st_v.items()}}

    # Symmetric H = max(h(A->B), h(B->A)) with the hybrid selector.
    sym = None
    try:
        if args.no_symmetric:
            raise hd.HDError(0, "skipped (--no-symmetric)")
        ms_ab, res_ab, _ = time_steps(hd, ctx, hd.Variant.HYBRID, da, db, args.steps,
                                      args.warmup, learned=False)
        ms_ba, res_ba, st_ba = time_steps(hd, ctx, hd.Variant.HYBRID, db, da, args.steps,
                                          args.warmup, learned=False)
        rab, rba = res_ab[-1], res_ba[-1]
        sym = {"ms": ms_ab + ms_ba, "points_per_s": (a.shape[0] + b.shape[0]) / ((ms_ab + ms_ba) / 1e3),
               "h_ab": rab.hd, "h_ba": rba.hd, "H": max(rab.hd, rba.hd),
               "direction": 1 if rba.cmax2_bits > rab.cmax2_bits else 0,
               "ms_ab": ms_ab, "ms_ba": ms_ba,
               "ba_stages_ms": {k: x["ms"] / max(1, x["calls"]) for k, x in st_ba.items()}}
        log(f"symmetric hybrid: {sym['ms']:.3f} ms (A->B {ms_ab:.3f}, B->A {ms_ba:.3f}), H={sym['H']}")
    except hd.HDError as e:
        log(f"symmetric hybrid unavailable: {e}")

    extra = None
    if not args.no_next:
        try:
            extra = next_rows(hd, ctx, da, db, a.shape[0], b.shape[0])
            log(f"next rows: {json.dumps(extra)}")
        except hd.HDError as e:
            log(f"next rows unavailable: {e}")

    use_learned = ms_learned is not None
    ms = ms_learned if use_learned else ms_plain
    stages = st_learned if use_learned else st_plain
    n_a = a.shape[0]
    value = n_a / (ms / 1e3)

    roofline = eb_roofline(st_plain, args.steps)

    cpu = None
    host_eb = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(a, b)
        log(f"cpu: {cpu['value']:.4g} A-points/s on {cpu['cores']} threads")
        log("host early break (execution = CPU)")
        host_eb = host_early_break(hd, a, b, (r0.cmax2_bits, r0.argmax, r0.nn))
        log(f"host early break: {host_eb['ms']:.1f} ms")

    line = {
        "metric": "A-points/sec, 10M x 10M 2D HD (directed, early break), learned-order vs plain",
        "value": value,
        "unit": "A-points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"gmm_{args.n // 1_000_000}m_2d: A {n_a} points from a 64-component "
                        f"Gaussian mixture (sigma 0.01), B = A + (0.002, 0.001) with 1% "
                        f"uniform outliers; directed h(A->B), early-break variant",
            "order": "learned" if use_learned else "plain",
            "plain_ms_per_step": ms_plain,
            "plain_value": n_a / (ms_plain / 1e3),
            "plain_spread": spread(res_plain),
            # the pipeline without the network: seed, a ranking round on a
            # sample of the pseudo-random order, top completions, budgeted
            # rounds on the matrix-core filter, segmented tail
            "plain_is": "ranked pipeline (pseudo-random orders of A and B, no network)",
            "single_pass_ms_per_step": ms_single,
            "single_pass_value": n_a / (ms_single / 1e3),
            "single_pass_spread": spread(res_single),
            "single_pass_is": "one pass over the pseudo-random order, cmax2 from 0, exact "
                              "scan only: the reference GPU kernel's schedule",
            "evals_single_pass": rs.evals,
            "learned_is": "A in descending predicted NN distance (the network), B in "
                          "the stratified order (Morton ranks spread by a golden-ratio "
                          "rotation), seed of 2048, rounds of 512 and 1536 steps, tail",
            "learned_ms_per_step": ms_learned,
            "learned_spread": spread(res_learned) if use_learned else None,
            "learned_speedup": (ms_plain / ms_learned) if use_learned else None,
            "learned_speedup_vs_single_pass": (ms_single / ms_learned) if use_learned else None,
            "learned_speedup_median": (spread(res_plain)["median_ms"]
                                       / spread(res_learned)["median_ms"])
            if use_learned else None,
            "model_train_s": train_s,
            "model_final_loss": final_loss,
            "cmax2_bits": r0.cmax2_bits,
            "hd": r0.hd,
            "argmax": r0.argmax,
            "nn": r0.nn,
            "evals_plain": r0.evals,
            "evals_learned": r1.evals if r1 is not None else None,
            "stages_ms": {k: v["ms"] / max(1, v["calls"]) for k, v in stages.items()},
            "other_variants": other,
            "symmetric_hybrid": sym,
            "next_rows": extra,
            "host_early_break": host_eb,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
