#!/usr/bin/env python3
"""Benchmark of the MI355X path-tracing engine (one GPU).

Workload (default): the full-material-coverage scene — 1024x1024 Perlin
terrain + 256 displaced icospheres (subdivision 5), 7,335,938 triangles,
materials drawn from the clearcoat / sheen / anisotropy / specular /
transmission+volume / iridescence / dispersion families, 2048x1024 log-normal
IBL with a sun disc, rendered at 1920x1080, 64 spp, max depth 12, webgl-full
profile. One step = one complete 64-spp render of that frame (a job).

Prints one JSON line: Mpaths/s (value), Mrays/s, p50/p99 job latency, the
roofline of the dominant kernel (closest-hit traversal, HBM-bound, from HIP
events on the engine stream) and the CPU baseline (the oracle on the host
cores over a bounded sample of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dspbr-pt_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Mpaths/s + Mrays/s at 1080p; p50/p99 job latency, batched vs plain, 1 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(name):
    from dspt import build_bvh, scenes
    t0 = time.time()
    if name == "full_materials":
        wl = scenes.full_materials()
    elif name == "sphere_field":
        wl = scenes.sphere_field()
    else:
        raise SystemExit(f"unknown workload {name}")
    t1 = time.time()
    wl.scene.bvhNodes, wl.scene.bvhTriangleIndices = build_bvh(wl.scene.bvhPositionBuffer, wl.scene.bvhIndexBuffer)
    log(f"[bench] workload {name}: {wl.scene.num_triangles} tris, generated in {t1 - t0:.1f}s, "
        f"BVH {len(wl.scene.bvhNodes) // 16} nodes in {time.time() - t1:.1f}s")
    return wl


def frame_block(wl, total_sum):
    from dspt import pack_frame_block
    s = wl.scene
    img = wl.ibl["image"]
    return pack_frame_block(wl.camera, wl.width, wl.height, max_bounces=wl.max_bounces,
                            ibl_size=(img["width"], img["height"]), ibl_total_sum=total_sum,
                            scene_counts=(s.num_triangles, len(s.bvhNodes) // 16, len(s.bvhTriangleIndices), 0))


def cpu_baseline(wl, block, budget_s=20.0):
    """Oracle (C++ restatement, FP32) on the host cores over a growing tile of
    the same frame at 1 spp until about budget_s of wall time is spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as O
    s = wl.scene
    threads = max(1, min(16, os.cpu_count() or 1))
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer(),
                        wl.ibl, 1 if wl.profile == "webgl-full" else 0)
    W, H = wl.width, wl.height
    rows, spent, paths, frame, frames_full = 8, 0.0, 0, 1, 0
    while spent < budget_s * 0.75:
        r = min(rows, H)
        y = max(0, H // 2 - r // 2)
        t0 = time.time()
        osc.render(block, W, H, frame, 1, rect=(0, y, W, y + r), record=False, threads=threads)
        dt = time.time() - t0
        spent += dt
        paths += W * r
        if r == H:
            frame += 1
            frames_full += 1
        elif dt < budget_s / 8:
            rows *= 2
        else:
            break
    return {"value": paths / spent / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP32 restatement over centred full-width row bands (doubling from 8 rows) then "
                      f"{frames_full} whole frames of the same {W}x{H} image, {paths} paths in {spent:.1f}s"}


def calibrate_copy(mib):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    per dispatch give the scale between the counters and real HBM bytes."""
    import torch
    n = mib << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": n, "bytes_written": n, "dispatches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="full_materials")
    ap.add_argument("--spp", type=int, default=None, help="override frames per job")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    from dspt import Engine, PROFILE_FULL, PROFILE_LEAN
    wl = make_workload(args.workload)
    spp = args.spp or wl.spp
    eng = Engine(PROFILE_FULL if wl.profile == "webgl-full" else PROFILE_LEAN)
    s = wl.scene
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer())
    img = wl.ibl["image"]
    total = eng.upload_ibl(img["data"], img["width"], img["height"])
    eng.resize(wl.width, wl.height)
    block = frame_block(wl, total)
    eng.set_frame(block)
    log(f"[bench] uploaded; device memory {eng.memory_usage() / 1e9:.2f} GB")

    for i in range(args.warmup):
        t0 = time.time()
        eng.render(1, spp)
        eng.synchronize()
        log(f"[bench] warmup {i}: {time.time() - t0:.2f}s")

    eng.reset_counters()
    eng.set_kernel_timing(True)
    lat = []
    eng.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        eng.render(1, spp)
        eng.synchronize()
        lat.append(time.perf_counter() - t0)
        log(f"[bench] step {i}: {lat[-1]:.3f}s")
    eng.synchronize()
    elapsed = time.perf_counter() - t_start
    c = eng.counters()
    ks = eng.kernel_stats()
    eng.set_kernel_timing(False)

    npaths = wl.width * wl.height * spp * args.steps
    rays = c["rays_primary"] + c["rays_extension"] + c["rays_shadow"]
    mpaths = npaths / elapsed / 1e6
    mrays = rays / elapsed / 1e6
    # dominant kernel: closest-hit traversal; algorithmic bytes = 64 B per node visit + 64 B per triangle test
    te = ks["trace_ext"]
    avg_ms = te["total_ms"] / max(1, te["launches"])
    alg_bytes = 64.0 * te["node_visits"] + 64.0 * te["tri_tests"]
    bytes_per_launch = alg_bytes / max(1, te["launches"])
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    kernel_ms = {k: round(v["total_ms"], 2) for k, v in ks.items()}

    out = {
        "metric": METRIC,
        "value": round(mpaths, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name} {wl.width}x{wl.height} {spp}spp depth{wl.max_bounces} "
                               f"{s.num_triangles} tris {wl.profile}", "mode": "plain (one job at a time)"},
        "mrays_per_s": round(mrays, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 1)},
        "counters": c,
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace_ext", "avg_launch_ms": round(avg_ms, 4),
                     "alg_bytes_per_launch": round(bytes_per_launch)},
    }
    if not args.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(wl, block, args.cpu_budget)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.4f} Mpaths/s in {time.time() - t0:.1f}s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
