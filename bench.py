#!/usr/bin/env python3
"""Benchmark of the MI355X path-tracing engine (one GPU).

Workload (default): the full-material-coverage scene — 1024x1024 Perlin
terrain + 256 displaced icospheres (subdivision 5), 7,335,938 triangles,
materials drawn from the clearcoat / sheen / anisotropy / specular /
transmission+volume / iridescence / dispersion families, 2048x1024 log-normal
IBL with a sun disc, rendered at 1920x1080, 64 spp, max depth 12, webgl-full
profile. One step = one complete 64-spp render of that frame (a job).

Prints one JSON line: Mpaths/s (value), Mrays/s, p50/p99 job latency, the
roofline of the dominant kernel and the CPU baseline (the oracle on the host
cores over a bounded sample of the same workload).

The timed steps run in the engine's shipping mode (bounce loop in captured
HIP graphs, no per-launch events). A second pass of the same steps with
per-launch HIP events on the engine stream gives each kernel's average
launch duration. The roofline of the closest-hit traversal (k_trace_ext,
HBM-bound) divides its measured memory traffic per launch by that
duration: the traffic comes from two child runs of this script under
rocprofv3 (--pmc FETCH_SIZE, then --pmc WRITE_SIZE), each scaled by what the
same counter reads for a plain copy kernel of known size in the same run.
The byte model of the traversal (64 B per node visit, 64 B per triangle
test) is reported beside it as bytes served (caches included). A third child
run reads each bouncing kernel's L2 hit rate (TCC_HIT/TCC_MISS), VALU busy
and lanes active (VALUBusy, VALUUtilization).

More objects:
  batched  the uneven 32-scene mix (10k-2M triangles, 256^2-1024^2,
           16-256 spp, 1024x512 IBL each) rendered through the shared path
           pool (least remaining work first; FIFO and round robin beside
           it) vs one job at a time through the single-job path, in
           frame-synchronous passes of a full pool, and with one frame per
           pass (the reference's render loop): Mpaths/s and p50/p99
           time-to-N-spp;
  serving  Poisson arrival streams over the full-material scene with
           resolution/spp mixes, admission by a device-memory budget,
           p50/p99 time-to-N-spp and time-to-first-frame against Mpaths/s;
  instanced  a top-level BVH over 4096 rigid instances of 16 meshes
           (~80M instanced triangles), transmission/volume/dispersion
           materials, 8192x4096 IBL, 3840x2160 at 256 spp, depth 16: alone
           and as one long job beside small-job traffic.
"""
import argparse
import csv
import glob
import json
import math
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dspbr-pt_amd"))

HBM_PEAK_GBS = 8000.0
COPY_CALIB_BYTES = 1 << 30  # the plain copy that calibrates the memory counters
PMC_KERNELS = ("k_trace_ext", "k_shade", "k_trace_shadow")
DEFAULT_MAX_PATHS = 64 << 20
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
    threads = host_threads()
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
                      f"{frames_full} whole frames of the same {W}x{H} image, {paths} paths in {spent:.1f}s, "
                      f"{threads} threads = every core this process may use (affinity "
                      f"{len(os.sched_getaffinity(0))}, os.cpu_count {os.cpu_count()}, "
                      f"OMP_NUM_THREADS {os.environ.get('OMP_NUM_THREADS')})"}


def host_threads():
    """The host cores this process may use: the CPU share the machine grants
    it (OMP_NUM_THREADS where the machine sets it, else the affinity mask)."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        return min(int(omp), aff)
    return aff


def batched_vs_plain(args):
    """Config-3 style comparison; every job is rendered by both schedules and
    one job's accumulation is compared bit for bit."""
    from dspt import (Engine, ORDER_BUCKETS, ORDER_NONE, ORDER_SHADE, PROFILE_FULL, SCHED_PASS, SCHED_STREAM, POLICY_FIFO,
                      POLICY_RR, POLICY_SRPT, build_bvh, scenes)
    from dspt.batch import BatchRenderer, JobSpec
    t0 = time.time()
    mix = scenes.batch_mix(seed=3, n_jobs=args.batch_jobs)
    for w in mix:
        w.scene.bvhNodes, w.scene.bvhTriangleIndices = build_bvh(w.scene.bvhPositionBuffer, w.scene.bvhIndexBuffer)
    log(f"[bench] batched mix: {len(mix)} scenes, {sum(w.scene.num_triangles for w in mix)} tris, "
        f"{sum(w.width * w.height * w.spp for w in mix) / 1e6:.1f} M paths, built in {time.time() - t0:.1f}s")
    eng = Engine(PROFILE_FULL)
    eng.set_schedule(SCHED_STREAM, 0, 0, 2.0)  # rings of 2 pools: the pool is shared by many jobs
    b = BatchRenderer(eng)
    specs = []
    for k, w in enumerate(mix):
        b.add_scene(k, w.scene)
        b.add_ibl(k, w.ibl)
        specs.append(JobSpec(k, k, w.camera, w.width, w.height, w.spp, max_bounces=w.max_bounces))
    total = sum(s.paths for s in specs)

    # autotune the pool on the first jobs' early chunks
    sub = specs[: min(8, len(specs))]
    best, table = b.autotune(lambda: [b.create_job(s) for s in sub],
                             candidates=[dict(max_paths=mp, trace_blocks=tb, chunk_iters=2, ray_order=ORDER_SHADE)
                                         for mp in (16 << 20, 32 << 20, 64 << 20) for tb in (1024, 2048)]
                             + [dict(max_paths=16 << 20, trace_blocks=2048, chunk_iters=2, ray_order=ro)
                                for ro in (ORDER_BUCKETS, ORDER_NONE)],
                             chunks=16, warm_chunks=4)
    log(f"[bench] autotune: best {best}; " + ", ".join(f"{r:.0f}" for r, _ in table))

    def run(mode_fn, **kw):
        done_at = {}
        eng.synchronize()
        eng.reset_counters()
        eng.set_kernel_timing(True)
        t_start = time.perf_counter()
        mode_fn(done_at, t_start, **kw)
        eng.synchronize()
        el = time.perf_counter() - t_start
        run.kms = {k: round(v["total_ms"], 1) for k, v in eng.kernel_stats().items()}
        run.cnt = eng.counters()
        eng.set_kernel_timing(False)
        lat = np.array(sorted(done_at.values()))
        return el, lat

    def batched(done_at, t_start, policy=POLICY_FIFO):
        ids = [b.create_job(s) for s in specs]
        eng.set_schedule(SCHED_STREAM, 0, best["chunk_iters"], 2.0)
        eng.set_policy(policy)
        live = set(ids)
        while live:
            eng.jobs_step(sorted(live))
            now = time.perf_counter() - t_start
            for j in list(live):
                d, tot = eng.job_status(j)
                if d >= tot:
                    done_at[j] = now
                    live.discard(j)
        run.ids = ids

    def single(done_at, t_start):
        # the caller's plain path: one job at a time through the single-job
        # API (each job alone in the streaming pool, as dspt_render renders)
        ids = []
        eng.set_schedule(SCHED_STREAM, 0, best["chunk_iters"], 2.0)
        for s in specs:
            t0 = time.perf_counter()
            j = b.create_job(s)
            eng.jobs_run([j])
            eng.synchronize()
            done_at[j] = time.perf_counter() - t_start
            single.alone[j] = (s.paths, time.perf_counter() - t0)
            ids.append(j)
        run.ids = ids
    single.alone = {}

    def plain(done_at, t_start, pass_frames=0):
        ids = []
        eng.set_schedule(SCHED_PASS, pass_frames)
        for s in specs:
            j = b.create_job(s)
            eng.jobs_run([j])
            eng.synchronize()
            done_at[j] = time.perf_counter() - t_start
            ids.append(j)
        eng.set_schedule(SCHED_STREAM)
        run.ids = ids

    # batched headline: all jobs in the pool, least remaining work first
    el_b, lat_b = run(batched, policy=POLICY_SRPT)
    ids_b, kms_b, cnt_b = run.ids, run.kms, run.cnt
    acc_b = eng.job_read_accum(ids_b[0])
    for j in ids_b:
        eng.job_destroy(j)
    el_r, lat_r = run(batched, policy=POLICY_RR)
    kms_r = run.kms
    for j in run.ids:
        eng.job_destroy(j)
    el_q, lat_q = run(batched, policy=POLICY_FIFO)
    acc_q = eng.job_read_accum(run.ids[0])
    kms_q = run.kms
    for j in run.ids:
        eng.job_destroy(j)
    eng.set_policy(POLICY_RR)
    el_p, lat_p = run(plain)
    ids_p, kms_p = run.ids, run.kms
    acc_p = eng.job_read_accum(ids_p[0])
    for j in ids_p:
        eng.job_destroy(j)
    el_s, lat_s = run(single)
    kms_s = run.kms
    acc_s = eng.job_read_accum(run.ids[0])
    alone = [single.alone[j] for j in run.ids]
    for j in run.ids:
        eng.job_destroy(j)
    # the reference's loop structure: one job at a time, one frame per pass
    el_f, lat_f = run(plain, pass_frames=1)
    kms_f = run.kms
    acc_f = eng.job_read_accum(run.ids[0])
    for j in run.ids:
        eng.job_destroy(j)
    q = lambda a, p: round(float(np.percentile(a, p)) * 1e3, 1)
    res = {
        "workload": f"{len(specs)} scenes {sum(w.scene.num_triangles for w in mix)} tris, "
                    f"{min(s.width for s in specs)}^2-{max(s.width for s in specs)}^2, "
                    f"{min(s.frames for s in specs)}-{max(s.frames for s in specs)} spp, depth 8, webgl-full",
        "paths": total,
        "batched_mpaths_per_s": round(total / el_b / 1e6, 2),
        "plain_mpaths_per_s": round(total / el_p / 1e6, 2),
        "speedup": round(el_p / el_b, 3),
        "batched_latency_ms": {"p50": q(lat_b, 50), "p99": q(lat_b, 99)},
        "plain_latency_ms": {"p50": q(lat_p, 50), "p99": q(lat_p, 99)},
        "batched_policy": "srpt",
        "round_robin": {"mpaths_per_s": round(total / el_r / 1e6, 2),
                        "latency_ms": {"p50": q(lat_r, 50), "p99": q(lat_r, 99)}},
        "fifo": {"mpaths_per_s": round(total / el_q / 1e6, 2),
                 "latency_ms": {"p50": q(lat_q, 50), "p99": q(lat_q, 99)},
                 "mode": "all jobs in the pool, whole runs of frames in submission order (FIFO)"},
        "kernel_ms": {"batched": kms_b, "round_robin": kms_r, "fifo": kms_q, "plain": kms_p},
        "plain_mode": "one job at a time, frame-synchronous passes of up to a full pool",
        "frame_loop_mode": "one job at a time, one frame per pass (the reference's render loop)",
        "identical_accumulation": bool(np.array_equal(acc_b, acc_p) and np.array_equal(acc_b, acc_f)
                                       and np.array_equal(acc_b, acc_s) and np.array_equal(acc_b, acc_q)),
        "single_job_api": {"mpaths_per_s": round(total / el_s / 1e6, 2),
                           "latency_ms": {"p50": q(lat_s, 50), "p99": q(lat_s, 99)}, "kernel_ms": kms_s,
                           "mode": "one job at a time, each alone in the streaming pool (the dspt_render path)",
                           "jobs_below_one_pool": sum(1 for p_, _ in alone if p_ < best["max_paths"]),
                           "work_share_below_one_pool": round(sum(p_ for p_, _ in alone if p_ < best["max_paths"])
                                                              / total, 4),
                           "time_share_below_one_pool": round(sum(t_ for p_, t_ in alone if p_ < best["max_paths"])
                                                              / sum(t_ for _, t_ in alone), 4)},
        "speedup_vs_single_job_api": round(el_s / el_b, 3),
        "work_per_path": {"rays": round((cnt_b["rays_primary"] + cnt_b["rays_extension"] + cnt_b["rays_shadow"])
                                        / total, 3),
                          "node_visits_per_ray": round(cnt_b["node_visits"] / max(1, cnt_b["rays_primary"]
                                                       + cnt_b["rays_extension"] + cnt_b["rays_shadow"]), 2),
                          "tri_tests_per_ray": round(cnt_b["tri_tests"] / max(1, cnt_b["rays_primary"]
                                                     + cnt_b["rays_extension"] + cnt_b["rays_shadow"]), 2)},
        "autotuned": best,
        "frame_loop": {"mpaths_per_s": round(total / el_f / 1e6, 2),
                       "latency_ms": {"p50": q(lat_f, 50), "p99": q(lat_f, 99)},
                       "speedup_batched": round(el_f / el_b, 3), "kernel_ms": kms_f},
    }
    log(f"[bench] batched {res['batched_mpaths_per_s']} vs plain {res['plain_mpaths_per_s']} vs single-job API "
        f"{res['single_job_api']['mpaths_per_s']} Mpaths/s")
    return res


def serving(wl, args):
    """Poisson arrivals over the resident full-material scene; each request
    draws a resolution and spp from the mixes below."""
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import BatchRenderer, JobSpec, Server, latency_stats
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    b.add_scene(0, wl.scene)
    b.add_ibl(0, wl.ibl)
    resolutions = [(256, 256), (512, 512), (1024, 576), (1920, 1080)]
    spps = [4, 8, 16, 32]
    # warm the pool and the kernels
    w0 = b.create_job(JobSpec(0, 0, wl.camera, 256, 256, 4, max_bounces=wl.max_bounces))
    eng.jobs_run([w0])
    eng.job_destroy(w0)
    eng.set_memory_budget(eng.memory_usage() + (6 << 30))
    out = []
    for rate in args.serve_rates:
        rng = np.random.default_rng(int(rate))
        t, reqs = 0.0, []
        while True:
            t += rng.exponential(1.0 / rate)
            if t > args.serve_seconds:
                break
            w, h = resolutions[int(rng.integers(0, len(resolutions)))]
            reqs.append(JobSpec(0, 0, wl.camera, w, h, spps[int(rng.integers(0, len(spps)))], max_bounces=wl.max_bounces,
                                arrival=t))
        if not reqs:
            continue
        c0 = eng.counters()["paths"]
        t0 = time.perf_counter()
        rec = Server(b).run(reqs)
        el = time.perf_counter() - t0
        st = latency_stats(rec)
        paths = sum(r.paths for r in reqs)
        st.update({"rate_jobs_per_s": rate, "mpaths_per_s": round(paths / el / 1e6, 2), "wall_s": round(el, 2),
                   "paths": paths})
        assert eng.counters()["paths"] - c0 >= paths
        st = {k: (round(v, 1) if isinstance(v, float) else v) for k, v in st.items()}
        log(f"[bench] serving {rate}/s: {st}")
        out.append(st)
    eng.set_memory_budget(0)
    return {"scene": f"{wl.name} {wl.scene.num_triangles} tris, depth {wl.max_bounces}",
            "resolutions": resolutions, "spp": spps, "budget_bytes_for_jobs": 6 << 30, "sweep": out}


def instanced(args):
    """Large instanced scene: a top-level BVH over 4096 seeded rigid instances
    of 16 base meshes (4k-64k triangles each), transmission+volume and
    dispersion materials, 8192x4096 log-normal IBL, 3840x2160, depth 16.
    Rendered alone as one long job, then again as a long job while small
    jobs from a batched mix arrive (Poisson) and share the pool."""
    from dspt import Engine, PROFILE_FULL, POLICY_SRPT, build_bvh, scenes
    from dspt.batch import BatchRenderer, JobSpec, Server, latency_stats
    t0 = time.time()
    wl = scenes.instanced_scene(spp=args.inst_spp)
    s = wl.scene
    log(f"[bench] instanced: {s.num_instances} instances of {len(s.mesh_count)} meshes, {s.scene.num_triangles} mesh tris, "
        f"{s.instanced_triangles} instanced tris, generated in {time.time() - t0:.1f}s")
    eng = Engine(PROFILE_FULL)
    eng.set_launch_config(max_paths=args.inst_max_paths)
    b = BatchRenderer(eng)
    t1 = time.time()
    b.add_scene("inst", s)
    build_s = time.time() - t1
    b.add_ibl("inst", wl.ibl)
    w0 = b.create_job(JobSpec("inst", "inst", wl.camera, 256, 256, 2, max_bounces=wl.max_bounces))
    eng.jobs_run([w0])
    eng.job_destroy(w0)
    long_spec = JobSpec("inst", "inst", wl.camera, wl.width, wl.height, wl.spp, max_bounces=wl.max_bounces)

    eng.reset_counters()
    eng.set_kernel_timing(True)
    eng.synchronize()
    t1 = time.perf_counter()
    j = b.create_job(long_spec)
    eng.jobs_run([j])
    eng.synchronize()
    el = time.perf_counter() - t1
    c = eng.counters()
    ks = eng.kernel_stats()
    eng.set_kernel_timing(False)
    eng.job_destroy(j)
    rays = c["rays_primary"] + c["rays_extension"] + c["rays_shadow"]
    alone = {"mpaths_per_s": round(long_spec.paths / el / 1e6, 2), "mrays_per_s": round(rays / el / 1e6, 2),
             "seconds": round(el, 2), "kernel_ms": {k: round(v["total_ms"], 1) for k, v in ks.items()},
             "node_visits_per_ray": round(c["node_visits"] / max(1, rays), 1),
             "tri_tests_per_ray": round(c["tri_tests"] / max(1, rays), 1)}
    log(f"[bench] instanced alone: {alone}")

    # the same long job beside small batched-mix traffic, shortest-remaining-first
    mix = scenes.batch_mix(seed=4, n_jobs=8, tri_range=(10_000, 200_000), side_range=(256, 512), spp_range=(16, 64))
    for k, w in enumerate(mix):
        w.scene.bvhNodes, w.scene.bvhTriangleIndices = build_bvh(w.scene.bvhPositionBuffer, w.scene.bvhIndexBuffer)
        b.add_scene(k, w.scene)
        b.add_ibl(k, w.ibl)
    rng = np.random.default_rng(9)
    reqs, t = [JobSpec("inst", "inst", wl.camera, wl.width, wl.height, wl.spp, max_bounces=wl.max_bounces)], 0.0
    while True:
        t += rng.exponential(1.0 / args.inst_traffic_rate)
        if t > 0.8 * el:
            break
        k = int(rng.integers(0, len(mix)))
        w = mix[k]
        reqs.append(JobSpec(k, k, w.camera, w.width, w.height, w.spp, max_bounces=w.max_bounces, arrival=t))
    eng.set_policy(POLICY_SRPT)
    t1 = time.perf_counter()
    rec = Server(b).run(reqs)
    el2 = time.perf_counter() - t1
    paths = sum(r.paths for r in reqs)
    small = latency_stats(rec[1:]) if len(rec) > 1 else None
    mixed = {"mpaths_per_s": round(paths / el2 / 1e6, 2), "seconds": round(el2, 2), "small_jobs": len(reqs) - 1,
             "small_rate_jobs_per_s": args.inst_traffic_rate, "policy": "srpt",
             "long_job_latency_ms": round((rec[0]["done"] - rec[0]["arrival"]) * 1e3, 1),
             "small_latency": {k: (round(v, 1) if isinstance(v, float) else v) for k, v in small.items()} if small else None}
    log(f"[bench] instanced + traffic: {mixed}")
    return {"scene": f"{s.num_instances} instances of {len(s.mesh_count)} meshes, {s.instanced_triangles} instanced tris "
                     f"({s.scene.num_triangles} stored), {wl.width}x{wl.height} {wl.spp}spp depth{wl.max_bounces}, "
                     f"IBL {wl.ibl['image']['width']}x{wl.ibl['image']['height']}",
            "bvh_build_s": round(build_s, 2), "alone": alone, "with_traffic": mixed}


def bvh_builds(wl, spp):
    from dspt import Engine, PROFILE_FULL, build_bvh
    s = wl.scene
    t0 = time.perf_counter()
    build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    host_ms = (time.perf_counter() - t0) * 1e3
    eng = Engine(PROFILE_FULL)
    trees, gms = {}, {}
    for name, method in (("gpu_lbvh", 0), ("gpu_ploc", 1)):
        eng.build_bvh_gpu(s.vertexBuffer, s.triangleIndexBuffer, 4, method)  # warm (code objects, allocator)
        best = None
        for _ in range(3):
            nodes, bidx, ms = eng.build_bvh_gpu(s.vertexBuffer, s.triangleIndexBuffer, 4, method)
            best = ms if best is None else min(best, ms)
        trees[name], gms[name] = (nodes, bidx), best
    img = wl.ibl["image"]
    rate = {}
    for name, (nd, bi) in [("host_sah", (s.bvhNodes, s.bvhTriangleIndices))] + list(trees.items()):
        e = Engine(PROFILE_FULL)
        e.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nd, bi, s.getFlatMaterialBuffer())
        tot = e.upload_ibl(img["data"], img["width"], img["height"])
        e.resize(wl.width, wl.height)
        e.set_frame(frame_block(wl, tot))
        e.render(1, spp)  # warm: pool, rings, code objects
        e.synchronize()
        t0 = time.perf_counter()
        e.render(1, spp)
        e.synchronize()
        rate[name] = round(wl.width * wl.height * spp / (time.perf_counter() - t0) / 1e6, 1)
        del e
    res = {"triangles": s.num_triangles, "build_ms": {"host_sah": round(host_ms, 1),
                                                      **{k: round(v, 2) for k, v in gms.items()}},
           "nodes": {"host_sah": len(s.bvhNodes) // 16, **{k: len(v[0]) // 16 for k, v in trees.items()}},
           "mpaths_per_s": rate}
    log(f"[bench] bvh builds: {res}")
    return res


def pmc_child(args):
    """Runs under rocprofv3 --pmc: the plain copy that calibrates the counter,
    then one render of the workload with the parent's launch configuration."""
    from dspt import Engine, PROFILE_FULL, PROFILE_LEAN
    from dspt._lib import selftest_copy
    selftest_copy(COPY_CALIB_BYTES, 3)
    wl = make_workload(args.workload)
    spp = args.spp or wl.spp
    eng = Engine(PROFILE_FULL if wl.profile == "webgl-full" else PROFILE_LEAN)
    s = wl.scene
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer())
    img = wl.ibl["image"]
    total = eng.upload_ibl(img["data"], img["width"], img["height"])
    eng.resize(wl.width, wl.height)
    eng.set_frame(frame_block(wl, total))
    eng.set_launch_config(max_paths=args.max_paths, shade_blocks=args.shade_blocks)
    eng.set_kernel_timing(True)  # eager launches: one profiled dispatch per kernel launch
    eng.render(1, spp)
    eng.synchronize()
    eng.close()
    log("[bench] pmc child: rendered one frame")


def rocprofv3_path():
    p = shutil.which("rocprofv3")
    if p:
        return p
    p = "/opt/rocm/bin/rocprofv3"
    return p if os.path.exists(p) else None


def _counter_means(outdir, counter):
    """Mean value per dispatch of `counter` for each kernel of interest and
    for the calibration copy, from rocprofv3's counter_collection.csv."""
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"no counter_collection.csv under {outdir}")
    per = {}
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "")
                if row.get("Counter_Name", counter) != counter:
                    continue
                val = float(row.get("Counter_Value", row.get(counter, "nan")))
                key = next((k for k in PMC_KERNELS + ("k_copy_calib",) if k in name), None)
                if key is None:
                    continue
                disp = row.get("Dispatch_Id") or row.get("Correlation_Id") or str(len(per.get(key, {})))
                per.setdefault(key, {})
                per[key][disp] = per[key].get(disp, 0.0) + val  # sum of per-instance rows of one dispatch
    return {k: (sum(v.values()) / len(v), len(v)) for k, v in per.items()}


def _pmc_run(args, counters):
    """One child run of this script under rocprofv3 --pmc with `counters`
    (all within one pass of the hardware) -> {counter: {kernel: (mean per
    dispatch, dispatches)}}, or an error string."""
    prof = rocprofv3_path()
    if not prof:
        return None, "rocprofv3 not found"
    out = tempfile.mkdtemp(prefix="dspt_pmc_")
    cmd = [prof, "--pmc", *counters, "-d", out, "-o", "pmc", "--output-format", "csv", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child", "--workload", args.workload,
           "--max-paths", str(args.max_paths), "--shade-blocks", str(args.shade_blocks)]
    if args.spp:
        cmd += ["--spp", str(args.spp)]
    log(f"[bench] counters {' '.join(counters)}: child run under rocprofv3")
    t0 = time.time()
    try:
        try:
            r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=None, timeout=300, cwd=ROOT)
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 {' '.join(counters)} run timed out"
        if r.returncode != 0 and not glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
            return None, f"rocprofv3 {' '.join(counters)} run failed ({r.returncode})"
        means = {c: _counter_means(out, c) for c in counters}
    except (RuntimeError, ValueError) as e:
        return None, str(e)
    finally:
        shutil.rmtree(out, ignore_errors=True)
    log(f"[bench] counters {' '.join(counters)} in {time.time() - t0:.1f}s")
    return means, None


def measure_traffic(args):
    """Memory traffic per launch of the bouncing kernels, in bytes, from two
    child runs under rocprofv3 (FETCH_SIZE, then WRITE_SIZE; one counter per
    run, each within one pass of the hardware), scaled by the same counters
    of the plain copy: bytes = counter x (copy bytes / copy counter)."""
    means = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        m, err = _pmc_run(args, [counter])
        if err:
            return None, err
        means[counter] = m[counter]
    f, w = means["FETCH_SIZE"], means["WRITE_SIZE"]
    if "k_copy_calib" not in f or "k_copy_calib" not in w:
        return None, "calibration copy not seen by the profiler"
    kf = COPY_CALIB_BYTES / (f["k_copy_calib"][0] * 1024.0)
    kw = COPY_CALIB_BYTES / (w["k_copy_calib"][0] * 1024.0)
    res = {"fetch_scale": round(kf, 4), "write_scale": round(kw, 4), "copy_bytes": COPY_CALIB_BYTES, "kernels": {}}
    for k in PMC_KERNELS:
        if k in f and k in w:
            rd, wr = f[k][0] * 1024.0 * kf, w[k][0] * 1024.0 * kw
            res["kernels"][k] = {"read_bytes": round(rd), "written_bytes": round(wr), "traffic_bytes": round(rd + wr),
                                 "dispatches": f[k][1]}
    return res, None


def measure_units(args):
    """L2 hit rate and VALU busy / lanes active of the bouncing kernels (one
    more child run): what bounds the traversal against the L2 and the shading
    against the VALU."""
    m, err = _pmc_run(args, ["TCC_HIT_sum", "TCC_MISS_sum", "VALUBusy", "VALUUtilization"])
    if err:
        return None, err
    res = {}
    for k in PMC_KERNELS:
        try:
            hit, miss = m["TCC_HIT_sum"][k][0], m["TCC_MISS_sum"][k][0]
            res[k] = {"l2_hit_rate": round(hit / max(1.0, hit + miss), 4), "valu_busy_pct": round(m["VALUBusy"][k][0], 1),
                      "lanes_active_pct": round(m["VALUUtilization"][k][0], 1)}
        except KeyError:
            continue
    return res, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="full_materials")
    ap.add_argument("--spp", type=int, default=None, help="override frames per job")
    ap.add_argument("--max-paths", type=int, default=DEFAULT_MAX_PATHS,
                    help="path pool of the single-job leg (64 Mi paths, about 20 GB of the 288 GB HBM)")
    ap.add_argument("--shade-blocks", type=int, default=0,
                    help="grid of the shading and refill kernels in the single-job leg (0: the engine default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-plain comparison")
    ap.add_argument("--batch-jobs", type=int, default=32)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving sweep")
    ap.add_argument("--serve-rates", type=float, nargs="*", default=[5.0, 25.0, 100.0, 200.0])
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--no-instanced", action="store_true", help="skip the large instanced scene")
    ap.add_argument("--no-bvh-build", action="store_true", help="skip the host vs GPU BVH build comparison")
    ap.add_argument("--inst-spp", type=int, default=256)
    ap.add_argument("--inst-max-paths", type=int, default=16 << 20, help="path pool of the instanced leg")
    ap.add_argument("--inst-traffic-rate", type=float, default=10.0, help="small jobs per second beside the long job")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the two rocprofv3 child runs that measure the traversal's memory traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine is measured on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    if args.pmc_child:
        pmc_child(args)
        return
    # counters first, in child processes, before this process touches the GPU
    traffic, traffic_err = (None, "skipped (--no-pmc)") if args.no_pmc else measure_traffic(args)
    if traffic_err:
        log(f"[bench] memory counters unavailable: {traffic_err}")
    units, units_err = (None, "skipped (--no-pmc)") if args.no_pmc else measure_units(args)
    if units_err:
        log(f"[bench] cache and VALU counters unavailable: {units_err}")

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
    eng.set_launch_config(max_paths=args.max_paths, shade_blocks=args.shade_blocks)
    log(f"[bench] uploaded; device memory {eng.memory_usage() / 1e9:.2f} GB; path pool {args.max_paths}")

    for i in range(args.warmup):
        t0 = time.time()
        eng.render(1, spp)
        eng.synchronize()
        log(f"[bench] warmup {i}: {time.time() - t0:.2f}s")

    # timed steps: the shipping launch mode (captured bounce graphs, no events)
    eng.reset_counters()
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

    # timing pass: the same steps again with HIP events on the engine stream
    # around every launch (eager launches), for per-kernel durations
    eng.reset_counters()
    eng.set_kernel_timing(True)
    eng.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        eng.render(1, spp)
        eng.synchronize()
    timing_elapsed = time.perf_counter() - t1
    ks = eng.kernel_stats()
    c_timing = eng.counters()
    eng.set_kernel_timing(False)
    log(f"[bench] timing pass: {timing_elapsed / args.steps:.3f}s per step")

    npaths = wl.width * wl.height * spp * args.steps
    rays = c["rays_primary"] + c["rays_extension"] + c["rays_shadow"]
    mpaths = npaths / elapsed / 1e6
    mrays = rays / elapsed / 1e6
    # dominant kernel: closest-hit traversal (HBM-bound)
    te = ks["trace_ext"]
    avg_ms = te["total_ms"] / max(1, te["launches"])
    served = (64.0 * te["node_visits"] + 64.0 * te["tri_tests"]) / max(1, te["launches"])
    kernel_ms = {k: round(v["total_ms"], 2) for k, v in ks.items()}
    tr = traffic["kernels"].get("k_trace_ext") if traffic else None
    hbm = tr["traffic_bytes"] if tr else None
    achieved = hbm / (avg_ms * 1e-3) / 1e9 if (hbm and avg_ms > 0) else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": hbm,
                "kernel": "k_trace_ext", "avg_launch_ms": round(avg_ms, 4), "launches": int(te["launches"]),
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE child runs of this script, "
                                   f"scaled by a {COPY_CALIB_BYTES >> 20} MiB copy" if tr else traffic_err),
                "bytes_served_per_launch": round(served),
                "served_gbs": round(served / (avg_ms * 1e-3) / 1e9, 1) if avg_ms > 0 else None,
                "served_model": "64 B per node visit + 64 B per triangle test (caches included)"}
    if traffic:
        per = {}
        for k, cls in (("k_trace_ext", "trace_ext"), ("k_shade", "shade"), ("k_trace_shadow", "trace_shadow")):
            t = traffic["kernels"].get(k)
            st = ks[cls]
            ms = st["total_ms"] / max(1, st["launches"])
            if t and ms > 0:
                per[k] = dict(t, avg_launch_ms=round(ms, 4), gbs=round(t["traffic_bytes"] / (ms * 1e-3) / 1e9, 1),
                              frac=round(t["traffic_bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
        roofline["kernels"] = per
        roofline["calibration"] = {k: traffic[k] for k in ("fetch_scale", "write_scale", "copy_bytes")}
    if units:
        for k, v in units.items():
            roofline.setdefault("kernels", {}).setdefault(k, {}).update(v)
        if "k_trace_ext" in units:
            roofline["l2_hit_rate"] = units["k_trace_ext"]["l2_hit_rate"]

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
                               f"{s.num_triangles} tris {wl.profile}",
                   "mode": "single job, streaming path pool, bounce loop in HIP graphs"},
        "mrays_per_s": round(mrays, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 1),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 1)},
        "counters": c,
        "work_per_path": {"rays": round(rays / npaths, 3), "node_visits_per_ray": round(c["node_visits"] / max(1, rays), 2),
                          "tri_tests_per_ray": round(c["tri_tests"] / max(1, rays), 2)},
        "timing_pass": {"ms_per_step": round(timing_elapsed / args.steps * 1e3, 2),
                        "identical_counters": c_timing == c,
                        "mode": "same steps, eager launches with HIP events around each"},
        "kernel_ms": kernel_ms,
        "roofline": roofline,
    }
    # display chain (ACES + gamma + FXAA) on the rendered frame; algorithmic
    # bytes per launch = 16 B accumulation read + 4 B canvas written per pixel
    eng.reset_counters()
    eng.set_kernel_timing(True)
    for _ in range(20):
        eng.postprocess(3, 1.0, True, True)
    kp = eng.kernel_stats()["post"]
    eng.set_kernel_timing(False)
    post_ms = kp["total_ms"] / max(1, kp["launches"])
    out["post"] = {"kernel": "k_postprocess", "ops": "AcesFilm + gamma + FXAA, 8-bit canvas",
                   "avg_launch_ms": round(post_ms, 4),
                   "achieved_gbs": round(20.0 * wl.width * wl.height / (post_ms * 1e-3) / 1e9, 1) if post_ms > 0 else None}
    log(f"[bench] post: {out['post']}")
    # BVH builds of the default scene: host binned SAH vs the GPU linear BVH,
    # and the render rate over each tree
    out["bvh_build"] = None if args.no_bvh_build else bvh_builds(wl, spp)
    out["batched"] = None if args.no_batched else batched_vs_plain(args)
    out["serving"] = None if args.no_serving else serving(wl, args)
    out["instanced"] = None if args.no_instanced else instanced(args)
    if not args.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(wl, block, args.cpu_budget)
        log(f"[bench] cpu baseline {out['cpu_baseline']['value']:.4f} Mpaths/s in {time.time() - t0:.1f}s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
