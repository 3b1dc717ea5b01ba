"""Batched and served rendering on top of the engine's job API.

BatchRenderer  many scenes / cameras / resolutions / frame counts rendered
               together through one path pool (`dspt_jobs_run` with the
               STREAM schedule: finished paths are refilled with samples of
               any job on the next bounce), with the pool size, traversal
               grid, chunk length and ray order chosen by a small autotuner.
               `run_plain` is the one-job-at-a-time loop with
               frame-synchronous passes (PASS schedule), optionally one frame
               per pass as in the reference's render loop.
Server         continuous batching: requests arrive over time, are admitted
               when the device-memory budget allows (the engine refuses a job
               with DSPT_ERR_MEMORY otherwise and the request waits), share
               the pool with the jobs already running, and report progressive
               per-frame completion; latency = arrival -> last frame folded.
"""
import time

import numpy as np

from . import _lib
from .renderer import pack_frame_block

ERR_MEMORY = -4


class JobSpec:
    """One render request: scene key, IBL key, camera, resolution, frames."""

    def __init__(self, scene, ibl, camera, width, height, frames, max_bounces=8, arrival=0.0, record=False, **frame_kw):
        self.scene, self.ibl, self.camera = scene, ibl, camera
        self.width, self.height, self.frames, self.max_bounces = width, height, frames, max_bounces
        self.arrival, self.record, self.frame_kw = arrival, record, frame_kw

    @property
    def paths(self):
        return self.width * self.height * self.frames


class BatchRenderer:
    def __init__(self, engine):
        self.eng = engine
        self.scenes = {}  # key -> (id, counts)
        self.ibls = {}    # key -> (id, (w, h), total)

    def add_scene(self, key, scene, gpu_bvh=False):
        """Make a scene resident. Without a prebuilt tree the BVH is built on
        the host (binned SAH), or on the GPU (PLOC, milliseconds) with
        gpu_bvh=True."""
        from ._lib import build_bvh
        from .scenes import InstancedScene
        if isinstance(scene, InstancedScene):
            sid = self.eng.scene_create_instanced(*scene.engine_args())
            self.scenes[key] = (sid, (scene.scene.num_triangles, 0, 0))
            return sid
        if gpu_bvh and scene.bvhNodes is None:
            sid, _ = self.eng.scene_create_gpu_bvh(scene.vertexBuffer, scene.triangleIndexBuffer,
                                                   scene.getFlatMaterialBuffer())
            if getattr(scene, "num_textures", 0):
                dims, texels = scene.getTextureArrays()
                self.eng.scene_set_textures(sid, scene.getFlatTextureInfoBuffer(), dims, texels)
            self.scenes[key] = (sid, (scene.num_triangles, 0, scene.num_triangles))
            return sid
        nodes, bidx = scene.bvhNodes, scene.bvhTriangleIndices
        if nodes is None:
            nodes, bidx = build_bvh(scene.bvhPositionBuffer, scene.bvhIndexBuffer)
        sid = self.eng.scene_create(scene.vertexBuffer, scene.triangleIndexBuffer, nodes, bidx,
                                    scene.getFlatMaterialBuffer())
        if getattr(scene, "num_textures", 0):
            dims, texels = scene.getTextureArrays()
            self.eng.scene_set_textures(sid, scene.getFlatTextureInfoBuffer(), dims, texels)
        self.scenes[key] = (sid, (scene.num_triangles, len(nodes) // 16, len(bidx)))
        return sid

    def add_ibl(self, key, ibl):
        img = ibl["image"]
        iid, tot = self.eng.ibl_create(img["data"], img["width"], img["height"])
        self.ibls[key] = (iid, (img["width"], img["height"]), tot)
        return iid

    def block(self, spec):
        sid, counts = self.scenes[spec.scene]
        iid, isize, tot = self.ibls[spec.ibl] if spec.ibl is not None else (-1, (0, 0), 0.0)
        kw = dict(spec.frame_kw)
        return pack_frame_block(spec.camera, spec.width, spec.height, max_bounces=spec.max_bounces, ibl_size=isize,
                                ibl_total_sum=tot, scene_counts=(counts[0], counts[1], counts[2], 0),
                                use_ibl=kw.pop("use_ibl", spec.ibl is not None),
                                show_background=kw.pop("show_background", spec.ibl is not None), **kw)

    def create_job(self, spec):
        sid = self.scenes[spec.scene][0]
        iid = self.ibls[spec.ibl][0] if spec.ibl is not None else -1
        return self.eng.job_create(sid, iid, self.block(spec), spec.width, spec.height, spec.frames, spec.record)

    def run_batched(self, job_ids):
        """All listed jobs through the pool together (STREAM schedule)."""
        self.eng.set_schedule(_lib.SCHED_STREAM)
        self.eng.jobs_run(job_ids)
        self.eng.synchronize()

    def run_plain(self, job_ids, pass_frames=0):
        """One job at a time in frame-synchronous passes (PASS schedule);
        pass_frames=1 renders one frame per pass like the reference's loop."""
        self.eng.set_schedule(_lib.SCHED_PASS, pass_frames)
        for j in job_ids:
            self.eng.jobs_run([j])
        self.eng.synchronize()
        self.eng.set_schedule(_lib.SCHED_STREAM)

    def autotune(self, make_jobs, candidates=None, chunks=24, warm_chunks=6):
        """Times candidate launch configurations on a fresh copy of the
        workload (make_jobs() -> job ids; destroyed afterwards): Mpaths/s over
        `chunks` chunks after `warm_chunks`. Keeps the fastest; returns
        (best, table)."""
        if candidates is None:
            candidates = [dict(max_paths=mp, trace_blocks=tb, chunk_iters=ci, ray_order=ro)
                          for mp in (8 << 20, 16 << 20, 32 << 20) for tb in (1024, 2048) for ci in (2, 4)
                          for ro in (_lib.ORDER_SHADE, _lib.ORDER_BUCKETS, _lib.ORDER_NONE)]
        table = []
        self.eng.set_schedule(_lib.SCHED_STREAM)
        for cand in candidates:
            self.eng.set_launch_config(cand["max_paths"], cand["trace_blocks"], 0)
            self.eng.set_schedule(_lib.SCHED_STREAM, 0, cand["chunk_iters"])
            self.eng.set_ray_order(cand.get("ray_order", _lib.ORDER_SHADE))
            ids = make_jobs()
            for _ in range(warm_chunks):
                self.eng.jobs_step(ids)
            self.eng.synchronize()
            p0 = self.eng.counters()["paths"]
            t0 = time.perf_counter()
            for _ in range(chunks):
                if self.eng.jobs_step(ids) == 0:
                    break
            self.eng.synchronize()
            dt = time.perf_counter() - t0
            rate = (self.eng.counters()["paths"] - p0) / dt / 1e6 if dt > 0 else 0.0
            table.append((rate, cand))
            for j in ids:
                self.eng.job_destroy(j)
        best = max(table, key=lambda x: x[0])[1]
        self.eng.set_launch_config(best["max_paths"], best["trace_blocks"], 0)
        self.eng.set_schedule(_lib.SCHED_STREAM, 0, best["chunk_iters"])
        self.eng.set_ray_order(best.get("ray_order", _lib.ORDER_SHADE))
        return best, table


class Server:
    """Continuous-batching request loop over resident scenes."""

    def __init__(self, batch):
        self.b = batch

    def run(self, specs, clock=time.perf_counter, on_progress=None, keep=False, on_finish=None):
        """specs: JobSpec list with .arrival in seconds from start. Returns per-job
        records {arrival, admitted, first_frame, done, frames, paths}.
        on_progress(job_id, frames_done, frames_total) after every step;
        on_finish(request_index, job_id) when a request's last frame is
        folded, before its job is destroyed (read the result there)."""
        eng = self.b.eng
        eng.set_schedule(_lib.SCHED_STREAM)
        pending = sorted(range(len(specs)), key=lambda i: specs[i].arrival)
        live = {}  # job id -> spec index
        rec = [dict(arrival=specs[i].arrival, admitted=None, first_frame=None, done=None, frames=specs[i].frames,
                    paths=specs[i].paths) for i in range(len(specs))]
        t0 = clock()
        k = 0
        while k < len(pending) or live:
            now = clock() - t0
            # admit arrived requests while memory allows
            while k < len(pending) and specs[pending[k]].arrival <= now:
                i = pending[k]
                try:
                    jid = self.b.create_job(specs[i])
                except _lib.DsptError as e:
                    if e.code == ERR_MEMORY and live:
                        break  # wait for a running job to finish
                    raise
                live[jid] = i
                rec[i]["admitted"] = now
                rec[i]["job"] = jid
                k += 1
            if not live:
                # idle until the next arrival
                nxt = specs[pending[k]].arrival
                while clock() - t0 < nxt:
                    time.sleep(min(0.002, max(0.0, nxt - (clock() - t0))))
                continue
            ids = list(live)
            eng.jobs_step(ids)
            now = clock() - t0
            for jid in ids:
                i = live[jid]
                d, tot = eng.job_status(jid)
                if d >= 1 and rec[i]["first_frame"] is None:
                    rec[i]["first_frame"] = now
                if on_progress:
                    on_progress(jid, d, tot)
                if d >= tot:
                    rec[i]["done"] = now
                    if on_finish:
                        on_finish(i, jid)
                    if not keep:
                        eng.job_destroy(jid)
                    del live[jid]
        eng.synchronize()
        return rec


def latency_stats(rec):
    lat = np.array([r["done"] - r["arrival"] for r in rec if r["done"] is not None])
    ttf = np.array([r["first_frame"] - r["arrival"] for r in rec if r["first_frame"] is not None])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "first_frame_p50_ms": float(np.percentile(ttf, 50) * 1e3),
            "first_frame_p99_ms": float(np.percentile(ttf, 99) * 1e3), "jobs": int(len(lat))}
