"""The serving and large-instanced workloads at their full sizes.

Serving (Poisson arrivals over the 7,335,938-triangle full-material scene with
its resolution / spp mix): requests that arrive while others run are admitted
as memory frees up (device budget plus free HBM), overlap on the GPU, and each
job's accumulation equals the same job rendered alone, bit for bit.

Large instanced scene (4096 instances of 16 meshes, 8192x4096 IBL, 3840x2160,
depth 16): the IBL's importance tables are bit-identical to the oracle's
build (renderer.ts:777-838: row sums and the marginal total in double, every
prefix value rounded to f32 as the reference rounds it); primary rays of two
windows of the 3840x2160 frame hit the same (instance, triangle) as the
oracle on the flattened scene, at the instancing bar of
test_gpu_instancing.py (the flattened scene is baked from the instances whose
world boxes those rays cross, so it answers exactly as the whole flattening
would); a full frame renders every path and is reproducible bit for bit.
"""
import math

import numpy as np
import pytest

import oracle_ref as O

pytestmark = pytest.mark.gpu


def test_serving_poisson_stream_full_scene():
    from dspt import Engine, PROFILE_FULL, POLICY_RR, build_bvh, scenes
    from dspt.batch import BatchRenderer, JobSpec, Server
    wl = scenes.full_materials()
    s = wl.scene
    assert s.num_triangles == 7335938
    s.bvhNodes, s.bvhTriangleIndices = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    eng = Engine(PROFILE_FULL)
    eng.set_policy(POLICY_RR)
    b = BatchRenderer(eng)
    b.add_scene(0, s)
    b.add_ibl(0, wl.ibl)
    resolutions = [(256, 256), (512, 512), (1024, 576), (1920, 1080)]
    spps = [4, 8, 16, 32]
    rng = np.random.default_rng(17)
    reqs, t = [], 0.0
    while len(reqs) < 40:
        t += rng.exponential(1.0 / 150.0)
        w, h = resolutions[int(rng.integers(0, 4))]
        reqs.append(JobSpec(0, 0, wl.camera, w, h, spps[int(rng.integers(0, 4))], max_bounces=wl.max_bounces,
                            arrival=t))
    # the path pool exists before the budget is set (it counts as engine memory)
    w0 = b.create_job(JobSpec(0, 0, wl.camera, 64, 64, 1, max_bounces=wl.max_bounces))
    eng.jobs_run([w0])
    eng.job_destroy(w0)
    # a budget below the stream's total, so admission makes requests wait
    eng.set_memory_budget(eng.memory_usage() + (3 << 30))
    print(f"serving: {len(reqs)} requests over {t:.3f}s", flush=True)
    mixed = {}
    rec = Server(b).run(reqs, on_finish=lambda i, jid: mixed.__setitem__(i, eng.job_read_accum(jid)))
    eng.set_memory_budget(0)
    assert all(r["done"] is not None for r in rec)
    # overlap: some request was admitted while another was still running
    spans = sorted((r["admitted"], r["done"]) for r in rec)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    # admission made some request wait after its arrival
    assert any(r["admitted"] > r["arrival"] + 1e-3 for r in rec)
    assert sorted(mixed) == list(range(len(reqs)))
    for k, spec in enumerate(reqs):
        acc = mixed[k]
        j = b.create_job(spec)
        eng.jobs_run([j])
        alone = eng.job_read_accum(j)
        eng.job_destroy(j)
        np.testing.assert_array_equal(acc.view(np.uint32), alone.view(np.uint32), err_msg=f"request {k}")
    eng.close()


def _window_rays(cam, W, H, x0, y0, w, h):
    """Camera rays through pixel centres of a window (renderer.frag:129-142 without the jitter)."""
    ys, xs = np.mgrid[y0:y0 + h, x0:x0 + w]
    vx = 2.0 * (xs + 0.5) / W - 1.0
    vy = 2.0 * (ys + 0.5) / H - 1.0
    film = math.tan(cam.fov * 0.5 * math.pi / 180.0) * cam.near
    d = np.stack([vx * (W / H) * film, vy * film, np.full_like(vx, -cam.near, dtype=np.float64)], -1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    M = np.asarray(cam.matrixWorld, np.float64).reshape(4, 4).T  # column-major elements
    d = d @ M[:3, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3] = np.asarray(cam.position, np.float32)
    rays[:, 3] = 1e5
    rays[:, 4:7] = d
    return rays


def _boxes_hit(rays, boxes):
    """Instances whose world box (padded) any ray crosses."""
    lo, hi = boxes[None, :, 0] - 1e-3, boxes[None, :, 1] + 1e-3
    crossed = np.zeros(len(boxes), bool)
    for k in range(0, len(rays), 256):
        r = rays[k:k + 256]
        o = r[:, None, 0:3].astype(np.float64)
        inv = 1.0 / np.where(r[:, 4:7] == 0, 1e-30, r[:, 4:7]).astype(np.float64)[:, None, :]
        t0, t1 = (lo - o) * inv, (hi - o) * inv
        tn = np.minimum(t0, t1).max(-1)
        tf = np.maximum(t0, t1).min(-1)
        crossed |= ((tn <= tf) & (tf >= 0)).any(0)
    return np.flatnonzero(crossed)


def test_instanced_full_size():
    from dspt import Engine, PROFILE_FULL, build_bvh, pack_frame_block, scenes
    wl = scenes.instanced_scene()
    s = wl.scene
    img = wl.ibl["image"]
    assert s.num_instances == 4096 and len(s.mesh_count) == 16 and (img["width"], img["height"]) == (8192, 4096)
    assert (wl.width, wl.height, wl.max_bounces) == (3840, 2160, 16)
    eng = Engine(PROFILE_FULL)
    eng.upload_scene_instanced(*s.engine_args())
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    # 1. the 8192x4096 importance tables
    print("config 5: IBL tables", flush=True)
    pdf, cdf, ypdf, ycdf, otot = O.ibl_cdf(img["data"], img["width"], img["height"])
    gp, gc, gyp, gyc = eng.read_ibl_tables()
    for a, b_ in ((gp, pdf), (gc, cdf), (gyp, ypdf), (gyc, ycdf)):
        np.testing.assert_array_equal(a.view(np.uint32), b_.view(np.uint32))
    assert tot == otot
    del pdf, cdf, gp, gc
    # 2. primary rays of two windows against the flattened scene
    W, H = wl.width, wl.height
    rays = np.concatenate([_window_rays(wl.camera, W, H, 1888, 1064, 64, 32),
                           _window_rays(wl.camera, W, H, 900, 400, 64, 32)])
    inst, tri, t, u, v = eng.trace_closest_instanced(rays)
    cand = _boxes_hit(rays, s.world_boxes())
    print(f"config 5: {len(rays)} window rays cross {len(cand)} of 4096 instance boxes", flush=True)
    flat, inst_of, mtri_of = s.flatten(instances=cand)
    nodes, bidx = build_bvh(flat.bvhPositionBuffer, flat.bvhIndexBuffer)
    osc = O.OracleScene(flat.vertexBuffer, flat.triangleIndexBuffer, nodes, bidx, flat.getFlatMaterialBuffer())
    ftri, ft, fu, fv, _ = osc.trace(rays)
    hit = ftri >= 0
    assert hit.mean() > 0.2
    exp_inst = np.where(hit, inst_of[np.maximum(ftri, 0)], -1)
    exp_tri = np.where(hit, mtri_of[np.maximum(ftri, 0)], -1)
    same = (inst == exp_inst) & (tri == exp_tri)
    assert same.mean() >= 0.999, f"{(~same).sum()} of {len(same)} rays differ"
    m = same & hit
    np.testing.assert_allclose(t[m], ft[m], rtol=1e-4, atol=5e-5)
    # 3. a full frame: every path, reproducible
    blk = pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                           ibl_total_sum=tot, scene_counts=(s.scene.num_triangles, 0, 0, 0))
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.reset_counters()
    eng.render(1, 1)
    a1 = eng.read_accum()
    c = eng.counters()
    assert c["paths"] == W * H and c["rays_primary"] == W * H
    eng.render(1, 1)
    np.testing.assert_array_equal(eng.read_accum().view(np.uint32), a1.view(np.uint32))
    assert np.isfinite(a1).all() and a1[..., :3].mean() > 0
    eng.close()
