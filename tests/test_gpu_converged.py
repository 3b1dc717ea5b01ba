"""Converged images against an FP64 render, and the full-size frame sample by sample.

The oracle is built twice from the same source: in f32 (the reference's
arithmetic, the parity target) and in f64 (a converged reference with the
same RNG sequence, so Monte Carlo noise cancels and what remains is
arithmetic). For every workload the engine's image must be as close to the
f64 render as the f32 restatement of the reference itself is:

    rmse(engine, f64) <= 1.5 * rmse(oracle f32, f64) + 1e-6     (relative RMSE)

The engine and the f32 oracle share one implementation of the
transcendentals (pin_math.hpp), so their images agree to the last bits; 1.5
is a margin, not a need. The measured values are printed.

Workloads: the 256x256 sphere field (16 spp, depth 8, webgl-lean) whole; the
1920x1080 full-material frame (7,335,938 triangles, 64 spp, depth 12,
webgl-full) on four 64x32 windows; three jobs of the 32-scene batched mix run
together through the job API (windows, full spp); a two-level instanced scene
against the f64 render of its flattening.

Per-sample bar at full size: on four windows of the 1080p frame (frame 1,
twelve bounces over curved specular, transmissive and thin-film lobes), every
sample takes the f32 oracle's branches (primary and secondary hit ids
included in the branch signature) and every sample is within 1e-4 relative.
For scale, the oracle's own f32 and f64 builds agree within 1e-4 on only
30-75 % of a window's samples: last-bit differences grow bounce to bounce.
"""
import numpy as np
import pytest

import oracle_ref as O

pytestmark = pytest.mark.gpu


def _rel_rmse(a, b):
    a = np.asarray(a, np.float64)[..., :3]
    b = np.asarray(b, np.float64)[..., :3]
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _agreement(s, sig, ref):
    """(same-branch fraction, fraction within 1e-4, within 1e-3) of samples s vs ref samples."""
    same = sig == ref["sigs"]
    a = s[same][:, :3].astype(np.float64)
    b = ref["samples"][same][:, :3].astype(np.float64)
    err = (np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max(axis=1)
    return same.mean(), (err <= 1e-4).mean(), (err <= 1e-3).mean()


def _check_rmse(name, gpu, of, od):
    e_gpu, e_ref = _rel_rmse(gpu, od), _rel_rmse(of, od)
    print(f"{name}: rel RMSE engine vs f64 {e_gpu:.3e}, oracle f32 vs f64 {e_ref:.3e}, engine vs f32 {_rel_rmse(gpu, of):.3e}")
    assert e_gpu <= 1.5 * e_ref + 1e-6, (e_gpu, e_ref)


def test_sphere_field_converged():
    from dspt import Engine, PROFILE_LEAN, build_bvh, pack_frame_block, scenes
    wl = scenes.sphere_field()
    s = wl.scene
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    mats = s.getFlatMaterialBuffer()
    eng = Engine(PROFILE_LEAN)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, mats)
    img = wl.ibl["image"]
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    blk = pack_frame_block(wl.camera, wl.width, wl.height, max_bounces=wl.max_bounces,
                           ibl_size=(img["width"], img["height"]), ibl_total_sum=tot)
    eng.resize(wl.width, wl.height)
    eng.set_frame(blk)
    eng.render(1, wl.spp)
    gpu = eng.read_accum()
    eng.close()
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, mats, wl.ibl, PROFILE_LEAN)
    of = osc.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8)["accum"]
    od = osc.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8, use_double=True)["accum"]
    _check_rmse(f"sphere field 256x256 {wl.spp} spp", gpu, of, od)


WINDOWS = ((0, 0), (928, 524), (1856, 1048), (300, 700))


def test_full_materials_1080p():
    from dspt import Engine, PROFILE_FULL, build_bvh, pack_frame_block, scenes
    wl = scenes.full_materials()
    s = wl.scene
    assert s.num_triangles == 7335938 and (wl.width, wl.height, wl.spp, wl.max_bounces) == (1920, 1080, 64, 12)
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    mats = s.getFlatMaterialBuffer()
    eng = Engine(PROFILE_FULL)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, mats)
    img = wl.ibl["image"]
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    W, H = wl.width, wl.height
    blk = pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                           ibl_total_sum=tot)
    eng.resize(W, H)
    eng.set_frame(blk)
    # frame 1 sample by sample, twice (same bits)
    eng.set_record_samples(True)
    eng.reset_counters()
    eng.render(1, 1)
    g_s, g_sig, g_prim = eng.read_samples(1)
    c = eng.counters()
    assert c["paths"] == W * H and c["rays_primary"] == W * H
    eng.render(1, 1)
    np.testing.assert_array_equal(eng.read_samples(1)[0].view(np.uint32), g_s.view(np.uint32))
    # the whole 64-spp frame
    eng.set_record_samples(False)
    eng.render(1, wl.spp)
    acc = eng.read_accum()
    eng.close()

    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, mats, wl.ibl, PROFILE_FULL)
    assert osc.total == tot
    gw, fw, dw = [], [], []
    for x0, y0 in WINDOWS:
        x1, y1 = x0 + 64, y0 + 32
        r = (x0, y0, x1, y1)
        of = osc.render(blk, W, H, 1, 1, rect=r, threads=8)
        od = osc.render(blk, W, H, 1, 1, rect=r, threads=8, use_double=True)
        win = (slice(None), slice(y0, y1), slice(x0, x1))
        np.testing.assert_array_equal(g_prim[win], of["prim"])
        gsame, g4, g3 = _agreement(g_s[win], g_sig[win], of)
        rsame, r4, r3 = _agreement(of["samples"], of["sigs"], od)
        print(f"window {r}: engine vs f32: branches {gsame:.4f}, within 1e-4 {g4:.4f}, 1e-3 {g3:.4f}; "
              f"f32 vs f64: branches {rsame:.4f}, within 1e-4 {r4:.4f}, 1e-3 {r3:.4f}")
        exact = (g_s[win].view(np.uint32) == of["samples"].view(np.uint32)).all(axis=-1).mean()
        print(f"window {r}: bit-identical samples {exact:.6f}")
        assert gsame == 1.0 and g4 == 1.0, (gsame, g4)
        # converged: the 64-spp window
        cf = osc.render(blk, W, H, 1, wl.spp, rect=r, record=False, threads=8)["accum"]
        cd = osc.render(blk, W, H, 1, wl.spp, rect=r, record=False, threads=8, use_double=True)["accum"]
        gw.append(acc[y0:y1, x0:x1]), fw.append(cf[y0:y1, x0:x1]), dw.append(cd[y0:y1, x0:x1])
    _check_rmse("full materials 1080p 64 spp (4 windows)", np.concatenate(gw), np.concatenate(fw), np.concatenate(dw))


def test_batched_mix_jobs_converged():
    """Three jobs of the 32-scene batched mix (the first jobs of its seeded
    draw: spheres 496^2 x 25 spp, terrain 704^2 x 227 spp, terrain with 1.1 M
    triangles 496^2 x 137 spp) rendered together through the job API."""
    from dspt import Engine, PROFILE_FULL, build_bvh
    from dspt.batch import BatchRenderer, JobSpec
    from dspt import scenes
    mix = scenes.batch_mix(seed=3, n_jobs=7)
    picks = [mix[1], mix[3], mix[6]]
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    specs, jobs = [], []
    for k, w in enumerate(picks):
        w.scene.bvhNodes, w.scene.bvhTriangleIndices = build_bvh(w.scene.bvhPositionBuffer, w.scene.bvhIndexBuffer)
        b.add_scene(k, w.scene)
        b.add_ibl(k, w.ibl)
        specs.append(JobSpec(k, k, w.camera, w.width, w.height, w.spp, max_bounces=w.max_bounces))
        jobs.append(b.create_job(specs[-1]))
    eng.jobs_run(jobs)
    accs = [eng.job_read_accum(j) for j in jobs]
    eng.close()
    for k, w in enumerate(picks):
        s = w.scene
        osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices,
                            s.getFlatMaterialBuffer(), w.ibl, PROFILE_FULL)
        blk = b.block(specs[k])
        x0, y0 = w.width // 2 - 32, w.height // 2 - 32
        r = (x0, y0, x0 + 64, y0 + 64)
        cf = osc.render(blk, w.width, w.height, 1, w.spp, rect=r, record=False, threads=8)["accum"]
        cd = osc.render(blk, w.width, w.height, 1, w.spp, rect=r, record=False, threads=8, use_double=True)["accum"]
        win = (slice(y0, y0 + 64), slice(x0, x0 + 64))
        _check_rmse(f"{w.name} {w.width}^2 {w.spp} spp (64x64 window)", accs[k][win], cf[win], cd[win])


def _bake_f32(wl, flat):
    """The flattened vertex stream with positions baked in f32 arithmetic
    (p.x*M0 + p.y*M1 + p.z*M2 + M3, each step rounded) instead of in f64
    rounded once: an equally valid baking, <= 1 ulp away."""
    inst = wl.scene
    src = np.asarray(inst.scene.vertexBuffer, np.float32).reshape(-1, 20)
    rows = []
    for i, m in enumerate(inst.inst_mesh):
        M = inst.o2w[i].astype(np.float32)
        p = src[inst.mesh_vfirst[m]:inst.mesh_vfirst[m] + inst.mesh_vcount[m], 0:3]
        rows.append(p[:, 0:1] * M[:, 0] + p[:, 1:2] * M[:, 1] + p[:, 2:3] * M[:, 2] + M[:, 3])
    vb = np.asarray(flat.vertexBuffer, np.float32).reshape(-1, 20).copy()
    vb[:, 0:3] = np.concatenate(rows).astype(np.float32)
    return vb


def test_instanced_scene_converged():
    """A two-level scene against the f64 render of its flattening (the
    reference's result for it). The engine's object-space walk rounds every
    intersection differently from baked vertices (the ray goes through the
    instance's inverse transform), so the gap allowed is the largest of the
    f32 oracle's over equally valid roundings of the flattened geometry:
    f64 baking rounded once, f32 baking step by step, and the baking with
    every coordinate moved by a random walk of 1, 2 or 4 ulps."""
    from dspt import Engine, PROFILE_FULL, build_bvh, scenes
    from dspt.batch import BatchRenderer, JobSpec
    wl = scenes.instanced_scene(seed=7, n_meshes=4, n_instances=96, tri_range=(300, 3000), extent=6.0, height=3.0,
                                width=64, height_px=48, spp=32, max_bounces=8, ibl_size=(128, 64))
    flat, _, _ = wl.scene.flatten()
    flat.bvhNodes, flat.bvhTriangleIndices = build_bvh(flat.bvhPositionBuffer, flat.bvhIndexBuffer)
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    b.add_scene("inst", wl.scene)
    b.add_ibl("inst", wl.ibl)
    spec = JobSpec("inst", "inst", wl.camera, wl.width, wl.height, wl.spp, max_bounces=wl.max_bounces)
    j = b.create_job(spec)
    eng.jobs_run([j])
    gpu = eng.job_read_accum(j)
    blk = b.block(spec)
    eng.close()
    mats = flat.getFlatMaterialBuffer()
    osc = O.OracleScene(flat.vertexBuffer, flat.triangleIndexBuffer, flat.bvhNodes, flat.bvhTriangleIndices, mats,
                        wl.ibl, PROFILE_FULL)
    of = osc.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8)["accum"]
    od = osc.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8, use_double=True)["accum"]
    vb = _bake_f32(wl, flat)
    pos = np.ascontiguousarray(vb[:, 0:3]).reshape(-1)
    n2, b2 = build_bvh(pos, flat.triangleIndexBuffer)
    alt = O.OracleScene(vb.reshape(-1), flat.triangleIndexBuffer, n2, b2, mats, wl.ibl, PROFILE_FULL)
    oa = alt.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8)["accum"]
    gaps = {"f64 baking": _rel_rmse(of, od), "f32 baking": _rel_rmse(oa, od)}
    rng = np.random.default_rng(5)
    base = np.asarray(flat.vertexBuffer, np.float32).reshape(-1, 20)
    for k in (1, 2, 4):
        vj = base.copy()
        p = vj[:, 0:3].copy()
        for _ in range(k):
            p = np.nextafter(p, rng.choice([-np.inf, np.inf], p.shape).astype(np.float32)).astype(np.float32)
        vj[:, 0:3] = p
        nj, bj = build_bvh(np.ascontiguousarray(p).reshape(-1), flat.triangleIndexBuffer)
        oj = O.OracleScene(vj.reshape(-1), flat.triangleIndexBuffer, nj, bj, mats, wl.ibl, PROFILE_FULL)
        gaps[f"{k}-ulp walk"] = _rel_rmse(oj.render(blk, wl.width, wl.height, 1, wl.spp, record=False, threads=8)["accum"], od)
    e_gpu = _rel_rmse(gpu, od)
    print(f"instanced {wl.width}x{wl.height} {wl.spp} spp: rel RMSE engine vs f64 {e_gpu:.3e}; f32 oracle vs f64 over "
          + ", ".join(f"{k} {v:.3e}" for k, v in gaps.items()))
    assert e_gpu <= 1.5 * max(gaps.values()) + 1e-6, (e_gpu, gaps)
