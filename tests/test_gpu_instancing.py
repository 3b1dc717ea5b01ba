"""Two-level (instanced) scenes against their flattened form.

The reference has no instancing: its loader bakes every node's world
matrix into the vertex stream (scene_loader.ts:173-222, 244-345), so the
reference's result for an instanced scene is the render of the flattened
scene, and the flattened scene runs through the single-level path that
test_gpu_parity.py pins to the oracle. The engine's two-level walk moves the
ray into object space instead of moving the triangles into world space, so
the two agree to rounding, not bit for bit. Bars:
  * traversal: hit/miss and (instance, mesh triangle) equal for >= 99.9 % of
    rays, t within 1e-4 relative or 5e-5 absolute where they agree;
  * per-sample radiance: within 1e-3 relative (or 1e-5 absolute) for
    >= 97 % of samples (paths whose branches flipped on a rounding
    difference differ by their whole value and are the rest);
  * a single-level job is bit-identical whether or not an instanced job
    shares the pool (the two-level kernels walk a one-level scene exactly as
    the one-level kernels do).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small(seed=7):
    from dspt import scenes
    return scenes.instanced_scene(seed=seed, n_meshes=4, n_instances=96, tri_range=(300, 3000), extent=6.0, height=3.0,
                                  width=48, height_px=48, spp=6, max_bounces=8, ibl_size=(128, 64))


@pytest.fixture(scope="module")
def inst_wl():
    from dspt import build_bvh
    wl = _small()
    flat, inst_of, mtri_of = wl.scene.flatten()
    flat.bvhNodes, flat.bvhTriangleIndices = build_bvh(flat.bvhPositionBuffer, flat.bvhIndexBuffer)
    return wl, flat, inst_of, mtri_of


def _rays(n, seed, cam):
    rng = np.random.default_rng(seed)
    half = n // 2
    org = np.empty((n, 3), np.float32)
    org[:half] = np.asarray(cam, np.float32)
    org[half:] = rng.uniform([-7, -1, -7], [7, 4, 7], (n - half, 3))
    tgt = rng.uniform([-6, 0, -6], [6, 3, 6], (n, 3)).astype(np.float32)
    d = tgt - org
    d[half:] = rng.normal(size=(n - half, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = 1e5
    rays[:, 4:7] = d
    return rays


def test_instanced_trace_matches_flattened(inst_wl):
    from dspt import Engine
    wl, flat, inst_of, mtri_of = inst_wl
    s = wl.scene
    e1 = Engine(0)
    e1.upload_scene_instanced(*s.engine_args())
    e2 = Engine(0)
    e2.upload_scene(flat.vertexBuffer, flat.triangleIndexBuffer, flat.bvhNodes, flat.bvhTriangleIndices,
                    flat.getFlatMaterialBuffer())
    rays = _rays(200_000, 3, wl.camera.position)
    inst, tri, t, u, v = e1.trace_closest_instanced(rays)
    ftri, ft, fu, fv = e2.trace_closest(rays)
    hit = ftri >= 0
    assert hit.mean() > 0.3 and (~hit).mean() > 0.05  # both cases exercised
    exp_inst = np.where(hit, inst_of[np.maximum(ftri, 0)], -1)
    exp_tri = np.where(hit, mtri_of[np.maximum(ftri, 0)], -1)
    same = (inst == exp_inst) & (tri == exp_tri)
    assert same.mean() >= 0.999, f"{(~same).sum()} of {len(same)} rays differ"
    m = same & hit
    # coordinates reach ~10, where a float32 ulp is ~1e-6; object-space rays
    # and baked vertices each round a few times
    np.testing.assert_allclose(t[m], ft[m], rtol=1e-4, atol=5e-5)
    # barycentrics of grazing hits are ill-conditioned (error ~ 1/|det|): a
    # few in 10^4 may exceed the bar, none by much
    du, dv = np.abs(u[m] - fu[m]), np.abs(v[m] - fv[m])
    ok = (du <= 2e-3 + 1e-3 * np.abs(fu[m])) & (dv <= 2e-3 + 1e-3 * np.abs(fv[m]))
    assert ok.mean() >= 0.9999 and max(du.max(), dv.max()) < 0.05
    assert np.all(inst[~hit & same] == -1)
    # the plain probe reports the mesh triangle too
    tri2, t2, _, _ = e1.trace_closest(rays)
    np.testing.assert_array_equal(tri2, tri)
    np.testing.assert_array_equal(t2, t)
    # occlusion agrees with the closest hit
    occ = e1.trace_occluded(rays)
    np.testing.assert_array_equal(occ, (tri >= 0).astype(np.int32))


def _job(eng, b, key, wl, record=True):
    from dspt.batch import JobSpec
    return b.create_job(JobSpec(key, key, wl.camera, wl.width, wl.height, wl.spp, max_bounces=wl.max_bounces,
                                record=record))


def test_instanced_render_matches_flattened(inst_wl):
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import BatchRenderer
    wl, flat, _, _ = inst_wl
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    b.add_scene("inst", wl.scene)
    b.add_scene("flat", flat)
    b.add_ibl("inst", wl.ibl)
    b.add_ibl("flat", wl.ibl)
    ji = _job(eng, b, "inst", wl)
    jf = _job(eng, b, "flat", wl)
    eng.jobs_run([ji, jf])
    si, _, _ = eng.job_read_samples(ji)
    sf, _, _ = eng.job_read_samples(jf)
    assert np.isfinite(si).all() and np.isfinite(sf).all()
    close = np.all(np.abs(si - sf) <= np.maximum(1e-3 * np.abs(sf), 1e-5), axis=-1)
    assert close.mean() >= 0.97, f"only {close.mean():.4f} of samples agree"
    ai, af = eng.job_read_accum(ji), eng.job_read_accum(jf)
    assert abs(ai[..., :3].mean() - af[..., :3].mean()) <= 0.05 * af[..., :3].mean()


def test_plain_job_unchanged_beside_instanced_job(inst_wl):
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import BatchRenderer
    wl, flat, _, _ = inst_wl
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    b.add_scene("inst", wl.scene)
    b.add_scene("flat", flat)
    b.add_ibl("inst", wl.ibl)
    b.add_ibl("flat", wl.ibl)
    alone = _job(eng, b, "flat", wl)
    eng.jobs_run([alone])  # one-level kernels
    a_acc = eng.job_read_accum(alone)
    a_s = eng.job_read_samples(alone)
    j1 = _job(eng, b, "flat", wl)
    j2 = _job(eng, b, "inst", wl)
    eng.jobs_run([j1, j2])  # two-level kernels for both
    np.testing.assert_array_equal(eng.job_read_accum(j1), a_acc)
    for x, y in zip(eng.job_read_samples(j1), a_s):
        np.testing.assert_array_equal(x, y)
    assert eng.job_status(j2) == (wl.spp, wl.spp)


@pytest.mark.parametrize("mode,index,tol", [("Albedo", 1, 1e-5), ("Normals", 4, 1e-4), ("Bitangents", 6, 1e-3),
                                            ("UV0", 8, 1e-4), ("IBL PDF", 10, 0.0)])
def test_instanced_debug_view_matches_flattened(inst_wl, mode, index, tol):
    """The debug program on a two-level scene reads the hit's attributes
    through the instance (object-space vertices, normals moved to world
    space): it equals the flattened scene's view, which the one-level path
    pins to the oracle, to rounding, wherever both rays hit the same
    triangle (>= 99.9 % of pixels); misses are identical."""
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import BatchRenderer, JobSpec
    wl, flat, _, _ = inst_wl
    eng = Engine(PROFILE_FULL)
    b = BatchRenderer(eng)
    b.add_scene("inst", wl.scene)
    b.add_scene("flat", flat)
    b.add_ibl("inst", wl.ibl)
    b.add_ibl("flat", wl.ibl)
    jobs = [b.create_job(JobSpec(k, k, wl.camera, wl.width, wl.height, 2, max_bounces=wl.max_bounces, record=True,
                                 debug_mode=index)) for k in ("inst", "flat")]
    eng.jobs_run(jobs)
    (si, _, pi), (sf, _, pf) = (eng.job_read_samples(j) for j in jobs)
    eng.close()
    miss = (pi < 0) & (pf < 0)
    hit = (pi >= 0) & (pf >= 0)
    assert miss.mean() > 0.05 and hit.mean() > 0.3
    assert (miss | hit).mean() >= 0.999
    np.testing.assert_array_equal(si[miss], sf[miss])
    close = np.all(np.abs(si[hit] - sf[hit]) <= np.maximum(tol * np.abs(sf[hit]), tol), axis=-1)
    assert close.mean() >= 0.999, f"{mode}: only {close.mean():.5f} of hit samples agree"
