"""The path pool's scheduling never changes a job's result.

Each work item (job, frame, pixel) is computed from its own seed, and frames
are folded into the running mean in frame order, so a job's samples and
accumulation must be bit-identical whether it runs alone, shares the pool
with other jobs, runs in frame-synchronous passes (one frame per pass, the
reference's loop), goes through the single-image API, or wraps a small frame
ring in a small pool. The single-image path is itself checked against the
oracle in test_gpu_parity.py. Also: memory admission, destroying a job with
work in flight, and the continuous-batching server loop.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix():
    from dspt import scenes
    return scenes.batch_mix(seed=11, n_jobs=3, tri_range=(2_000, 20_000), side_range=(32, 64), spp_range=(3, 6),
                            ibl_size=(128, 64), max_bounces=6)


@pytest.fixture(scope="module")
def mix():
    from dspt import build_bvh
    wls = _mix()
    for w in wls:
        w.scene.bvhNodes, w.scene.bvhTriangleIndices = build_bvh(w.scene.bvhPositionBuffer, w.scene.bvhIndexBuffer)
    return wls


def _batch(eng, wls):
    from dspt.batch import BatchRenderer, JobSpec
    b = BatchRenderer(eng)
    specs = []
    for k, w in enumerate(wls):
        b.add_scene(k, w.scene)
        b.add_ibl(k, w.ibl)
        specs.append(JobSpec(k, k, w.camera, w.width, w.height, w.spp, max_bounces=w.max_bounces, record=True))
    return b, specs


def _result(eng, jid):
    acc = eng.job_read_accum(jid)
    s, sig, prim = eng.job_read_samples(jid)
    return acc, s, sig, prim


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_job_result_independent_of_schedule(mix):
    from dspt import Engine, PROFILE_FULL, SCHED_PASS, SCHED_STREAM
    eng = Engine(PROFILE_FULL)
    b, specs = _batch(eng, mix)
    alone = b.create_job(specs[0])
    eng.jobs_run([alone])
    ref = _result(eng, alone)
    assert eng.job_status(alone) == (specs[0].frames, specs[0].frames)
    assert np.isfinite(ref[0]).all() and ref[0][..., 3].min() > 0

    ids = [b.create_job(s) for s in specs]  # job 0 again, together with the others
    eng.jobs_run(ids)
    _same(_result(eng, ids[0]), ref)

    eng.set_schedule(SCHED_PASS, 1)  # one frame per pass
    j = b.create_job(specs[0])
    eng.jobs_run([j])
    _same(_result(eng, j), ref)
    eng.set_schedule(SCHED_STREAM)

    # single-image API on the same scene: frames 1..n of the implicit job
    w = mix[0]
    eng2 = Engine(PROFILE_FULL)
    eng2.upload_scene(w.scene.vertexBuffer, w.scene.triangleIndexBuffer, w.scene.bvhNodes, w.scene.bvhTriangleIndices,
                      w.scene.getFlatMaterialBuffer())
    img = w.ibl["image"]
    eng2.upload_ibl(img["data"], img["width"], img["height"])
    eng2.resize(w.width, w.height)
    eng2.set_frame(b.block(specs[0]))
    eng2.set_record_samples(True)
    eng2.render(1, specs[0].frames)
    eng2.synchronize()
    s2, sig2, prim2 = eng2.read_samples(specs[0].frames)
    np.testing.assert_array_equal(eng2.read_accum(), ref[0])
    np.testing.assert_array_equal(s2, ref[1])
    np.testing.assert_array_equal(sig2, ref[2])


def test_small_pool_and_ring_wrap(mix):
    """A 65536-path pool and a 1-pool ring force slot reuse within frames and
    ring-slot reuse across frames."""
    from dspt import Engine, PROFILE_FULL, SCHED_STREAM
    from dspt.batch import JobSpec
    w = mix[1]
    eng = Engine(PROFILE_FULL)
    b, _ = _batch(eng, [w])
    spec = JobSpec(0, 0, w.camera, 192, 128, 9, max_bounces=w.max_bounces, record=True)
    ref_job = b.create_job(spec)
    eng.jobs_run([ref_job])
    ref = _result(eng, ref_job)

    eng.set_launch_config(65536, 256, 256)
    eng.set_schedule(SCHED_STREAM, 0, 3, 1.0)
    j = b.create_job(spec)  # ring of ceil(65536 / 24576) + 2 = 5 frames < 9
    eng.jobs_run([j])
    _same(_result(eng, j), ref)


def test_destroy_with_work_in_flight(mix):
    from dspt import Engine, PROFILE_FULL
    eng = Engine(PROFILE_FULL)
    b, specs = _batch(eng, mix)
    ref_job = b.create_job(specs[2])
    eng.jobs_run([ref_job])
    ref = _result(eng, ref_job)
    a = b.create_job(specs[1])
    c = b.create_job(specs[2])
    for _ in range(3):
        eng.jobs_step([a, c])
    eng.job_destroy(a)
    eng.jobs_run([c])
    _same(_result(eng, c), ref)


def test_admission_by_memory_budget(mix):
    from dspt import DsptError, Engine, PROFILE_FULL
    from dspt.batch import JobSpec
    eng = Engine(PROFILE_FULL)
    b, specs = _batch(eng, mix[:1])
    small = b.create_job(specs[0])
    eng.set_memory_budget(eng.memory_usage() + (1 << 20))
    big = JobSpec(0, 0, mix[0].camera, 1024, 1024, 64, max_bounces=4)
    with pytest.raises(DsptError) as e:
        b.create_job(big)
    assert e.value.code == -4
    eng.job_destroy(small)
    eng.set_memory_budget(0)
    j = b.create_job(big)  # unlimited again
    eng.job_destroy(j)


def test_server_loop_completes_every_request(mix):
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import JobSpec, Server, latency_stats
    eng = Engine(PROFILE_FULL)
    b, specs = _batch(eng, mix)
    reqs = []
    rng = np.random.default_rng(5)
    t = 0.0
    for k in range(12):
        t += rng.exponential(0.004)
        s = specs[k % len(specs)]
        reqs.append(JobSpec(s.scene, s.ibl, s.camera, s.width, s.height, s.frames, max_bounces=s.max_bounces, arrival=t))
    rec = Server(b).run(reqs)
    assert all(r["done"] is not None and r["done"] >= r["arrival"] for r in rec)
    st = latency_stats(rec)
    assert st["jobs"] == len(reqs) and st["p99_ms"] >= st["p50_ms"] > 0


def test_bounce_graph_replays_equal_eager_launches(mix):
    """The bounce loop captured as a HIP graph, replayed chunk after chunk
    (both state-buffer parities with an odd chunk length, one- and two-level
    kernels, recaptured after a pool reallocation), gives the results of
    eager launches bit for bit."""
    from dspt import Engine, PROFILE_FULL, SCHED_STREAM, scenes
    from dspt.batch import JobSpec
    inst = scenes.instanced_scene(seed=9, n_meshes=3, n_instances=40, tri_range=(300, 1500), extent=5.0, height=2.0,
                                  width=40, height_px=32, spp=5, max_bounces=6, ibl_size=(128, 64))
    results = {}
    for graphs in (False, True):
        eng = Engine(PROFILE_FULL)
        eng.set_bounce_graphs(graphs)
        b, specs = _batch(eng, mix)
        b.add_scene("inst", inst.scene)
        b.add_ibl("inst", inst.ibl)
        specs.append(JobSpec("inst", "inst", inst.camera, inst.width, inst.height, inst.spp,
                             max_bounces=inst.max_bounces, record=True))
        out = []
        ids = [b.create_job(s) for s in specs[:-1]]  # one-level kernels only
        eng.jobs_run(ids)
        out += [_result(eng, j) for j in ids]
        eng.set_schedule(SCHED_STREAM, chunk_iters=3)  # chunks start on either state buffer
        ids = [b.create_job(s) for s in specs]  # with the two-level job
        eng.jobs_run(ids)
        out += [_result(eng, j) for j in ids]
        eng.set_launch_config(max_paths=65536)  # a new pool: graphs captured again
        ids = [b.create_job(s) for s in specs]
        eng.jobs_run(ids)
        out += [_result(eng, j) for j in ids]
        results[graphs] = out
        eng.close()
    assert len(results[True]) == len(results[False])
    for a, g in zip(results[False], results[True]):
        _same(a, g)


def test_ray_order_keeps_results(mix):
    """Bucketing the extension queue by (octant, origin cell) changes which
    lane walks which ray, not any path's result: accumulations, samples and
    the work counters equal those of the compaction order, with and without
    the captured graphs, one- and two-level."""
    from dspt import Engine, ORDER_BUCKETS, ORDER_NONE, ORDER_SHADE, PROFILE_FULL, SCHED_STREAM, scenes
    from dspt.batch import JobSpec
    inst = scenes.instanced_scene(seed=9, n_meshes=3, n_instances=40, tri_range=(300, 1500), extent=5.0, height=2.0,
                                  width=40, height_px=32, spp=5, max_bounces=6, ibl_size=(128, 64))
    runs = {}
    for order, graphs in ((ORDER_NONE, True), (ORDER_BUCKETS, True), (ORDER_BUCKETS, False), (ORDER_SHADE, True)):
        eng = Engine(PROFILE_FULL)
        eng.set_ray_order(order)
        eng.set_bounce_graphs(graphs)
        eng.set_launch_config(max_paths=65536)  # several refills per frame: sorted and unsorted entries mix
        eng.set_schedule(SCHED_STREAM, chunk_iters=3)
        b, specs = _batch(eng, mix)
        b.add_scene("inst", inst.scene)
        b.add_ibl("inst", inst.ibl)
        specs.append(JobSpec("inst", "inst", inst.camera, inst.width, inst.height, inst.spp,
                             max_bounces=inst.max_bounces, record=True))
        eng.reset_counters()
        ids = [b.create_job(s) for s in specs]
        eng.jobs_run(ids)
        c = eng.counters()
        runs[(order, graphs)] = ([_result(eng, j) for j in ids], {k: c[k] for k in ("paths", "rays_extension",
                                                                                   "rays_shadow", "node_visits",
                                                                                   "tri_tests", "shade_hits")})
        eng.close()
    ref, ref_cnt = runs[(ORDER_NONE, True)]
    for key, (res, cnt) in runs.items():
        assert cnt == ref_cnt, key
        for a, g in zip(ref, res):
            _same(a, g)


def test_large_pool_keeps_results(mix):
    """A 67 M-path job (1024x1024, 64 frames) fills a 64 Mi-path pool, so
    positions up to the top of the pool are used; its accumulation and work
    counters equal those of the same job in a 16 Mi-path pool."""
    from dspt import Engine, PROFILE_FULL
    from dspt.batch import JobSpec
    w = mix[0]
    runs = []
    for mp in (64 << 20, 16 << 20):
        eng = Engine(PROFILE_FULL)
        eng.set_launch_config(max_paths=mp)
        b, _ = _batch(eng, [w])
        eng.reset_counters()
        j = b.create_job(JobSpec(0, 0, w.camera, 1024, 1024, 64, max_bounces=w.max_bounces))
        eng.jobs_run([j])
        c = eng.counters()
        runs.append((eng.job_read_accum(j), {k: c[k] for k in ("paths", "rays_extension", "rays_shadow",
                                                               "node_visits", "tri_tests", "shade_hits")}))
        assert eng.launch_config()["max_paths"] == mp
        eng.close()
    assert runs[0][1]["paths"] == 1024 * 1024 * 64
    assert runs[0][1] == runs[1][1]
    np.testing.assert_array_equal(runs[0][0], runs[1][0])


def test_render_calls_resize_ring_and_continue(mix):
    """The single-image API: a short call, then a longer restart (the ring
    grows to keep more frames in flight), then a continuation; each matches
    the same frames rendered in one call, bit for bit."""
    from dspt import Engine, PROFILE_FULL, pack_frame_block
    w = mix[0]
    s = w.scene
    img = w.ibl["image"]

    def engine():
        eng = Engine(PROFILE_FULL)
        eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices,
                         s.getFlatMaterialBuffer())
        tot = eng.upload_ibl(img["data"], img["width"], img["height"])
        eng.resize(w.width, w.height)
        eng.set_frame(pack_frame_block(w.camera, w.width, w.height, max_bounces=w.max_bounces,
                                       ibl_size=(img["width"], img["height"]), ibl_total_sum=tot))
        eng.set_launch_config(max_paths=65536)  # a small pool: the ring size matters
        return eng

    one = engine()
    one.render(1, 9)
    ref = one.read_accum()
    one.render(1, 5)
    ref5 = one.read_accum()
    one.close()
    eng = engine()
    eng.render(1, 2)  # sizes the implicit job's ring for 2 frames
    eng.render(1, 5)  # restart: a larger ring
    np.testing.assert_array_equal(eng.read_accum(), ref5)
    eng.render(6, 4)  # continue with frames 6..9 on the running mean
    np.testing.assert_array_equal(eng.read_accum(), ref)
    eng.close()


def test_launch_switch_errors():
    """The launch switches check their arguments."""
    from dspt import DsptError, Engine, ORDER_SHADE
    eng = Engine(0)
    with pytest.raises(DsptError):
        eng.set_ray_order(7)
    eng.set_ray_order(ORDER_SHADE)
    eng.set_bounce_graphs(False)
    eng.set_bounce_graphs(True)
    with pytest.raises(DsptError):
        eng.set_launch_config(max_paths=(1 << 30) + 1)
    eng.close()
