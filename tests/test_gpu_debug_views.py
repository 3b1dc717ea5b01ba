"""The debug program (SH/integrator/debug.glsl:1-83) on the GPU against the
oracle, whose debug views are pinned in test_debug_views.py: every mode of
debugModes (renderer.ts:181) and an unknown name (u_debug_mode = -1), with
the background from the IBL and without, sample by sample (branches and
values, the bar of test_gpu_parity.py) and the accumulation bit for bit; and
the Python drop-in's debugMode property renders the same frames."""
import numpy as np
import pytest

from test_debug_views import MODES, H, W, debug_block, debug_workload, oracle_scene
from test_gpu_parity import _check_samples, _upload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    return debug_workload()


@pytest.fixture(scope="module")
def pair(wl):
    from dspt import Engine
    eng = Engine(0)
    total = _upload(eng, wl)
    osc = oracle_scene(wl)
    assert total == osc.total
    eng.resize(W, H)
    yield eng, osc
    eng.close()


@pytest.mark.parametrize("mode,bg", [(m, True) for m in MODES[1:]] + [("Wireframe", True), ("None", True),
                                                                       ("Albedo", False), ("IBL PDF", False),
                                                                       ("Normals", False)])
def test_debug_view_matches_oracle(wl, pair, mode, bg):
    eng, osc = pair
    frames = 2
    blk = debug_block(wl, osc.total, mode, show_background=bg)
    eng.set_frame(blk)
    eng.set_record_samples(True)
    eng.render(1, frames)
    acc = eng.read_accum()
    g_s, g_sig, g_prim = eng.read_samples(frames)
    o = osc.render(blk, W, H, 1, frames, threads=8)
    _check_samples(g_s, g_sig, g_prim, o)
    np.testing.assert_array_equal(acc.view(np.uint32), o["accum"].view(np.uint32))
    hit = (g_prim >= 0).all(axis=0)  # hit in both frames
    assert 0.2 < hit.mean() < 0.9
    if mode not in ("None", "IBL PDF", "IBL CDF"):
        assert (acc[..., 3][hit] == 1).all()  # one hit, alpha 1: no bounce reached the sample


def test_python_renderer_debug_mode(wl):
    """PathtracingRenderer.debugMode packs the mode's index (renderer.ts:406)
    and renders the debug program; 'None' renders the path integrator again."""
    from dspt import Engine, pack_frame_block
    from dspt.renderer import PathtracingRenderer
    N = 4
    r = PathtracingRenderer("webgl-lean", width=W, height=H)
    r.maxBounces = wl.max_bounces
    r.setScene(wl.scene)
    r.setIBL(wl.ibl)
    out = {}
    for mode in ("Tangents", "Wireframe", "None"):
        r.debugMode = mode
        r.render(wl.camera, N)
        out[mode] = r.readAccumulation().copy()
    eng = Engine(0)
    total = _upload(eng, wl)
    eng.resize(W, H)
    s, img = wl.scene, wl.ibl["image"]
    for mode in ("Tangents", "Wireframe", "None"):
        eng.set_frame(pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces,
                                       ibl_size=(img["width"], img["height"]), ibl_total_sum=total,
                                       scene_counts=(s.num_triangles, len(s.bvhNodes) // 16,
                                                     len(s.bvhTriangleIndices), 0),
                                       debug_mode=MODES.index(mode) if mode in MODES else -1))
        eng.render(1, N - 1)
        np.testing.assert_array_equal(out[mode].view(np.uint32), eng.read_accum().view(np.uint32), err_msg=mode)
    eng.close()
    assert not np.array_equal(out["Tangents"], out["None"])


def test_debug_and_path_jobs_share_the_pool(wl):
    """A debug job and path-tracing jobs in one batched pool: the shading
    kernel's debug instance runs them together, and every job's
    accumulation, samples, signatures and primary hits equal the same job run
    alone (the path job's alone runs the plain instance)."""
    from dspt import Engine
    from dspt.batch import BatchRenderer, JobSpec
    eng = Engine(0)
    b = BatchRenderer(eng)
    b.add_scene(0, wl.scene)
    b.add_ibl(0, wl.ibl)
    specs = [JobSpec(0, 0, wl.camera, W, H, 3, max_bounces=wl.max_bounces, record=True, debug_mode=4),
             JobSpec(0, 0, wl.camera, W, H, 3, max_bounces=wl.max_bounces, record=True),
             JobSpec(0, 0, wl.camera, 40, 24, 2, max_bounces=wl.max_bounces, record=True, debug_mode=10,
                     ibl_rotation_rad=1.3)]
    alone = []
    for s in specs:
        j = b.create_job(s)
        eng.jobs_run([j])
        alone.append(j)
    mixed = [b.create_job(s) for s in specs]
    b.run_batched(mixed)
    for a, m in zip(alone, mixed):
        for x, y in zip((eng.job_read_accum(a), *eng.job_read_samples(a)),
                        (eng.job_read_accum(m), *eng.job_read_samples(m))):
            np.testing.assert_array_equal(x, y)
    eng.close()
