"""The engine's display chain (k_postprocess, one fused pass) against the
oracle's restatement of renderer.ts:665-700, on an accumulation the engine
rendered.

Bar: the float display buffer and the 8-bit canvas are bit-exact in every
case. f32 add/mul/div are correctly rounded on both sides, and pow (gamma,
Cineon) is the one pinned implementation (pin_math.hpp) that the kernel and
the oracle share.
"""
import numpy as np
import pytest

import oracle_ref as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rendered():
    from dspt import Engine, PROFILE_LEAN, build_bvh, pack_frame_block, scenes
    wl = scenes.sphere_field(n_spheres=16, width=96, height=64, spp=4, ibl_size=(128, 64))
    s = wl.scene
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    eng = Engine(PROFILE_LEAN)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    img = wl.ibl["image"]
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    eng.resize(wl.width, wl.height)
    eng.set_frame(pack_frame_block(wl.camera, wl.width, wl.height, max_bounces=6, ibl_size=(img["width"], img["height"]),
                                   ibl_total_sum=tot, scene_counts=(s.num_triangles, len(nodes) // 16, len(bidx), 0)))
    eng.render(1, wl.spp)
    eng.synchronize()
    return eng, eng.read_accum()


CASES = [(mode, gamma, fxaa, size) for mode in (-1, 0, 1, 2, 3, 4) for gamma in (False, True) for fxaa in (False, True)
         for size in (None, (192, 128), (61, 37))]


@pytest.mark.parametrize("mode,gamma,fxaa,size", CASES)
def test_display_chain_matches_oracle(rendered, mode, gamma, fxaa, size):
    eng, acc = rendered
    assert acc[..., :3].max() > 1.0  # the operators see values above 1
    u8, f = eng.postprocess(mode, 0.8, gamma, fxaa, size, want_float=True)
    rf, ru8 = O.postprocess(acc, mode, exposure=0.8, gamma=gamma, fxaa=fxaa, out_size=size)
    assert f.shape == rf.shape and u8.shape == ru8.shape
    np.testing.assert_array_equal(f.view(np.uint32), rf.view(np.uint32))
    np.testing.assert_array_equal(u8, ru8)


def test_job_postprocess(rendered):
    """A job's display chain reads that job's accumulation."""
    from dspt import scenes
    from dspt.batch import BatchRenderer, JobSpec
    from dspt import Engine, PROFILE_LEAN
    wl = scenes.sphere_field(n_spheres=8, width=40, height=30, spp=3, ibl_size=(64, 32))
    eng = Engine(PROFILE_LEAN)
    b = BatchRenderer(eng)
    b.add_scene(0, wl.scene)
    b.add_ibl(0, wl.ibl)
    j = b.create_job(JobSpec(0, 0, wl.camera, wl.width, wl.height, wl.spp, max_bounces=5))
    eng.jobs_run([j])
    acc = eng.job_read_accum(j)
    u8, f = eng.job_postprocess(j, 1, 1.5, False, True, want_float=True)
    rf, ru8 = O.postprocess(acc, 1, exposure=1.5, gamma=False, fxaa=True)
    np.testing.assert_array_equal(f, rf)
    np.testing.assert_array_equal(u8, ru8)
