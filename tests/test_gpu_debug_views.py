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
