"""Renderer settings that change the integrator, sample by sample against the
oracle on the GPU: a point light (SH/lighting.glsl:2-22, visible and behind
spheres), a rotated IBL (SH/lighting.glsl:56-89, u_ibl_rotation), the IBL off
with a background colour (SH/integrator/misptdl.glsl, pow(bg, 2.2) on
misses) and the clamp off (clampThreshold = 0). Each must agree with the
oracle on every sample (branches and values, the bar of test_gpu_parity.py)
and must change the image against the default settings."""
import numpy as np
import pytest

import oracle_ref as O
from test_gpu_parity import _check_samples, _oracle_scene, _upload

pytestmark = pytest.mark.gpu


def _render(wl, profile, W, H, frames, **kw):
    from dspt import Engine, pack_frame_block
    from dspt.scene import Light  # noqa: F401
    eng = Engine(profile)
    _upload(eng, wl)
    osc = _oracle_scene(wl, profile)
    s = wl.scene
    img = wl.ibl["image"]
    blk = pack_frame_block(wl.camera, W, H, max_bounces=kw.pop("max_bounces", wl.max_bounces),
                           ibl_size=(img["width"], img["height"]), ibl_total_sum=osc.total,
                           scene_counts=(s.num_triangles, len(s.bvhNodes) // 16, len(s.bvhTriangleIndices), 0), **kw)
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.set_record_samples(True)
    eng.render(1, frames)
    acc = eng.read_accum()
    g_s, g_sig, g_prim = eng.read_samples(frames)
    eng.close()
    o = osc.render(blk, W, H, 1, frames, threads=8)
    _check_samples(g_s, g_sig, g_prim, o)
    np.testing.assert_array_equal(acc.view(np.uint32), o["accum"].view(np.uint32))
    return acc, g_s, g_prim


@pytest.fixture(scope="module")
def baseline(sphere_field_small):
    return _render(sphere_field_small, 0, 48, 40, 2)


def test_point_light(sphere_field_small, baseline):
    from dspt.scene import Light
    wl = sphere_field_small
    # above the sphere field: spheres shadow part of the ground, the rest sees it
    light = Light(position=(0.0, 6.0, 1.0), emission=(40.0, 35.0, 30.0))
    acc, s, prim = _render(wl, 0, 48, 40, 2, light=light)
    assert np.abs(acc - baseline[0]).max() > 1e-3
    # the light alone (IBL off, black background): radiance only where it is
    # seen, so some hit pixels stay black (shadowed or facing away) and others are lit
    acc2, s2, prim2 = _render(wl, 0, 48, 40, 2, light=light, use_ibl=False, show_background=False,
                              background_color=(0, 0, 0, 1), max_bounces=1)
    hit = prim2[0] >= 0
    lum = s2[0][..., :3].sum(-1)
    assert (lum[hit] > 0).mean() > 0.1 and (lum[hit] == 0).mean() > 0.01


def test_rotated_ibl(sphere_field_small, baseline):
    acc, _, _ = _render(sphere_field_small, 0, 48, 40, 2, ibl_rotation_rad=0.7)
    assert np.abs(acc - baseline[0]).max() > 1e-3
    acc2, _, _ = _render(sphere_field_small, 0, 48, 40, 2, ibl_rotation_rad=-2.5)
    assert np.abs(acc2 - acc).max() > 1e-3


def test_ibl_off_background_colour(sphere_field_small):
    bg = (0.2, 0.4, 0.6, 1.0)
    # u_background_from_ibl = showBackground: off, so primary misses see the colour
    acc, s, prim = _render(sphere_field_small, 0, 48, 40, 2, use_ibl=False, show_background=False, background_color=bg)
    miss = prim[0] < 0
    assert miss.any()
    want = np.power(np.float32(bg[:3]), np.float32(2.2))
    np.testing.assert_allclose(s[0][miss][:, :3], np.broadcast_to(want, (miss.sum(), 3)), rtol=1e-6)
    # IBL lighting off but the IBL as background
    acc2, s2, prim2 = _render(sphere_field_small, 0, 48, 40, 2, use_ibl=False, show_background=True, background_color=bg)
    assert np.abs(s2[0][miss][:, :3] - want).max() > 1e-3


def test_clamp_off(sphere_field_small):
    acc, s, _ = _render(sphere_field_small, 0, 48, 40, 4, clamp_threshold=0.0)
    # the 5e4 sun disc makes some samples exceed the default clamp of 3
    assert s[..., :3].max() > 3.0


def test_full_profile_settings():
    from dspt import build_bvh, scenes
    from dspt.scene import Light
    wl = scenes.full_materials(grid=48, n_objects=24, subdiv=2, width=40, height=28, spp=2, ibl_size=(256, 128))
    wl.scene.bvhNodes, wl.scene.bvhTriangleIndices = build_bvh(wl.scene.bvhPositionBuffer, wl.scene.bvhIndexBuffer)
    _render(wl, 1, 40, 28, 2, light=Light(position=(0.0, 9.0, 0.0), emission=(60, 60, 60)), ibl_rotation_rad=1.3,
            clamp_threshold=0.0)
