"""Material textures (scene_data_webgl2.ts:233-363, material.glsl:5-110).

Oracle pins (CPU): texel centres reproduce the stored bytes / 255 exactly;
half-way points are the two-texel mean; REPEAT makes uv and uv + 1
equivalent; the layer rounds to nearest and clamps; an array index out of
range and untextured slots give vec4(1); uv set 1 reads (0, 0).
GPU: textured scenes render like the oracle (hit ids exact, same-branch
samples within tolerance), and a scene whose materials reference textures it
does not have is refused.
"""
import numpy as np
import pytest

import oracle_ref as O


def _scene_with_textures(arrays, infos):
    from dspt import MaterialData
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    v = np.zeros((3, 20), np.float32)
    v[:, 0:3] = P
    v[:, 4:7] = [0, 0, 1]
    nodes = np.zeros(16, np.float32)
    nodes[3 * 4 + 3] = 0.0  # leaf: cnt in c.w, first in d.w
    nodes[2 * 4 + 3] = 1.0
    sc = O.OracleScene(v.reshape(-1), np.arange(3, dtype=np.uint32), nodes, np.zeros(1, np.int32), MaterialData().data)
    # arrays: [h, w, 4] (one layer) or [layers, h, w, 4]; dims = (w, h, layers)
    dims = np.array([(a.shape[2], a.shape[1], a.shape[0]) if a.ndim == 4 else (a.shape[1], a.shape[0], 1) for a in arrays],
                    np.int32)
    texels = np.concatenate([a.reshape(-1) for a in arrays]).astype(np.uint8)
    sc.set_textures(np.asarray(infos, np.float32).reshape(-1), dims, texels)
    return sc


def _info(array=0, layer=0, offset=(0, 0), scale=(1, 1), uv_set=0):
    return [offset[0], offset[1], array, layer, scale[0], scale[1], uv_set, 0]


def test_texel_centres_and_midpoints():
    img = np.array([[[10, 20, 30, 255], [200, 100, 50, 0]],
                    [[0, 255, 128, 64], [1, 2, 3, 4]]], np.uint8)  # 2 x 2, row 0 first
    sc = _scene_with_textures([img], [_info()])
    centres = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]], np.float32)
    got = sc.tex_value(0, centres)
    want = np.array([img[0, 0], img[0, 1], img[1, 0], img[1, 1]], np.float32) / np.float32(255)
    np.testing.assert_array_equal(got, want)
    mid = sc.tex_value(0, np.array([[0.5, 0.25]], np.float32))[0]
    np.testing.assert_allclose(mid, (want[0] + want[1]) / 2, rtol=0, atol=1e-7)
    quad = sc.tex_value(0, np.array([[0.5, 0.5]], np.float32))[0]
    np.testing.assert_allclose(quad, want.mean(axis=0), rtol=0, atol=1e-7)


def test_repeat_wrap_and_scale_offset():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (8, 16, 4), dtype=np.uint8)
    sc = _scene_with_textures([img], [_info(), _info(offset=(0.5, 0.25), scale=(2, 3))])
    uv = rng.uniform(0, 1, (200, 2)).astype(np.float32)
    a = sc.tex_value(0, uv)
    for shift in ((1, 0), (0, 1), (-2, 3)):
        b = sc.tex_value(0, uv + np.array(shift, np.float32))
        np.testing.assert_allclose(a, b, rtol=0, atol=2e-5)
    # scale/offset: texinfo 1 at uv equals texinfo 0 at uv*scale+offset
    c = sc.tex_value(1, uv)
    d = sc.tex_value(0, uv * np.array([2, 3], np.float32) + np.array([0.5, 0.25], np.float32))
    np.testing.assert_array_equal(c, d)


def test_layers_arrays_and_fallbacks():
    l0 = np.full((4, 4, 4), 51, np.uint8)
    l1 = np.full((4, 4, 4), 204, np.uint8)
    arr = np.stack([l0, l1])  # one 4x4 array with two layers
    other = np.full((2, 2, 4), 102, np.uint8)
    infos = [_info(0, 0.4), _info(0, 0.6), _info(0, 7.0), _info(5, 0), _info(1, 0), _info(0, 1, uv_set=1)]
    sc = _scene_with_textures([arr, other], infos)
    uv = np.array([[0.3, 0.7]], np.float32)
    assert sc.tex_value(0, uv)[0, 0] == np.float32(51) / np.float32(255)   # layer 0.4 -> 0
    assert sc.tex_value(1, uv)[0, 0] == np.float32(204) / np.float32(255)  # 0.6 -> 1
    assert sc.tex_value(2, uv)[0, 0] == np.float32(204) / np.float32(255)  # clamped to the last layer
    np.testing.assert_array_equal(sc.tex_value(3, uv), np.ones((1, 4), np.float32))  # array out of range
    assert sc.tex_value(4, uv)[0, 0] == np.float32(102) / np.float32(255)
    np.testing.assert_array_equal(sc.tex_value(-1, uv), np.ones((1, 4), np.float32))  # untextured slot
    # uv set 1 is never written by the render state: (0, 0) whatever uv0 is
    np.testing.assert_array_equal(sc.tex_value(5, uv), sc.tex_value(5, np.zeros((1, 2), np.float32)))


def test_scene_texture_arrays_group_by_resolution():
    from dspt import scenes
    s = scenes.textured_spheres().scene
    dims, texels = s.getTextureArrays()
    assert dims.tolist() == [[64, 64, 3], [32, 32, 2], [48, 96, 1]]
    assert texels.size == int((dims[:, 0] * dims[:, 1] * dims[:, 2]).sum()) * 4
    info = s.getFlatTextureInfoBuffer().reshape(-1, 8)
    np.testing.assert_array_equal(info[1], np.float32([0.25, 0.1, 0, 0, 2, 3, 0, 0]))  # the checker again, offset + repeat


def _upload(eng, wl):
    from dspt import build_bvh
    s = wl.scene
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    s.bvhNodes, s.bvhTriangleIndices = nodes, bidx
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    img = wl.ibl["image"]
    return eng.upload_ibl(img["data"], img["width"], img["height"])


@pytest.mark.gpu
def test_textured_render_matches_oracle():
    from dspt import Engine, PROFILE_FULL, pack_frame_block, scenes
    wl = scenes.textured_spheres(n_spheres=22)
    W, H, F = 64, 48, 2
    eng = Engine(PROFILE_FULL)
    tot = _upload(eng, wl)
    s = wl.scene
    dims, texels = s.getTextureArrays()
    eng.scene_set_textures(0, s.getFlatTextureInfoBuffer(), dims, texels)
    img = wl.ibl["image"]
    blk = pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                           ibl_total_sum=tot, scene_counts=(s.num_triangles, len(s.bvhNodes) // 16,
                                                             len(s.bvhTriangleIndices), 0))
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.set_record_samples(True)
    eng.render(1, F)
    eng.synchronize()
    g_s, g_sig, g_prim = eng.read_samples(F)
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer(),
                        wl.ibl, 1)
    osc.set_textures(s.getFlatTextureInfoBuffer(), dims, texels)
    o = osc.render(blk, W, H, 1, F, threads=8)
    from test_gpu_parity import _check_samples
    _check_samples(g_s, g_sig, g_prim, o)
    # the textures change the image: the same scene without them is refused
    eng2 = Engine(PROFILE_FULL)
    _upload(eng2, wl)
    eng2.resize(W, H)
    eng2.set_frame(blk)
    from dspt import DsptError
    with pytest.raises(DsptError) as e:
        eng2.render(1, 1)
    assert e.value.code == -3


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6, 8, 9, 13])
def test_textured_debug_views_match_oracle(mode):
    """Debug views of textured materials (debug.glsl:1-83 over the render
    state of material.glsl:5-110): albedo, metalness and roughness read
    through their textures, normals / tangents / bitangents after the normal
    map, uv0, clearcoat and specular tint, sample for sample against the
    oracle (the bar of test_gpu_parity.py) and the accumulation bit for bit."""
    from dspt import Engine, PROFILE_FULL, pack_frame_block, scenes
    from test_gpu_parity import _check_samples
    wl = scenes.textured_spheres(n_spheres=22)
    W, H, F = 64, 48, 2
    eng = Engine(PROFILE_FULL)
    tot = _upload(eng, wl)
    s = wl.scene
    dims, texels = s.getTextureArrays()
    eng.scene_set_textures(0, s.getFlatTextureInfoBuffer(), dims, texels)
    img = wl.ibl["image"]
    blk = pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                           ibl_total_sum=tot, scene_counts=(s.num_triangles, len(s.bvhNodes) // 16,
                                                             len(s.bvhTriangleIndices), 0), debug_mode=mode)
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.set_record_samples(True)
    eng.render(1, F)
    acc = eng.read_accum()
    g_s, g_sig, g_prim = eng.read_samples(F)
    eng.close()
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer(),
                        wl.ibl, 1)
    osc.set_textures(s.getFlatTextureInfoBuffer(), dims, texels)
    o = osc.render(blk, W, H, 1, F, threads=8)
    _check_samples(g_s, g_sig, g_prim, o)
    np.testing.assert_array_equal(acc.view(np.uint32), o["accum"].view(np.uint32))
    assert (g_prim >= 0).mean() > 0.2
