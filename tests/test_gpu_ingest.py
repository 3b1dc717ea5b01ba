"""Files to image on the GPU: a glTF scene (node hierarchy, KHR materials, a
PNG texture) and a Radiance HDR environment written here are loaded by the
viewport (loadSceneFromFile / loadIblFromFile), rendered by the engine, and
compared with the oracle rendering the same loaded buffers (primary hit ids
exact, same-branch samples within the full-profile bars of
test_gpu_parity.py). Then the display chain writes a PNG of the canvas."""
import math

import numpy as np
import pytest

import oracle_ref as O

pytestmark = pytest.mark.gpu


def _rgbe(img):
    """f32 RGB -> RGBE bytes (Radiance: e = ceil(log2 max) + 128, m = v / 2^(e-128) * 256)."""
    mx = img.max(axis=-1)
    out = np.zeros(img.shape[:2] + (4,), np.uint8)
    ok = mx > 1e-32
    e = np.zeros_like(mx, dtype=np.int64)
    e[ok] = np.floor(np.log2(mx[ok])).astype(np.int64) + 1
    scale = np.where(ok, np.ldexp(1.0, (-e).astype(np.int64)) * 256.0, 0.0)
    out[..., :3] = np.clip(np.floor(img * scale[..., None]), 0, 255).astype(np.uint8)
    out[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    return out


def _write_files(tmp_path):
    from test_loaders import _Gltf, _png_bytes, _hdr_bytes
    from dspt import scenes
    import base64
    g = _Gltf()
    sp, sn, suv, sf = scenes.uv_sphere(1.0, 24, 12)
    pa = g.acc(sp.astype(np.float32), 5126, "VEC3")
    na = g.acc(sn.astype(np.float32), 5126, "VEC3")
    ua = g.acc(suv.astype(np.float32), 5126, "VEC2")
    ia = g.acc(sf.astype(np.uint32).reshape(-1), 5125, "SCALAR")
    gp = np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float32)
    gpa = g.acc(gp, 5126, "VEC3")
    gna = g.acc(np.tile(np.array([0, 1, 0], np.float32), (4, 1)), 5126, "VEC3")
    gua = g.acc(np.array([[0, 0], [3, 0], [3, 3], [0, 3]], np.float32), 5126, "VEC2")
    gia = g.acc(np.array([0, 2, 1, 0, 3, 2], np.uint16), 5123, "SCALAR")
    chk = np.zeros((16, 16, 4), np.uint8)
    chk[...] = 255
    chk[(np.add.outer(np.arange(16) // 4, np.arange(16) // 4) % 2).astype(bool), :3] = 40
    g.js["images"] = [{"uri": "data:image/png;base64," + base64.b64encode(_png_bytes(chk)).decode()}]
    g.js["textures"] = [{"source": 0}]
    g.js["materials"] = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.9, 0.9, 1], "metallicFactor": 0.0, "roughnessFactor": 0.8,
                                  "baseColorTexture": {"index": 0}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.3, 0.2, 1], "metallicFactor": 0.0, "roughnessFactor": 0.3},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 1.0, "clearcoatRoughnessFactor": 0.1}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [1, 1, 1, 1], "metallicFactor": 0.0, "roughnessFactor": 0.05},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 1.0},
                        "KHR_materials_volume": {"thicknessFactor": 1.0, "attenuationDistance": 1.5,
                                                 "attenuationColor": [0.9, 0.7, 0.5]},
                        "KHR_materials_ior": {"ior": 1.5}}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.9, 0.8, 0.5, 1], "metallicFactor": 1.0, "roughnessFactor": 0.35},
         "extensions": {"KHR_materials_anisotropy": {"anisotropyStrength": 0.6, "anisotropyRotation": 0.5}}},
    ]
    g.js["meshes"] = [{"primitives": [{"attributes": {"POSITION": gpa, "NORMAL": gna, "TEXCOORD_0": gua}, "indices": gia,
                                       "material": 0}]}] + \
                     [{"primitives": [{"attributes": {"POSITION": pa, "NORMAL": na, "TEXCOORD_0": ua}, "indices": ia,
                                       "material": m}]} for m in (1, 2, 3)]
    g.js["nodes"] = [{"mesh": 0}, {"children": [2, 3, 4], "translation": [0, 0.8, 0]},
                     {"mesh": 1, "translation": [-2.2, 0, 0], "scale": [0.8, 0.8, 0.8]},
                     {"mesh": 2, "translation": [0, 0, 0], "rotation": [0, math.sin(0.3), 0, math.cos(0.3)]},
                     {"mesh": 3, "translation": [2.2, 0, 0], "scale": [0.8, 1.0, 0.8]}]
    g.js["scenes"] = [{"nodes": [0, 1]}]
    g.write(str(tmp_path / "scene.glb"), glb=True)
    sky = scenes.sky_ibl(64, 32)["image"]
    img = sky["data"].reshape(32, 64, 4)[..., :3].astype(np.float64)
    (tmp_path / "sky.hdr").write_bytes(_hdr_bytes(_rgbe(img)))


def test_files_render_like_the_oracle(tmp_path):
    from dspt import ViewportCamera
    from dspt.renderer import PathtracingViewport
    _write_files(tmp_path)
    W, H, F = 64, 48, 2
    vp = PathtracingViewport(W, H, materialProfile="webgl-full")
    vp.renderer.maxBounces = 6
    loaded = vp.loadSceneFromFile(str(tmp_path / "scene.glb"))
    ibl = vp.loadIblFromFile(str(tmp_path / "sky.hdr"))
    s = loaded["scene"]
    assert s.num_triangles == 2 + 3 * 24 * 11 * 2 and s.num_materials == 4 and s.num_textures == 1
    cam = ViewportCamera(position=(0, 3, 8), target=(0, 0.8, 0), fov=45, near=0.01)
    r = vp.renderer
    eng = r.engine
    eng.set_record_samples(True)
    r.render(cam, F + 1)
    g_s, g_sig, g_prim = eng.read_samples(F)
    blk = r.frame_block(cam, 1)
    from dspt import build_bvh
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer(), ibl, 1)
    dims, texels = s.getTextureArrays()
    osc.set_textures(s.getFlatTextureInfoBuffer(), dims, texels)
    o = osc.render(blk, W, H, 1, F, threads=8)
    assert (g_prim >= 0).mean() > 0.5
    from test_gpu_parity import _check_samples
    _check_samples(g_s, g_sig, g_prim, o)
    # display chain -> PNG
    from dspt.image import read_png
    r.tonemapping = "AcesFilm"
    vp.saveImage(str(tmp_path / "out.png"))
    png = read_png(str(tmp_path / "out.png"))
    np.testing.assert_array_equal(png, r.readDisplay()[::-1])
    assert png[..., :3].std() > 5
