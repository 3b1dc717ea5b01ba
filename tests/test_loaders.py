"""Ingest: the Radiance HDR parser (ibl_loader.ts:16-190) and the glTF
loader (scene_loader.ts:99-633), on files written here.

HDR pins: RLE scanlines written from the Radiance format's definition decode
to the encoded bytes; colour = m * 2^(e-128) / 255 (closed form), e = 0 ->
black, alpha 1; each error branch of the reference parser raises.
glTF pins: world positions equal the node chain's affine applied in float64
(TRS composed from the quaternion's rotation matrix, computed here with
scipy's Rotation, and a `matrix` node); normals follow the inverse
transpose and stay unit (non-uniform scale); triangle order follows the
reference's LIFO walk; normalised u16 attributes dequantise to v/65535;
materials and KHR extensions land in their MaterialData slots; texture
transforms and texCoord land in TexInfo, one layer per image; .glb and
.gltf give the same scene; the instanced load flattens to the plain load.
"""
import base64
import json
import math
import os
import struct

import numpy as np
import pytest


# ----------------------------------------------------------------- HDR
def _rle_channel(vals):
    """New-style RLE of one channel: runs of >= 3 equal bytes as (128+n, v),
    the rest as literal packets (n, bytes...), n <= 127."""
    out, i, n = bytearray(), 0, len(vals)
    while i < n:
        j = i
        while j < n and j - i < 127 and vals[j] == vals[i]:
            j += 1
        if j - i >= 3:
            out += bytes([128 + (j - i), vals[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 127 and not (j + 2 < n and vals[j] == vals[j + 1] == vals[j + 2]):
            j += 1
        j = max(j, i + 1)
        out += bytes([j - i]) + bytes(vals[i:j])
        i = j
    return bytes(out)


def _hdr_bytes(rgbe, header_extra=b"", fmt=b"FORMAT=32-bit_rle_rgbe\n"):
    h, w, _ = rgbe.shape
    out = b"#?RADIANCE\n" + header_extra + fmt + b"\n" + f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 0xFF])
        for ch in range(4):
            out += _rle_channel(list(rgbe[y, :, ch]))
    return out


def test_hdr_decode():
    from dspt.loaders import parseIblFromHdrBuffer
    rng = np.random.default_rng(0)
    h, w = 5, 37
    rgbe = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    rgbe[:, 10:20] = rgbe[:, 10:11]  # runs
    rgbe[2, :, 3] = 0  # e == 0 -> black
    rgbe[1, 3, 3] = 255
    r = parseIblFromHdrBuffer(_hdr_bytes(rgbe, b"GAMMA=2.2\nEXPOSURE=0.5\n"))
    assert (r["image"]["width"], r["image"]["height"]) == (w, h)
    assert r["gamma"] == 2.2 and r["exposure"] == 0.5
    img = r["image"]["data"].reshape(h, w, 4)
    e = rgbe[..., 3].astype(np.float64)
    ref = rgbe[..., :3].astype(np.float64) * (2.0 ** (e - 128) / 255.0)[..., None]
    ref[e == 0] = 0
    np.testing.assert_array_equal(img[..., :3], ref.astype(np.float32))
    assert np.all(img[..., 3] == 1) and np.all(img[2, :, :3] == 0)


@pytest.mark.parametrize("case", ["token", "format", "width", "scanline", "eof"])
def test_hdr_errors(case):
    from dspt.loaders import HdrError, parseIblFromHdrBuffer
    rgbe = np.full((2, 16, 4), 128, np.uint8)
    data = _hdr_bytes(rgbe)
    if case == "token":
        data = b"XX" + data[2:]
    elif case == "format":
        data = data.replace(b"FORMAT=32-bit_rle_rgbe", b"FORMAT=32-bit_rle_xyze")
    elif case == "width":
        data = _hdr_bytes(np.full((2, 4, 4), 128, np.uint8))
    elif case == "scanline":
        i = data.index(b"+X 16\n") + 6
        data = data[:i] + bytes([1, 2]) + data[i + 2:]
    else:
        data = data[:-5]
    with pytest.raises(HdrError):
        parseIblFromHdrBuffer(data)


# ----------------------------------------------------------------- glTF
class _Gltf:
    """Minimal glTF writer for the tests."""

    def __init__(self):
        self.bin = bytearray()
        self.js = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                   "nodes": [], "scenes": [], "materials": [], "textures": [], "images": []}

    def view(self, data):
        while len(self.bin) % 4:
            self.bin += b"\0"
        self.js["bufferViews"].append({"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)})
        self.bin += data
        return len(self.js["bufferViews"]) - 1

    def acc(self, arr, ctype, typ, normalized=False):
        arr = np.ascontiguousarray(arr)
        v = self.view(arr.tobytes())
        a = {"bufferView": v, "componentType": ctype, "count": int(arr.shape[0]), "type": typ}
        if normalized:
            a["normalized"] = True
        if ctype == 5126 and typ == "VEC3":
            a["min"] = arr.min(0).tolist()
            a["max"] = arr.max(0).tolist()
        self.js["accessors"].append(a)
        return len(self.js["accessors"]) - 1

    def write(self, path, glb=False):
        js = json.loads(json.dumps(self.js))
        if glb:
            js["buffers"] = [{"byteLength": len(self.bin)}]
            jb = json.dumps(js).encode()
            jb += b" " * (-len(jb) % 4)
            bb = bytes(self.bin) + b"\0" * (-len(self.bin) % 4)
            body = struct.pack("<II", len(jb), 0x4E4F534A) + jb + struct.pack("<II", len(bb), 0x004E4942) + bb
            open(path, "wb").write(struct.pack("<4sII", b"glTF", 2, 12 + len(body)) + body)
        else:
            js["buffers"] = [{"byteLength": len(self.bin),
                              "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()}]
            open(path, "w").write(json.dumps(js))


def _png_bytes(img):
    import tempfile
    from dspt.image import write_png
    with tempfile.NamedTemporaryFile(suffix=".png", delete=False) as f:
        p = f.name
    write_png(p, img)
    b = open(p, "rb").read()
    os.remove(p)
    return b


def _quat(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return list(axis * math.sin(ang / 2)) + [math.cos(ang / 2)]


def _build():
    g = _Gltf()
    # mesh 0: a quad with normals, uv, tangents, RGB colours
    qp = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    qn = np.tile(np.array([0, 1, 0], np.float32), (4, 1))
    quv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    qt = np.tile(np.array([1, 0, 0, -1], np.float32), (4, 1))
    qc = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], np.float32)
    qi = np.array([0, 2, 1, 0, 3, 2], np.uint16)
    prim0 = {"attributes": {"POSITION": g.acc(qp, 5126, "VEC3"), "NORMAL": g.acc(qn, 5126, "VEC3"),
                            "TEXCOORD_0": g.acc(quv, 5126, "VEC2"), "TANGENT": g.acc(qt, 5126, "VEC4"),
                            "COLOR_0": g.acc(qc, 5126, "VEC3")}, "indices": g.acc(qi, 5123, "SCALAR"), "material": 1}
    # mesh 1: a triangle with normalised u16 positions and u8 normalised uv, u32 indices; plus a
    # primitive without indices (skipped, as the reference does)
    tp = np.array([[0, 0, 0], [65535, 0, 0], [0, 65535, 32768]], np.uint16)
    tuv = np.array([[0, 0], [255, 0], [0, 255]], np.uint8)
    ti = np.array([0, 1, 2], np.uint32)
    prim1 = {"attributes": {"POSITION": g.acc(tp, 5123, "VEC3", normalized=True),
                            "TEXCOORD_0": g.acc(tuv, 5121, "VEC2", normalized=True)},
             "indices": g.acc(ti, 5125, "SCALAR"), "material": 0}
    prim_skip = {"attributes": {"POSITION": g.acc(qp, 5126, "VEC3")}}
    g.js["meshes"] = [{"primitives": [prim0]}, {"primitives": [prim1, prim_skip]}]
    png = _png_bytes(np.arange(4 * 4 * 4, dtype=np.uint8).reshape(4, 4, 4))
    g.js["images"] = [{"uri": "data:image/png;base64," + base64.b64encode(png).decode()}]
    g.js["textures"] = [{"source": 0}]
    g.js["materials"] = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 0.125, 0.5], "metallicFactor": 0.3,
                                  "roughnessFactor": 0.7,
                                  "baseColorTexture": {"index": 0, "texCoord": 0,
                                                       "extensions": {"KHR_texture_transform": {"offset": [0.25, 0.5],
                                                                                                "scale": [2, 3]}}}},
         "alphaMode": "MASK", "alphaCutoff": 0.3, "doubleSided": True, "emissiveFactor": [1, 0.5, 0.25],
         "normalTexture": {"index": 0, "scale": 0.8, "texCoord": 1},
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                        "KHR_materials_ior": {"ior": 1.7},
                        "KHR_materials_transmission": {"transmissionFactor": 0.9},
                        "KHR_materials_volume": {"thicknessFactor": 0.5, "attenuationDistance": 2.0,
                                                 "attenuationColor": [0.9, 0.8, 0.7]},
                        "KHR_materials_clearcoat": {"clearcoatFactor": 0.6, "clearcoatRoughnessFactor": 0.2},
                        "KHR_materials_sheen": {"sheenColorFactor": [0.1, 0.2, 0.3], "sheenRoughnessFactor": 0.4},
                        "KHR_materials_anisotropy": {"anisotropyStrength": 0.5, "anisotropyRotation": 0.3},
                        "KHR_materials_iridescence": {"iridescenceFactor": 0.7, "iridescenceIor": 1.4},
                        "KHR_materials_dispersion": {"dispersion": 2.5},
                        "KHR_materials_specular": {"specularFactor": 0.6, "specularColorFactor": [0.9, 0.9, 0.8]},
                        "KHR_materials_diffuse_transmission": {"diffuseTransmissionFactor": 0.3}}},
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}},
    ]
    # nodes: 0 = root (translation), children 1 (TRS, mesh 0) and 2 (matrix, mesh 1); 3 = second root (mesh 0)
    M2 = np.array([[2, 0, 0, 1], [0, 1, 0, 2], [0, 0, 0.5, 3], [0, 0, 0, 1]], np.float64)
    g.js["nodes"] = [
        {"translation": [0, 1, 0], "children": [1, 2]},
        {"mesh": 0, "translation": [3, 0, 0], "rotation": _quat([0, 0, 1], 0.7), "scale": [1, 2, 3]},
        {"mesh": 1, "matrix": M2.T.reshape(-1).tolist()},
        {"mesh": 0, "translation": [-5, 0, 0]},
    ]
    g.js["scenes"] = [{"nodes": [0, 3]}]
    return g, dict(qp=qp, qn=qn, qt=qt, qc=qc, qi=qi, tp=tp, tuv=tuv, M2=M2)


def _trs(t, q, s):
    from scipy.spatial.transform import Rotation
    M = np.eye(4)
    M[:3, :3] = Rotation.from_quat(q).as_matrix() @ np.diag(s)
    M[:3, 3] = t
    return M


def test_gltf_world_space_bake(tmp_path):
    from dspt.loaders import loadSceneFromFile
    g, d = _build()
    path = str(tmp_path / "s.gltf")
    g.write(path)
    scene, prof = loadSceneFromFile(path)
    vb = scene.vertexBuffer.reshape(-1, 20)
    tri = scene.triangleIndexBuffer.reshape(-1, 3)
    assert len(vb) == 4 + 3 + 4 and len(tri) == 2 + 1 + 2  # the index-less primitive is skipped
    root = np.eye(4)
    root[:3, 3] = [0, 1, 0]
    n1 = root @ _trs([3, 0, 0], _quat([0, 0, 1], 0.7), [1, 2, 3])
    n2 = root @ d["M2"]
    n3 = np.eye(4)
    n3[:3, 3] = [-5, 0, 0]
    # LIFO walk: root 3 first (last root), then root 0, whose children pop 2 then 1
    blocks = [(n3, d["qp"], 0), (n2, d["tp"].astype(np.float64) / 65535.0, 1), (n1, d["qp"], 0)]
    v = 0
    for M, p, kind in blocks:
        k = len(p)
        ref = p @ M[:3, :3].T + M[:3, 3]
        np.testing.assert_allclose(vb[v:v + k, 0:3], ref, rtol=1e-6, atol=1e-6)
        if kind == 0:
            nrm = d["qn"] @ np.linalg.inv(M[:3, :3])  # (M^-T n) as rows
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            np.testing.assert_allclose(vb[v:v + k, 4:7], nrm, atol=1e-6)
            np.testing.assert_allclose(np.linalg.norm(vb[v:v + k, 12:15], axis=1), 1, atol=1e-6)
            assert np.all(vb[v:v + k, 15] == -1)
            np.testing.assert_array_equal(vb[v:v + k, 16:19], d["qc"])
            assert np.all(vb[v:v + k, 19] == 1)
            assert np.all(vb[v:v + k, 3] == 0)  # material 1 of the file is used first -> index 0
        else:
            np.testing.assert_allclose(vb[v:v + k, 8:10], d["tuv"] / 255.0, atol=1e-7)
            assert np.all(vb[v:v + k, 3] == 1)
            assert np.all(vb[v:v + k, 4:7] == 0) and np.all(vb[v:v + k, 16:20] == 1)
        v += k
    np.testing.assert_array_equal(tri[0:2], d["qi"].reshape(-1, 3))
    np.testing.assert_array_equal(tri[2], np.array([0, 1, 2]) + 4)
    np.testing.assert_array_equal(tri[3:5], d["qi"].reshape(-1, 3) + 7)
    np.testing.assert_array_equal(scene.bvhPositionBuffer.reshape(-1, 3), vb[:, 0:3])
    assert scene.bvhIndexBuffer is scene.triangleIndexBuffer


def test_gltf_materials_and_textures(tmp_path):
    from dspt.loaders import loadSceneFromFile
    g, _ = _build()
    path = str(tmp_path / "s.gltf")
    g.write(path)
    scene, _ = loadSceneFromFile(path)
    assert scene.num_materials == 2
    m1, m0 = scene.materials  # first use: file material 1, then 0
    assert m1.baseColorTextureId == 0 and m1.metallicFactor == 1.0 and m1.roughnessFactor == 1.0
    f32 = np.float32
    assert tuple(m0.baseColorFactor) == (f32(0.5), f32(0.25), f32(0.125))
    assert m0.cutoutOpacity == 0.5 and m0.alphaCutoff == f32(0.3) and m0.doubleSided == 1
    assert m0.metallicFactor == f32(0.3) and m0.roughnessFactor == f32(0.7) and m0.normalScale == f32(0.8)
    assert tuple(m0.emissiveFactor) == (4.0, 2.0, 1.0)
    assert m0.ior == f32(1.7) and m0.transmissionFactor == f32(0.9) and m0.thinWalled == 0
    assert m0.attenuationDistance == 2.0 and tuple(m0.attenuationColor) == (f32(0.9), f32(0.8), f32(0.7))
    assert m0.clearcoatFactor == f32(0.6) and m0.sheenRoughnessFactor == f32(0.4)
    assert m0.anisotropy == 0.5 and tuple(m0.anisotropyDirection) == (f32(math.cos(0.3)), f32(math.sin(0.3)), 0)
    assert m0.iridescenceFactor == f32(0.7) and m0.iridescenceIor == f32(1.4) and m0.dispersion == 2.5
    assert m0.specularFactor == f32(0.6) and m0.diffuseTransmissionFactor == f32(0.3)
    # texture infos: plain (material 1), transformed base colour, normal map on uv set 1; one image layer
    assert scene.num_textures == 3
    t_plain, t_base, t_norm = scene.texInfos
    assert m0.baseColorTextureId == 1 and m0.normalTextureId == 2
    # TexInfo f32[8]: offset.xy, array, layer, scale.xy, uv set
    assert t_base.data[[0, 1, 4, 5, 6]].tolist() == [0.25, 0.5, 2, 3, 0]
    assert t_norm.data[6] == 1 and t_plain.data[[4, 5]].tolist() == [1, 1]
    assert t_plain.data[3] == 0 and t_base.data[3] == 0 and t_norm.data[3] == 1
    dims, texels = scene.getTextureArrays()
    # srgb (base colour) and linear (normal map) copies of the same image are separate layers
    assert dims.tolist() == [[4, 4, 2]]
    np.testing.assert_array_equal(texels[:64], np.arange(64, dtype=np.uint8))


def test_glb_equals_gltf(tmp_path):
    from dspt.loaders import loadSceneFromFile
    g, _ = _build()
    g.write(str(tmp_path / "a.gltf"))
    g.write(str(tmp_path / "a.glb"), glb=True)
    a, _ = loadSceneFromFile(str(tmp_path / "a.gltf"))
    b, _ = loadSceneFromFile(str(tmp_path / "a.glb"))
    np.testing.assert_array_equal(a.vertexBuffer, b.vertexBuffer)
    np.testing.assert_array_equal(a.triangleIndexBuffer, b.triangleIndexBuffer)
    np.testing.assert_array_equal(a.getFlatMaterialBuffer(), b.getFlatMaterialBuffer())


def test_instanced_load_flattens_to_plain_load(tmp_path):
    from dspt.loaders import loadSceneFromFile
    g, _ = _build()
    path = str(tmp_path / "s.gltf")
    g.write(path)
    plain, _ = loadSceneFromFile(path)
    inst, _ = loadSceneFromFile(path, instanced=True)
    assert inst.num_instances == 3 and len(inst.mesh_count) == 2  # mesh 0 shared by two nodes
    flat, inst_of, _ = inst.flatten()
    np.testing.assert_array_equal(flat.triangleIndexBuffer, plain.triangleIndexBuffer)
    a, b = flat.vertexBuffer.reshape(-1, 20), plain.vertexBuffer.reshape(-1, 20)
    np.testing.assert_allclose(a, b, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(flat.getFlatMaterialBuffer(), plain.getFlatMaterialBuffer())
    np.testing.assert_array_equal(inst_of, [0, 0, 1, 2, 2])
