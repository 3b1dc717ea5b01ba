"""Ingest: Radiance HDR images -> IBLs and glTF 2.0 files -> scenes, from
local files (no network).

parseIblFromHdrBuffer / loadIblFromFile
    packages/lib/loaders/ibl_loader.ts:16-190: the header up to the
    '-Y h +X w' line (GAMMA= / EXPOSURE= kept as metadata), FORMAT must be
    32-bit_rle_rgbe, new-style RLE scanlines only (width 8..32767), colour =
    m * 2^(e-128) / 255 (e > 0) in doubles rounded to f32, alpha 1, rows in
    file order.

loadSceneFromFile(path, instanced=False)
    packages/lib/loaders/scene_loader.ts:99-633 restated over a small glTF
    reader of our own (the reference uses @gltf-transform 4.3.0, which is not
    available here): the first scene's node graph is walked depth first with
    a LIFO stack (the last root first; children popped last first), every
    node's world matrix = parent * local (f32 entries, sums in doubles), and
    each indexed primitive with POSITION is baked into the 20-float vertex
    stream: positions by the world matrix, normals and tangents by its normal
    matrix and renormalised (scene_loader.ts:244-345, 577-611), uv0 / uv1,
    COLOR_0 (RGB gets alpha 1). Quantised / normalised integer attributes are
    dequantised first (as the reference's dequantize() transform). Materials
    (scene_loader.ts:357-478) map the core PBR model and the KHR_materials_*
    extensions onto MaterialData; textures go through addTexture with their
    KHR_texture_transform offset/scale and texCoord. KHR_lights_punctual on a
    mesh node adds a Light at the node's world position.
    instanced=True keeps every (node, primitive) as an instance of an
    object-space mesh instead (dspt.scenes.InstancedScene): its flatten() is
    the same scene in the same triangle order.

Images: PNG through dspt.image.read_png; other formats (JPEG, WebP) through
Pillow when it is importable, else an error naming the texture.
"""
import base64
import json
import math
import os
import struct
import time

import numpy as np

from .scene import Light, MaterialData, PathtracingSceneData, VERTEX_STRIDE


# ----------------------------------------------------------------- HDR
class HdrError(ValueError):
    pass


def _hdr_error(msg):
    raise HdrError(f"HDR parse error: {msg}")


def parseIblFromHdrBuffer(data):
    """bytes -> {'image': {'data': f32[4*w*h], 'width', 'height'}, 'header', 'gamma', 'exposure'}."""
    data = bytes(data)
    n = len(data)
    pos = 0

    def read_line():
        nonlocal pos
        if pos >= n:
            return ""
        end = data.find(b"\n", pos)
        end = n if end < 0 else end + 1
        s = data[pos:end].decode("latin-1")
        pos = end
        return s

    magic = read_line()
    if not magic.startswith("#?"):
        _hdr_error("bad initial token")
    header, gamma, exposure, fmt, width, height = magic, 1.0, 1.0, "", 0, 0
    import re
    dims = re.compile(r"^\s*-Y\s+(\d+)\s+\+X\s+(\d+)\s*$")
    while True:
        line = read_line()
        if not line:
            _hdr_error("no header found")
        header += line
        if line.startswith("GAMMA="):
            gamma = float(line[6:].split()[0]) if line[6:].split() else float("nan")
            continue
        if line.startswith("EXPOSURE="):
            exposure = float(line[9:].split()[0]) if line[9:].split() else float("nan")
            continue
        if line.startswith("FORMAT="):
            fmt = line[7:].strip()
            continue
        m = dims.match(line)
        if m:
            height, width = int(m.group(1)), int(m.group(2))
            break
    if fmt != "32-bit_rle_rgbe":
        _hdr_error(f'unsupported format "{fmt}"')
    if width <= 0 or height <= 0:
        _hdr_error("invalid image dimensions")
    if width < 8 or width > 0x7FFF:
        _hdr_error("unsupported scanline width for RLE")
    buf = np.frombuffer(data, np.uint8)
    px = np.zeros((height, width, 4), np.uint8)
    line_buf = np.zeros((4, width), np.uint8)
    for y in range(height):
        if pos + 4 > n:
            _hdr_error("unexpected end of file")
        a, b, c, d = (int(v) for v in buf[pos:pos + 4])
        pos += 4
        if a != 2 or b != 2 or ((c << 8) | d) != width:
            _hdr_error("bad rgbe scanline format")
        for ch in range(4):
            x = 0
            while x < width:
                if pos >= n:
                    _hdr_error("unexpected end of file in scanline")
                count = int(buf[pos])
                pos += 1
                if count > 128:
                    run = count - 128
                    if run <= 0 or x + run > width or pos >= n:
                        _hdr_error("bad scanline data")
                    line_buf[ch, x:x + run] = buf[pos]
                    pos += 1
                    x += run
                else:
                    if count <= 0 or x + count > width or pos + count > n:
                        _hdr_error("bad scanline data")
                    line_buf[ch, x:x + count] = buf[pos:pos + count]
                    pos += count
                    x += count
        px[y] = line_buf.T
    e = px[..., 3].astype(np.float64)
    scale = np.where(e > 0, np.power(2.0, e - 128.0) / 255.0, 0.0)
    img = np.ones((height, width, 4), np.float32)
    img[..., :3] = (px[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)
    img[..., :3][e == 0] = 0.0
    return {"image": {"data": img.reshape(-1), "width": width, "height": height}, "header": header, "gamma": gamma,
            "exposure": exposure}


def loadIblFromFile(path):
    with open(path, "rb") as f:
        return parseIblFromHdrBuffer(f.read())


# ----------------------------------------------------------------- glTF reader
_COMP = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32, 5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfFile:
    """JSON + binary buffers of a .gltf (external or data-URI buffers) or .glb."""

    def __init__(self, path):
        self.dir = os.path.dirname(os.path.abspath(path))
        raw = open(path, "rb").read()
        bin_chunk = None
        if raw[:4] == b"glTF":
            _, version, length = struct.unpack("<4sII", raw[:12])
            if version != 2:
                raise ValueError(f"GLB version {version} is not supported")
            pos, js = 12, None
            while pos < length:
                clen, ctype = struct.unpack("<II", raw[pos:pos + 8])
                body = raw[pos + 8:pos + 8 + clen]
                if ctype == 0x4E4F534A:
                    js = json.loads(body.decode("utf-8"))
                elif ctype == 0x004E4942 and bin_chunk is None:
                    bin_chunk = body
                pos += 8 + clen
            if js is None:
                raise ValueError("GLB without a JSON chunk")
            self.js = js
        else:
            self.js = json.loads(raw.decode("utf-8"))
        if not str(self.js.get("asset", {}).get("version", "2.0")).startswith("2"):
            raise ValueError("only glTF 2.x is supported")
        self.buffers = []
        for i, b in enumerate(self.js.get("buffers", [])):
            uri = b.get("uri")
            if uri is None:
                if bin_chunk is None:
                    raise ValueError(f"buffer {i} has no uri and there is no GLB binary chunk")
                self.buffers.append(bin_chunk)
            else:
                self.buffers.append(self._load_uri(uri))
        self._acc_cache = {}

    def _load_uri(self, uri):
        if uri.startswith("data:"):
            head, _, payload = uri.partition(",")
            return base64.b64decode(payload) if head.endswith(";base64") else payload.encode("latin-1")
        if "://" in uri:
            raise ValueError(f"remote resource {uri!r}: only local files can be loaded")
        from urllib.parse import unquote
        with open(os.path.join(self.dir, unquote(uri)), "rb") as f:
            return f.read()

    def view_bytes(self, vi):
        v = self.js["bufferViews"][vi]
        buf = self.buffers[v["buffer"]]
        off = v.get("byteOffset", 0)
        return buf[off:off + v["byteLength"]], v.get("byteStride")

    def accessor(self, ai, dequantize=True):
        """Accessor -> f32 (dequantised, as the reference's dequantize()
        transform leaves it) or the raw integer array, shape [count, ncomp]."""
        key = (ai, dequantize)
        if key in self._acc_cache:
            return self._acc_cache[key]
        a = self.js["accessors"][ai]
        dt = np.dtype(_COMP[a["componentType"]]).newbyteorder("<")
        nc = _NCOMP[a["type"]]
        count = a["count"]
        if "bufferView" in a:
            data, stride = self.view_bytes(a["bufferView"])
            off = a.get("byteOffset", 0)
            esize = dt.itemsize * nc
            stride = stride or esize
            if stride == esize:
                arr = np.frombuffer(data, dt, count * nc, off).reshape(count, nc)
            else:
                rows = np.frombuffer(data, np.uint8, (count - 1) * stride + esize, off) if count else np.zeros(0, np.uint8)
                idx = (np.arange(count)[:, None] * stride + np.arange(esize)[None, :]).reshape(-1)
                arr = rows[idx].view(dt).reshape(count, nc) if count else np.zeros((0, nc), dt)
            arr = arr.copy()
        else:
            arr = np.zeros((count, nc), dt)
        sp = a.get("sparse")
        if sp:
            ii = sp["indices"]
            idt = np.dtype(_COMP[ii["componentType"]]).newbyteorder("<")
            d, _ = self.view_bytes(ii["bufferView"])
            idx = np.frombuffer(d, idt, sp["count"], ii.get("byteOffset", 0)).astype(np.int64)
            vv = sp["values"]
            d, _ = self.view_bytes(vv["bufferView"])
            vals = np.frombuffer(d, dt, sp["count"] * nc, vv.get("byteOffset", 0)).reshape(-1, nc)
            arr[idx] = vals
        if dequantize and arr.dtype != np.float32:
            f = arr.astype(np.float32)
            if a.get("normalized", False):
                if arr.dtype == np.int8:
                    f = np.maximum(f / np.float32(127.0), np.float32(-1.0))
                elif arr.dtype == np.uint8:
                    f = f / np.float32(255.0)
                elif arr.dtype == np.int16:
                    f = np.maximum(f / np.float32(32767.0), np.float32(-1.0))
                elif arr.dtype == np.uint16:
                    f = f / np.float32(65535.0)
            arr = f
        self._acc_cache[key] = arr
        return arr

    def image_bytes(self, ii):
        im = self.js["images"][ii]
        if "bufferView" in im:
            data, _ = self.view_bytes(im["bufferView"])
            return bytes(data), im.get("mimeType", "")
        uri = im["uri"]
        mime = im.get("mimeType", "")
        if uri.startswith("data:"):
            mime = mime or uri[5:uri.find(";")]
        elif not mime:
            ext = os.path.splitext(uri)[1].lower()
            mime = {".png": "image/png", ".jpg": "image/jpeg", ".jpeg": "image/jpeg", ".webp": "image/webp",
                    ".avif": "image/avif"}.get(ext, "")
        return self._load_uri(uri), mime


# ----------------------------------------------------------------- matrices (scene_loader.ts:555-611)
def _identity():
    m = np.zeros(16, np.float32)
    m[0] = m[5] = m[10] = m[15] = 1
    return m


def _node_matrix(node):
    """Column-major f32 local matrix: `matrix` as given, else T * R * S
    composed in doubles (scene_loader.ts:555-571)."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).astype(np.float32)
    t = node.get("translation", [0, 0, 0])
    qx, qy, qz, qw = node.get("rotation", [0, 0, 0, 1])
    sx, sy, sz = node.get("scale", [1, 1, 1])
    x2, y2, z2 = qx * 2, qy * 2, qz * 2
    xx, xy, xz = qx * x2, qx * y2, qx * z2
    yy, yz, zz = qy * y2, qy * z2, qz * z2
    wx, wy, wz = qw * x2, qw * y2, qw * z2
    m = [0.0] * 16
    m[0], m[1], m[2] = (1 - (yy + zz)) * sx, (xy + wz) * sx, (xz - wy) * sx
    m[4], m[5], m[6] = (xy - wz) * sy, (1 - (xx + zz)) * sy, (yz + wx) * sy
    m[8], m[9], m[10] = (xz + wy) * sz, (yz - wx) * sz, (1 - (xx + yy)) * sz
    m[12], m[13], m[14], m[15] = t[0], t[1], t[2], 1
    return np.asarray(m, np.float64).astype(np.float32)


def _mul4(a, b):
    """multiplyMat4: column-major, each entry summed in doubles, stored f32."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    out = np.zeros(16, np.float64)
    for row in range(4):
        for col in range(4):
            s = 0.0
            for k in range(4):
                s += a[row + k * 4] * b[k + col * 4]
            out[row + col * 4] = s
    return out.astype(np.float32)


def _normal_matrix(e):
    """getNormalMatrix: the inverse transpose of the upper 3x3 (f32 result),
    identity when |det| <= 1e-12."""
    e = e.astype(np.float64)
    a00, a01, a02 = e[0], e[4], e[8]
    a10, a11, a12 = e[1], e[5], e[9]
    a20, a21, a22 = e[2], e[6], e[10]
    b01 = a22 * a11 - a12 * a21
    b11 = -a22 * a10 + a12 * a20
    b21 = a21 * a10 - a11 * a20
    det = a00 * b01 + a01 * b11 + a02 * b21
    if abs(det) <= 1e-12:
        return np.asarray([1, 0, 0, 0, 1, 0, 0, 0, 1], np.float32)
    det = 1 / det
    return np.asarray([b01 * det, (-a22 * a01 + a02 * a21) * det, (a12 * a01 - a02 * a11) * det,
                       b11 * det, (a22 * a00 - a02 * a20) * det, (-a12 * a00 + a02 * a10) * det,
                       b21 * det, (-a21 * a00 + a01 * a20) * det, (a11 * a00 - a01 * a10) * det], np.float64).astype(np.float32)


def _xf_points(p, m):
    """transformPoint over rows of p (f32), doubles, left-to-right sums."""
    p = p.astype(np.float64)
    m = m.astype(np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([((m[0] * x + m[4] * y) + m[8] * z) + m[12], ((m[1] * x + m[5] * y) + m[9] * z) + m[13],
                     ((m[2] * x + m[6] * y) + m[10] * z) + m[14]], 1).astype(np.float32)


def _xf_dirs(d, n):
    """transformDir: normal matrix, then normalised when longer than 1e-8."""
    d = d.astype(np.float64)
    n = n.astype(np.float64)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    r = np.stack([(n[0] * x + n[3] * y) + n[6] * z, (n[1] * x + n[4] * y) + n[7] * z, (n[2] * x + n[5] * y) + n[8] * z], 1)
    ln = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
    ok = ln > 1e-8
    r[ok] = r[ok] / ln[ok, None]
    return r.astype(np.float32)


def _affine_rows(m):
    """Column-major mat4 -> row-major 3x4 object->world affine."""
    return np.asarray(m, np.float32).reshape(4, 4).T[:3].copy()


# ----------------------------------------------------------------- images / textures
def _decode_image(data, mime, name):
    if mime == "image/png":
        from .image import read_png
        return read_png(data)
    if mime not in ("image/jpeg", "image/webp", "image/avif"):
        raise ValueError(f'Unsupported texture "{name}" ({mime or "unknown mime"})')
    try:
        import io
        from PIL import Image
    except ImportError:
        raise ValueError(f'texture "{name}" ({mime}) needs Pillow, which is not importable') from None
    with Image.open(io.BytesIO(data)) as im:
        return np.asarray(im.convert("RGBA"), np.uint8).copy()


def _digest(data):
    """Content key of an image (the reference keys its texture cache by a
    hash of the bytes, scene_loader.ts:524-533; any collision-free digest
    gives the same deduplication)."""
    import hashlib
    return hashlib.sha1(data).hexdigest()[:16]


class _Loader:
    def __init__(self, g, scene):
        self.g = g
        self.scene = scene
        self.tex_cache = {}
        self.mat_cache = {}
        self.img_cache = {}

    def _image(self, ii):
        if ii not in self.img_cache:
            data, mime = self.g.image_bytes(ii)
            name = self.g.js["images"][ii].get("name", self.g.js["images"][ii].get("uri", ""))
            self.img_cache[ii] = (data, mime, name, _digest(data))
        return self.img_cache[ii]

    def texture(self, info, color_space):
        """ensureTexture (scene_loader.ts:117-129) + createSceneTexture
        (:491-522): one TexInfo per (image, colour space, uv set, transform);
        one texture-array layer per (image, colour space, size)."""
        if not info or "index" not in info:
            return -1
        tex = self.g.js["textures"][info["index"]]
        src = tex.get("source")
        if src is None:
            for ext in tex.get("extensions", {}).values():
                if "source" in ext:
                    src = ext["source"]
                    break
        if src is None:
            raise ValueError(f"texture {info['index']} has no image source")
        data, mime, name, h = self._image(src)
        tt = info.get("extensions", {}).get("KHR_texture_transform", {})
        off = tt.get("offset", [0, 0])
        scale = tt.get("scale", [1, 1])
        uv = info.get("texCoord", 0)
        key = (f"{name}:{mime}:{color_space}:{h}", uv, tuple(off), tuple(scale))
        if key not in self.tex_cache:
            img = _decode_image(data, mime, name)
            uuid = f"{name}:{mime}:{color_space}:{img.shape[1]}x{img.shape[0]}:{h}"
            self.tex_cache[key] = self.scene.addTexture(img, offset=(off[0], off[1]), repeat=(scale[0], scale[1]),
                                                        uvSet=uv, key=uuid)
        return self.tex_cache[key]

    def material(self, mi):
        """ensureMaterial + createMaterialData (scene_loader.ts:357-478), in
        first-use order; the glTF defaults are those the reader reports."""
        if mi in self.mat_cache:
            return self.mat_cache[mi]
        m = self.g.js["materials"][mi]
        pbr = m.get("pbrMetallicRoughness", {})
        ext = m.get("extensions", {})
        d = MaterialData()
        bc = pbr.get("baseColorFactor", [1, 1, 1, 1])
        d.baseColorFactor = [bc[0], bc[1], bc[2]]
        mode = m.get("alphaMode", "OPAQUE")
        d.cutoutOpacity = 1.0 if mode == "OPAQUE" else bc[3]
        d.alphaCutoff = m.get("alphaCutoff", 0.5) if mode == "MASK" else (1.0 if mode == "OPAQUE" else 0.0)
        d.metallicFactor = pbr.get("metallicFactor", 1.0)
        d.roughnessFactor = pbr.get("roughnessFactor", 1.0)
        d.normalScale = m.get("normalTexture", {}).get("scale", 1.0)
        d.emissiveFactor = list(m.get("emissiveFactor", [0, 0, 0]))
        d.doubleSided = 1 if m.get("doubleSided", False) else 0
        d.baseColorTextureId = self.texture(pbr.get("baseColorTexture"), "srgb")
        d.metallicRoughnessTextureId = self.texture(pbr.get("metallicRoughnessTexture"), "linear")
        d.normalTextureId = self.texture(m.get("normalTexture"), "linear")
        d.emissiveTextureId = self.texture(m.get("emissiveTexture"), "srgb")
        if "KHR_materials_emissive_strength" in ext:
            s = ext["KHR_materials_emissive_strength"].get("emissiveStrength", 1.0)
            d.emissiveFactor = [c * s for c in d.emissiveFactor]
        if "KHR_materials_ior" in ext:
            d.ior = ext["KHR_materials_ior"].get("ior", 1.5)
        if "KHR_materials_specular" in ext:
            e = ext["KHR_materials_specular"]
            d.specularFactor = e.get("specularFactor", 1.0)
            d.specularColorFactor = e.get("specularColorFactor", [1, 1, 1])
            d.specularTextureId = self.texture(e.get("specularTexture"), "linear")
            d.specularColorTextureId = self.texture(e.get("specularColorTexture"), "srgb")
        if "KHR_materials_transmission" in ext:
            e = ext["KHR_materials_transmission"]
            d.transmissionFactor = e.get("transmissionFactor", 0.0)
            d.transmissionTextureId = self.texture(e.get("transmissionTexture"), "linear")
        if "KHR_materials_volume" in ext:
            e = ext["KHR_materials_volume"]
            d.thinWalled = 1 if e.get("thicknessFactor", 0.0) == 0 else 0
            d.attenuationColor = e.get("attenuationColor", [1, 1, 1])
            dist = e.get("attenuationDistance", math.inf)
            d.attenuationDistance = np.finfo(np.float64).max if (not dist or dist == 0) else dist
        if "KHR_materials_clearcoat" in ext:
            e = ext["KHR_materials_clearcoat"]
            d.clearcoatFactor = e.get("clearcoatFactor", 0.0)
            d.clearcoatRoughnessFactor = e.get("clearcoatRoughnessFactor", 0.0)
            d.clearcoatNormalTextureScale = e.get("clearcoatNormalTexture", {}).get("scale", 1.0)
            d.clearcoatTextureId = self.texture(e.get("clearcoatTexture"), "linear")
            d.clearcoatRoughnessTextureId = self.texture(e.get("clearcoatRoughnessTexture"), "linear")
            d.clearcoatNormalTextureId = self.texture(e.get("clearcoatNormalTexture"), "linear")
        if "KHR_materials_sheen" in ext:
            e = ext["KHR_materials_sheen"]
            d.sheenColorFactor = e.get("sheenColorFactor", [0, 0, 0])
            d.sheenRoughnessFactor = e.get("sheenRoughnessFactor", 0.0)
            d.sheenColorTextureId = self.texture(e.get("sheenColorTexture"), "srgb")
            d.sheenRoughnessTextureId = self.texture(e.get("sheenRoughnessTexture"), "linear")
        if "KHR_materials_anisotropy" in ext:
            e = ext["KHR_materials_anisotropy"]
            rot = e.get("anisotropyRotation", 0.0)
            d.anisotropy = e.get("anisotropyStrength", 0.0)
            d.anisotropyDirection = [math.cos(rot), math.sin(rot), 0.0]
            d.anisotropyTextureId = self.texture(e.get("anisotropyTexture"), "linear")
        if "KHR_materials_iridescence" in ext:
            e = ext["KHR_materials_iridescence"]
            d.iridescenceFactor = e.get("iridescenceFactor", 0.0)
            d.iridescenceIor = e.get("iridescenceIor", 1.3)
            d.iridescenceThicknessMinimum = e.get("iridescenceThicknessMinimum", 100.0)
            d.iridescenceThicknessMaximum = e.get("iridescenceThicknessMaximum", 400.0)
            d.iridescenceTextureId = self.texture(e.get("iridescenceTexture"), "linear")
            d.iridescenceThicknessTextureId = self.texture(e.get("iridescenceThicknessTexture"), "linear")
        if "KHR_materials_dispersion" in ext:
            d.dispersion = ext["KHR_materials_dispersion"].get("dispersion", 0.0)
        if "KHR_materials_diffuse_transmission" in ext:
            e = ext["KHR_materials_diffuse_transmission"]
            d.diffuseTransmissionFactor = e.get("diffuseTransmissionFactor", 0.0)
            d.diffuseTransmissionColorFactor = e.get("diffuseTransmissionColorFactor", [1, 1, 1])
            d.diffuseTransmissionTextureId = self.texture(e.get("diffuseTransmissionTexture"), "linear")
            d.diffuseTransmissionColorTextureId = self.texture(e.get("diffuseTransmissionColorTexture"), "srgb")
        idx = self.scene.addMaterial(d)
        self.mat_cache[mi] = idx
        return idx

    def primitive_arrays(self, prim):
        """Object-space vertex stream rows (before the world transform),
        indices; None for a primitive the reference skips (no indices or no
        POSITION, scene_loader.ts:257-267)."""
        at = prim.get("attributes", {})
        if "indices" not in prim or "POSITION" not in at:
            return None
        g = self.g
        pos = g.accessor(at["POSITION"])
        V = len(pos)
        vb = np.zeros((V, VERTEX_STRIDE), np.float32)
        vb[:, 0:3] = pos[:, :3]
        if "NORMAL" in at:
            vb[:, 4:7] = g.accessor(at["NORMAL"])[:, :3]
        if "TEXCOORD_0" in at:
            vb[:, 8:10] = g.accessor(at["TEXCOORD_0"])[:, :2]
        if "TEXCOORD_1" in at:
            vb[:, 10:12] = g.accessor(at["TEXCOORD_1"])[:, :2]
        if "TANGENT" in at:
            vb[:, 12:16] = g.accessor(at["TANGENT"])[:, :4]
        vb[:, 16:20] = 1.0
        if "COLOR_0" in at:
            c = g.accessor(at["COLOR_0"])
            vb[:, 16:16 + c.shape[1]] = c
        idx = g.accessor(prim["indices"], dequantize=False).reshape(-1).astype(np.uint32)
        return vb, idx, "NORMAL" in at, "TANGENT" in at


def _walk(g):
    """(node index, world matrix) in the reference's traversal order."""
    scenes = g.js.get("scenes", [])
    if not scenes:
        raise ValueError("This model contains no scene.")
    nodes = g.js.get("nodes", [])
    stack = [(n, _identity()) for n in scenes[0].get("nodes", [])]
    out = []
    while stack:
        ni, parent = stack.pop()
        node = nodes[ni]
        world = _mul4(parent, _node_matrix(node))
        for c in node.get("children", []):
            stack.append((c, world))
        out.append((ni, world))
    return out


def loadSceneFromFile(path, instanced=False, progress=None):
    """glTF / GLB file -> (PathtracingSceneData, or InstancedScene when
    instanced, {profile ms})."""
    prof = {}
    t0 = time.perf_counter()
    if progress:
        progress({"phase": "Reading files"})
    g = GltfFile(path)
    prof["readMs"] = (time.perf_counter() - t0) * 1e3
    t1 = time.perf_counter()
    if progress:
        progress({"phase": "Parsing scene"})
    if instanced:
        from .scenes import InstancedScene
        out = InstancedScene()
        scene = out.scene
    else:
        scene = PathtracingSceneData()
    L = _Loader(g, scene)
    lights_ext = g.js.get("extensions", {}).get("KHR_lights_punctual", {}).get("lights", [])
    verts, tris, bvh_pos = [], [], []
    nv = nt = 0
    base = {}  # (mesh, primitive) -> base-mesh index (instanced)
    o2w, inst_mesh = [], []
    for ni, world in _walk(g):
        node = g.js["nodes"][ni]
        if "mesh" not in node:
            continue
        li = node.get("extensions", {}).get("KHR_lights_punctual", {}).get("light")
        if li is not None and li < len(lights_ext):
            lt = lights_ext[li]
            col, inten = lt.get("color", [1, 1, 1]), lt.get("intensity", 1.0)
            light = Light(position=(float(world[12]), float(world[13]), float(world[14])),
                          emission=(col[0] * inten, col[1] * inten, col[2] * inten))
            light.type = 0 if lt.get("type") == "point" else 1
            scene.addLight(light)
        mesh = g.js["meshes"][node["mesh"]]
        for pi, prim in enumerate(mesh.get("primitives", [])):
            mat = L.material(prim["material"]) if "material" in prim else 0
            arr = L.primitive_arrays(prim)
            if arr is None:
                continue
            vb, idx, has_n, has_t = arr
            vb = vb.copy()
            vb[:, 3] = mat
            if instanced:
                key = (node["mesh"], pi, mat)
                if key not in base:
                    base[key] = len(out.mesh_first)
                    out.mesh_vfirst.append(nv)
                    out.mesh_vcount.append(len(vb))
                    out.mesh_first.append(nt)
                    out.mesh_count.append(len(idx) // 3)
                    verts.append(vb)
                    tris.append(idx.astype(np.int64) + nv)
                    nv += len(vb)
                    nt += len(idx) // 3
                o2w.append(_affine_rows(world))
                inst_mesh.append(base[key])
                continue
            nm = _normal_matrix(world)
            p = _xf_points(vb[:, 0:3], world)
            vb[:, 0:3] = p
            if has_n:
                vb[:, 4:7] = _xf_dirs(vb[:, 4:7], nm)
            if has_t:
                vb[:, 12:15] = _xf_dirs(vb[:, 12:15], nm)
            verts.append(vb)
            tris.append(idx.astype(np.int64) + nv)
            bvh_pos.append(p)
            nv += len(vb)
            nt += len(idx) // 3
    if not verts:
        raise ValueError("the scene has no indexed triangle primitives")
    if scene.num_materials == 0:
        scene.addMaterial(MaterialData())  # primitives without a material use index 0
    V = np.concatenate(verts)
    scene.vertexBuffer = V.reshape(-1)
    scene.triangleIndexBuffer = np.concatenate(tris).astype(np.uint32).reshape(-1)
    scene.bvhPositionBuffer = np.ascontiguousarray(V[:, 0:3]).reshape(-1)
    scene.bvhIndexBuffer = scene.triangleIndexBuffer
    prof["parseMs"] = (time.perf_counter() - t1) * 1e3
    if instanced:
        out.o2w = np.asarray(o2w, np.float32).reshape(-1, 3, 4)
        out.inst_mesh = np.asarray(inst_mesh, np.int32)
        return out, prof
    return scene, prof
