"""Host-side scene model mirroring the reference's data formats.

MaterialData      packages/lib/material.ts:2-481 (68 f32, same slots/defaults)
PathtracingSceneData  packages/lib/scene_data.ts:23-132 (20-f32 vertex stream,
                  u32 triangle indices, BVH position/index buffers)
ViewportCamera    packages/lib/pathtracing_viewport.ts:68-98 (look-at camera,
                  column-major matrixWorld in a Float32Array)
"""
import math

import numpy as np

VERTEX_STRIDE = 20  # scene_data.ts:21

# (name, first slot, width) of MaterialData, material.ts:24-390
_MAT_FIELDS = [
    ("baseColorFactor", 0, 3), ("metallicFactor", 3, 1), ("roughnessFactor", 4, 1), ("anisotropy", 5, 1),
    ("anisotropyRotation", 6, 1), ("transmissionFactor", 7, 1), ("cutoutOpacity", 8, 1), ("doubleSided", 9, 1),
    ("normalScale", 10, 1), ("ior", 11, 1), ("specularColorFactor", 12, 3), ("specularFactor", 15, 1),
    ("sheenColorFactor", 16, 3), ("sheenRoughnessFactor", 19, 1), ("emissiveFactor", 20, 3),
    ("clearcoatNormalTextureScale", 23, 1), ("clearcoatFactor", 24, 1), ("clearcoatRoughnessFactor", 25, 1),
    ("diffuseTransmissionFactor", 26, 1), ("alphaCutoff", 27, 1), ("attenuationColor", 28, 3),
    ("attenuationDistance", 31, 1), ("multiscatterColorFactor", 32, 3), ("thinWalled", 35, 1),
    ("anisotropyDirection", 36, 3), ("diffuseTransmissionTextureId", 39, 1), ("iridescenceFactor", 40, 1),
    ("iridescenceIor", 41, 1), ("iridescenceThicknessMinimum", 42, 1), ("iridescenceThicknessMaximum", 43, 1),
    ("baseColorTextureId", 44, 1), ("metallicRoughnessTextureId", 45, 1), ("normalTextureId", 46, 1),
    ("emissiveTextureId", 47, 1), ("specularTextureId", 48, 1), ("specularColorTextureId", 49, 1),
    ("transmissionTextureId", 50, 1), ("clearcoatTextureId", 51, 1), ("clearcoatRoughnessTextureId", 52, 1),
    ("clearcoatNormalTextureId", 53, 1), ("sheenColorTextureId", 54, 1), ("sheenRoughnessTextureId", 55, 1),
    ("anisotropyTextureId", 56, 1), ("anisotropyDirectionTextureId", 57, 1), ("iridescenceTextureId", 58, 1),
    ("iridescenceThicknessTextureId", 59, 1), ("diffuseTransmissionColorFactor", 60, 3),
    ("diffuseTransmissionColorTextureId", 63, 1), ("dispersion", 64, 1),
]

# constructor defaults, material.ts:415-480 (attenuationDistance = Number.MAX_VALUE -> +inf in f32)
_MAT_DEFAULTS = {
    "baseColorFactor": [1, 1, 1], "metallicFactor": 0, "roughnessFactor": 0, "anisotropy": 0.0,
    "anisotropyRotation": 0, "transmissionFactor": 0, "cutoutOpacity": 1, "doubleSided": 1, "normalScale": 1,
    "ior": 1.49, "specularColorFactor": [1, 1, 1], "specularFactor": 1, "sheenRoughnessFactor": 0,
    "sheenColorFactor": [0, 0, 0], "clearcoatNormalTextureScale": 1, "emissiveFactor": [0, 0, 0],
    "clearcoatFactor": 0, "clearcoatRoughnessFactor": 0, "diffuseTransmissionFactor": 0, "alphaCutoff": 0,
    "attenuationDistance": np.finfo(np.float64).max, "attenuationColor": [1, 1, 1],
    "multiscatterColorFactor": [1, 1, 1], "thinWalled": 1, "anisotropyDirection": [1, 0, 0],
    "diffuseTransmissionTextureId": -1, "iridescenceFactor": 0.0, "iridescenceIor": 1.3,
    "iridescenceThicknessMinimum": 100.0, "iridescenceThicknessMaximum": 400.0,
    "diffuseTransmissionColorFactor": [1, 1, 1], "diffuseTransmissionColorTextureId": -1, "dispersion": 0,
}


class MaterialData:
    """68-float material record with the reference's slots and defaults."""

    SIZE = 68

    def __init__(self, **kw):
        object.__setattr__(self, "data", np.zeros(self.SIZE, np.float32))
        object.__setattr__(self, "dirty", False)
        for name, _, _ in _MAT_FIELDS:
            if name.endswith("TextureId"):
                setattr(self, name, -1)
        for k, v in _MAT_DEFAULTS.items():
            setattr(self, k, v)
        for k, v in kw.items():
            setattr(self, k, v)

    def __setattr__(self, name, value):
        for fname, slot, width in _MAT_FIELDS:
            if fname == name:
                with np.errstate(over="ignore"):
                    if width == 1:
                        self.data[slot] = np.float32(float(value))
                    else:
                        self.data[slot:slot + width] = np.asarray(value, dtype=np.float64)[:width]
                object.__setattr__(self, "dirty", True)
                return
        object.__setattr__(self, name, value)

    def __getattr__(self, name):
        for fname, slot, width in _MAT_FIELDS:
            if fname == name:
                return float(self.data[slot]) if width == 1 else [float(x) for x in self.data[slot:slot + width]]
        raise AttributeError(name)


class TexInfo:
    """8-float texture record of material.ts:483-523: offset.xy, texture-array
    index, layer, scale.xy (the texture's repeat), uv set, pad."""

    def __init__(self):
        self.data = np.zeros(8, np.float32)
        self.texOffset = (0, 0)
        self.texArrayIdx = 255
        self.texIdx = 255
        self.texScale = (1, 1)
        self.uvSet = 255

    def __setattr__(self, name, value):
        slots = {"texOffset": (0, 2), "texArrayIdx": (2, 1), "texIdx": (3, 1), "texScale": (4, 2), "uvSet": (6, 1)}
        if name in slots:
            a, n = slots[name]
            self.data[a:a + n] = np.asarray(value, np.float64).reshape(-1)[:n] if n > 1 else np.float32(value)
            return
        object.__setattr__(self, name, value)


class Light:
    """Point light record of scene_data.ts:3-8."""

    def __init__(self, position=(1, 1, 1), emission=(1, 1, 1)):
        self.position = list(position)
        self.type = 0
        self.emission = list(emission)


class PathtracingSceneData:
    """Flat scene buffers handed to the renderer (scene_data.ts:23-132)."""

    def __init__(self):
        self.vertexBuffer = None         # f32[20*V]
        self.triangleIndexBuffer = None  # u32[3*T]
        self.bvhPositionBuffer = None    # f32[3*V]
        self.bvhIndexBuffer = None       # u32[3*T]
        self.materials = []
        self.lights = []
        self.texInfos = []
        self.texArrays = {}  # "w,h" -> list of (key, RGBA8 image h x w x 4), insertion ordered
        # engine extension: a prebuilt bvh_gpu node array + index array skips the build
        self.bvhNodes = None
        self.bvhTriangleIndices = None

    @property
    def num_triangles(self):
        if self.triangleIndexBuffer is not None:
            return len(self.triangleIndexBuffer) // 3
        return 0 if self.vertexBuffer is None else len(self.vertexBuffer) // (VERTEX_STRIDE * 3)

    @property
    def num_materials(self):
        return len(self.materials)

    def addMaterial(self, mat):
        self.materials.append(mat)
        return len(self.materials) - 1

    def addTexture(self, image, offset=(0, 0), repeat=(1, 1), uvSet=0, key=None):
        """scene_data.ts:71-101: images of one resolution share a texture array;
        returns the TexInfo id a material slot refers to. image: u8[h, w, 4]
        (row 0 first); key identifies an image already in its array."""
        img = np.ascontiguousarray(image, np.uint8)
        h, w = img.shape[0], img.shape[1]
        res = f"{w},{h}"
        info = TexInfo()
        if res in self.texArrays:
            arr = self.texArrays[res]
            idx = next((i for i, (k, _) in enumerate(arr) if key is not None and k == key), -1)
            if idx < 0:
                arr.append((key, img))
                idx = len(arr) - 1
            info.texArrayIdx = list(self.texArrays).index(res)
            info.texIdx = idx
        else:
            self.texArrays[res] = [(key, img)]
            info.texArrayIdx = len(self.texArrays) - 1
            info.texIdx = 0
        info.texOffset = offset
        info.texScale = repeat
        info.uvSet = uvSet
        self.texInfos.append(info)
        return len(self.texInfos) - 1

    @property
    def num_textures(self):
        return len(self.texInfos)

    def getFlatTextureInfoBuffer(self):
        if not self.texInfos:
            return np.zeros(0, np.float32)
        return np.concatenate([t.data for t in self.texInfos]).astype(np.float32)

    def getTextureArrays(self):
        """(dims i32[A, 3] = width, height, layers; texels u8, arrays concatenated)."""
        dims, parts = [], []
        for res, arr in self.texArrays.items():
            w, h = (int(x) for x in res.split(","))
            dims.append((w, h, len(arr)))
            parts += [img.reshape(-1) for _, img in arr]
        if not dims:
            return np.zeros((0, 3), np.int32), np.zeros(0, np.uint8)
        return np.asarray(dims, np.int32), np.concatenate(parts)

    def addLight(self, light):
        self.lights.append(light)
        return len(self.lights) - 1

    def getFlatMaterialBuffer(self):
        if not self.materials:
            return np.zeros(0, np.float32)
        return np.concatenate([m.data for m in self.materials]).astype(np.float32)

    def getPositionBuffer(self):
        v = np.asarray(self.vertexBuffer, np.float32).reshape(-1, VERTEX_STRIDE)
        return np.ascontiguousarray(v[:, :3]).reshape(-1)


def _normalize(v, fallback):
    length = math.hypot(v[0], v[1], v[2])
    if length <= 1e-10:
        return list(fallback)
    return [v[0] / length, v[1] / length, v[2] / length]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


class ViewportCamera:
    """Look-at camera, pathtracing_viewport.ts:68-98 (JS doubles stored into f32)."""

    def __init__(self, position=(0, 0, 6), target=(0, 0, 0), up=(0, 1, 0), fov=45.0, near=0.01):
        self.fov = float(fov)
        self.near = float(near)
        self.position = [0.0, 0.0, 6.0]
        self.matrixWorld = np.zeros(16, np.float32)
        self.setLookAt(position, target, up)

    def setLookAt(self, position, target, up=(0, 1, 0)):
        self.position = [float(x) for x in position]
        fwd = _normalize([target[i] - position[i] for i in range(3)], [0, 0, -1])
        right = _normalize(_cross(fwd, up), [1, 0, 0])
        cam_up = _normalize(_cross(right, fwd), [0, 1, 0])
        back = [-x for x in fwd]
        e = np.zeros(16, np.float64)
        e[0], e[4], e[8], e[12] = right[0], cam_up[0], back[0], position[0]
        e[1], e[5], e[9], e[13] = right[1], cam_up[1], back[1], position[1]
        e[2], e[6], e[10], e[14] = right[2], cam_up[2], back[2], position[2]
        e[3], e[7], e[11], e[15] = 0, 0, 0, 1
        self.matrixWorld = e.astype(np.float32)
