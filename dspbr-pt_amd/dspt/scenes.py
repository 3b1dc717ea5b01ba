"""Seeded procedural workloads (synthetic inputs of the benchmark configurations).

Every generator returns a PathtracingSceneData whose buffers follow the
reference's layouts (20-float vertex stream, u32 indices, 68-float
materials), plus an IBL dict {'image': {'data', 'width', 'height'}} and a
camera. Draws use numpy's PCG64 with fixed seeds; the draw order is part of
the workload definition.

  sphere_field   64 UV spheres (32x16) over a 20x20 ground quad, 61,442 tris,
                 baseColor/metallic/roughness materials, sky + sun IBL (256^2, 16 spp, depth 8)
  full_materials 1024x1024 Perlin terrain + 256 displaced icospheres (subdiv 5),
                 7,335,938 tris, materials drawn from the clearcoat / sheen /
                 anisotropy / specular / transmission+volume / iridescence /
                 dispersion families, log-normal IBL + sun (1920x1080, 64 spp, depth 12)
  batch_scene    one scene of the uneven batched mix (spheres / terrain / soup)
"""
import math

import numpy as np

from .scene import MaterialData, PathtracingSceneData, ViewportCamera, VERTEX_STRIDE


# ----------------------------------------------------------------- geometry
def uv_sphere(radius, width_segments=32, height_segments=16):
    """Sphere with (ws+1)(hs+1) vertices and ws*(hs-1)*2 triangles (pole rows emit one triangle per quad)."""
    ws, hs = width_segments, height_segments
    u = np.arange(ws + 1) / ws
    v = np.arange(hs + 1) / hs
    uu, vv = np.meshgrid(u, v)
    phi = uu * 2.0 * math.pi
    theta = vv * math.pi
    nx = -np.cos(phi) * np.sin(theta)
    ny = np.cos(theta)
    nz = np.sin(phi) * np.sin(theta)
    n = np.stack([nx, ny, nz], -1).reshape(-1, 3)
    p = n * radius
    uv = np.stack([uu, 1.0 - vv], -1).reshape(-1, 2)
    idx = []
    grid = np.arange((hs + 1) * (ws + 1)).reshape(hs + 1, ws + 1)
    for iy in range(hs):
        for ix in range(ws):
            a, b = grid[iy, ix + 1], grid[iy, ix]
            c, d = grid[iy + 1, ix], grid[iy + 1, ix + 1]
            if iy != 0:
                idx.append((a, b, d))
            if iy != hs - 1:
                idx.append((b, c, d))
    return p, n, uv, np.array(idx, np.int64)


def icosphere(subdiv):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
             (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
             (8, 6, 7), (9, 8, 1)]
    v = np.array(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(faces, np.int64)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        es = np.sort(e, axis=1)
        uniq, inv = np.unique(es, axis=0, return_inverse=True)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        inv = inv.reshape(3, -1) + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = inv[0], inv[1], inv[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def vertex_normals(p, f):
    fn = np.cross(p[f[:, 1]] - p[f[:, 0]], p[f[:, 2]] - p[f[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, f[:, k], fn)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return n / np.maximum(ln, 1e-20)


def perlin2d(shape, scale, seed, octaves=4):
    """Fractal gradient noise on a regular grid (vectorised)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    h, w = shape
    out = np.zeros(shape, np.float64)
    amp, freq = 1.0, 1.0
    for _ in range(octaves):
        gsz = int(max(2, scale * freq)) + 2
        ang = rng.uniform(0, 2 * math.pi, (gsz, gsz))
        gx, gy = np.cos(ang), np.sin(ang)
        ys = np.linspace(0, scale * freq, h, endpoint=False)
        xs = np.linspace(0, scale * freq, w, endpoint=False)
        X, Y = np.meshgrid(xs, ys)
        x0, y0 = np.floor(X).astype(int), np.floor(Y).astype(int)
        fx, fy = X - x0, Y - y0

        def dotg(ix, iy, dx, dy):
            return gx[iy, ix] * dx + gy[iy, ix] * dy

        n00 = dotg(x0, y0, fx, fy)
        n10 = dotg(x0 + 1, y0, fx - 1, fy)
        n01 = dotg(x0, y0 + 1, fx, fy - 1)
        n11 = dotg(x0 + 1, y0 + 1, fx - 1, fy - 1)
        sx = fx * fx * fx * (fx * (fx * 6 - 15) + 10)
        sy = fy * fy * fy * (fy * (fy * 6 - 15) + 10)
        nx0 = n00 + sx * (n10 - n00)
        nx1 = n01 + sx * (n11 - n01)
        out += amp * (nx0 + sy * (nx1 - nx0))
        amp *= 0.5
        freq *= 2.0
    return out


class MeshBuilder:
    """Accumulates objects into the flat reference buffers."""

    def __init__(self):
        self.pos, self.nrm, self.uv, self.idx, self.mat = [], [], [], [], []
        self.nverts = 0

    def add(self, p, n, uv, f, mat_index):
        self.pos.append(np.asarray(p, np.float64))
        self.nrm.append(np.asarray(n, np.float64))
        self.uv.append(np.asarray(uv, np.float64) if uv is not None else np.zeros((len(p), 2)))
        self.idx.append(np.asarray(f, np.int64) + self.nverts)
        self.mat.append(np.full(len(p), mat_index, np.float64))
        self.nverts += len(p)

    def finish(self, scene):
        p = np.concatenate(self.pos)
        n = np.concatenate(self.nrm)
        uv = np.concatenate(self.uv)
        m = np.concatenate(self.mat)
        V = len(p)
        vb = np.zeros((V, VERTEX_STRIDE), np.float32)
        vb[:, 0:3] = p
        vb[:, 3] = m
        vb[:, 4:7] = n
        vb[:, 8:10] = uv
        vb[:, 16:20] = 1.0  # vertex colour default (scene_loader.ts:340-343)
        scene.vertexBuffer = vb.reshape(-1)
        tri = np.concatenate(self.idx).astype(np.uint32).reshape(-1)
        scene.triangleIndexBuffer = tri
        scene.bvhPositionBuffer = np.ascontiguousarray(vb[:, 0:3]).reshape(-1)
        scene.bvhIndexBuffer = tri
        return scene


# ----------------------------------------------------------------- IBLs
def equirect_dirs(w, h):
    x = (np.arange(w) + 0.5) / w
    y = (np.arange(h) + 0.5) / h
    phi = x[None, :] * 2 * math.pi
    theta = y[:, None] * math.pi
    return np.stack(np.broadcast_arrays(np.sin(theta) * np.cos(phi), np.cos(theta) + 0 * phi, np.sin(theta) * np.sin(phi)), -1)


def add_sun(img, dirs, elevation_deg=45.0, azimuth_deg=0.0, diameter_deg=2.0, radiance=5e4):
    el, az = math.radians(elevation_deg), math.radians(azimuth_deg)
    sun = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    cosang = dirs @ sun
    mask = cosang >= math.cos(math.radians(diameter_deg * 0.5))
    img[mask, :3] = radiance
    return img


def sky_ibl(w=512, h=256):
    d = equirect_dirs(w, h)
    up = np.clip(d[..., 1], -1, 1)
    t = np.clip(up, 0, 1)[..., None]
    zen, hor, gnd = np.array([0.25, 0.45, 1.0]), np.array([0.9, 0.9, 1.0]), np.array([0.3, 0.27, 0.22])
    sky = hor * (1 - t) + zen * t
    col = np.where(up[..., None] >= 0, sky, gnd * (0.5 + 0.5 * (1 + up[..., None])))
    img = np.concatenate([col, np.ones((h, w, 1))], -1)
    add_sun(img, d)
    return {"image": {"data": img.astype(np.float32).reshape(-1), "width": w, "height": h}}


def lognormal_ibl(w, h, seed, mu=0.0, sigma=2.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    lum = np.exp(rng.normal(mu, sigma, (h, w)))
    tint = rng.uniform(0.6, 1.0, (h, w, 3))
    img = np.concatenate([lum[..., None] * tint, np.ones((h, w, 1))], -1)
    add_sun(img, equirect_dirs(w, h))
    return {"image": {"data": img.astype(np.float32).reshape(-1), "width": w, "height": h}}


# ----------------------------------------------------------------- materials
def basic_material(rng):
    """baseColor U[0,1]^3, metallic Bernoulli(0.3), roughness U[0.05,1] (in that draw order)."""
    m = MaterialData()
    m.baseColorFactor = rng.uniform(0, 1, 3)
    m.metallicFactor = 1.0 if rng.uniform() < 0.3 else 0.0
    m.roughnessFactor = rng.uniform(0.05, 1.0)
    return m


FAMILIES = ["clearcoat", "sheen", "anisotropy", "specular", "transmission", "iridescence", "dispersion"]


def family_material(rng, family=None):
    """Material from one lobe family with seeded parameters (full profile)."""
    if family is None:
        family = FAMILIES[int(rng.integers(0, len(FAMILIES)))]
    m = basic_material(rng)
    if family == "clearcoat":
        m.clearcoatFactor = rng.uniform(0.5, 1.0)
        m.clearcoatRoughnessFactor = rng.uniform(0.0, 0.3)
    elif family == "sheen":
        m.metallicFactor = 0.0
        m.sheenColorFactor = rng.uniform(0, 1, 3)
        m.sheenRoughnessFactor = rng.uniform(0.2, 1.0)
    elif family == "anisotropy":
        m.metallicFactor = 1.0
        m.roughnessFactor = rng.uniform(0.1, 0.6)
        m.anisotropy = rng.uniform(0.3, 1.0)
        a = rng.uniform(0, 2 * math.pi)
        m.anisotropyDirection = [math.cos(a), math.sin(a), 0.0]
    elif family == "specular":
        m.metallicFactor = 0.0
        m.specularFactor = rng.uniform(0.0, 1.0)
        m.specularColorFactor = rng.uniform(0, 1, 3)
        m.ior = rng.uniform(1.3, 2.0)
    elif family == "transmission":
        m.metallicFactor = 0.0
        m.transmissionFactor = 1.0
        m.thinWalled = 0
        m.roughnessFactor = rng.uniform(0.0, 0.3)
        m.ior = rng.uniform(1.3, 1.8)
        m.attenuationDistance = rng.uniform(0.1, 2.0)
        m.attenuationColor = rng.uniform(0.3, 1.0, 3)
    elif family == "iridescence":
        m.iridescenceFactor = rng.uniform(0.5, 1.0)
        m.iridescenceIor = rng.uniform(1.2, 1.8)
        m.iridescenceThicknessMinimum = rng.uniform(100, 400)
        m.iridescenceThicknessMaximum = rng.uniform(400, 1200)
    elif family == "dispersion":
        m.metallicFactor = 0.0
        m.transmissionFactor = 1.0
        m.thinWalled = 0
        m.roughnessFactor = rng.uniform(0.0, 0.2)
        m.ior = rng.uniform(1.4, 1.8)
        m.dispersion = rng.uniform(0.5, 5.0)
        m.attenuationDistance = rng.uniform(0.1, 2.0)
        m.attenuationColor = rng.uniform(0.3, 1.0, 3)
    else:
        raise ValueError(family)
    return m


# ----------------------------------------------------------------- workloads
class Workload:
    def __init__(self, name, scene, ibl, camera, width, height, spp, max_bounces, profile):
        self.name, self.scene, self.ibl, self.camera = name, scene, ibl, camera
        self.width, self.height, self.spp, self.max_bounces, self.profile = width, height, spp, max_bounces, profile


def sphere_field(seed=1, n_spheres=64, width=256, height=256, spp=16, max_bounces=8, ibl_size=(512, 256)):
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = PathtracingSceneData()
    mb = MeshBuilder()
    gm = scene.addMaterial(basic_material(rng))
    gp = np.array([[-10, 0, -10], [10, 0, -10], [10, 0, 10], [-10, 0, 10]], np.float64)
    mb.add(gp, np.tile([0, 1, 0], (4, 1)), [[0, 0], [1, 0], [1, 1], [0, 1]], [[0, 2, 1], [0, 3, 2]], gm)
    sp, sn, suv, sf = uv_sphere(1.0)
    for _ in range(n_spheres):
        c = np.array([rng.uniform(-5, 5), rng.uniform(0, 3), rng.uniform(-5, 5)])
        r = rng.uniform(0.2, 0.6)
        mi = scene.addMaterial(basic_material(rng))
        mb.add(sp * r + c, sn, suv, sf, mi)
    mb.finish(scene)
    cam = ViewportCamera(position=(0, 3, 12), target=(0, 1, 0), fov=45, near=0.01)
    return Workload("sphere_field", scene, sky_ibl(*ibl_size), cam, width, height, spp, max_bounces, "webgl-lean")


def full_materials(seed=2, grid=1024, n_objects=256, subdiv=5, width=1920, height=1080, spp=64, max_bounces=12,
                   ibl_size=(2048, 1024)):
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = PathtracingSceneData()
    mb = MeshBuilder()
    extent = 40.0
    hgt = perlin2d((grid, grid), 6.0, seed + 1000) * 3.0
    xs = np.linspace(-extent, extent, grid)
    X, Z = np.meshgrid(xs, xs)
    P = np.stack([X, hgt, Z], -1)
    dx = np.gradient(hgt, xs, axis=1)
    dz = np.gradient(hgt, xs, axis=0)
    Nrm = np.stack([-dx, np.ones_like(dx), -dz], -1)
    Nrm /= np.linalg.norm(Nrm, axis=-1, keepdims=True)
    ii = np.arange(grid * grid).reshape(grid, grid)
    a, b, c, d = ii[:-1, :-1].ravel(), ii[:-1, 1:].ravel(), ii[1:, 1:].ravel(), ii[1:, :-1].ravel()
    tf = np.concatenate([np.stack([a, d, b], 1), np.stack([b, d, c], 1)])
    tm = scene.addMaterial(family_material(rng))
    mb.add(P.reshape(-1, 3), Nrm.reshape(-1, 3), None, tf, tm)
    iv, iff = icosphere(subdiv)
    for k in range(n_objects):
        cx, cz = rng.uniform(-30, 30), rng.uniform(-30, 30)
        r = rng.uniform(0.5, 2.5)
        gx = int(np.clip((cx + extent) / (2 * extent) * (grid - 1), 0, grid - 1))
        gz = int(np.clip((cz + extent) / (2 * extent) * (grid - 1), 0, grid - 1))
        cy = hgt[gz, gx] + r * rng.uniform(0.6, 1.4)
        amp = rng.uniform(0.02, 0.12) * r
        ph = rng.uniform(0, 2 * math.pi, 3)
        disp = 1.0 + amp / r * np.sin(5 * iv[:, 0] + ph[0]) * np.sin(5 * iv[:, 1] + ph[1]) * np.sin(5 * iv[:, 2] + ph[2])
        p = iv * (r * disp)[:, None] + np.array([cx, cy, cz])
        n = vertex_normals(p, iff)
        mi = scene.addMaterial(family_material(rng))
        mb.add(p, n, None, iff, mi)
    mb.finish(scene)
    cam = ViewportCamera(position=(0, 18, 45), target=(0, 0, 0), fov=45, near=0.01)
    return Workload("full_materials", scene, lognormal_ibl(*ibl_size, seed=seed + 2000), cam, width, height, spp,
                    max_bounces, "webgl-full")


def triangle_soup(rng, n, extent=5.0):
    c = rng.uniform(-extent, extent, (n, 3)) + np.array([0, extent, 0])
    e = rng.uniform(0.01, 0.5, n)
    d1 = rng.normal(size=(n, 3))
    d2 = rng.normal(size=(n, 3))
    d1 /= np.linalg.norm(d1, axis=1, keepdims=True)
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    p = np.stack([c, c + d1 * e[:, None], c + d2 * e[:, None]], 1).reshape(-1, 3)
    f = np.arange(3 * n).reshape(n, 3)
    nn = np.cross(d1, d2)
    nn /= np.maximum(np.linalg.norm(nn, axis=1, keepdims=True), 1e-12)
    return p, np.repeat(nn, 3, axis=0), f


def batch_scene(seed, n_tris, kind):
    """One scene of the uneven batched mix: kind in {'spheres', 'terrain', 'soup'}, about n_tris triangles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = PathtracingSceneData()
    mb = MeshBuilder()
    gm = scene.addMaterial(family_material(rng))
    gp = np.array([[-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20]], np.float64)
    mb.add(gp, np.tile([0, 1, 0], (4, 1)), None, [[0, 2, 1], [0, 3, 2]], gm)
    if kind == "spheres":
        sp, sn, suv, sf = uv_sphere(1.0)
        k = max(1, n_tris // len(sf))
        for _ in range(k):
            c = np.array([rng.uniform(-6, 6), rng.uniform(0, 4), rng.uniform(-6, 6)])
            r = rng.uniform(0.2, 0.8)
            mb.add(sp * r + c, sn, suv, sf, scene.addMaterial(family_material(rng)))
    elif kind == "terrain":
        g = max(8, int(math.sqrt(n_tris / 2)) + 1)
        hgt = perlin2d((g, g), 4.0, seed + 7) * 1.5
        xs = np.linspace(-8, 8, g)
        X, Z = np.meshgrid(xs, xs)
        P = np.stack([X, hgt, Z], -1).reshape(-1, 3)
        ii = np.arange(g * g).reshape(g, g)
        a, b, c, d = ii[:-1, :-1].ravel(), ii[:-1, 1:].ravel(), ii[1:, 1:].ravel(), ii[1:, :-1].ravel()
        tf = np.concatenate([np.stack([a, d, b], 1), np.stack([b, d, c], 1)])
        mb.add(P, vertex_normals(P, tf), None, tf, scene.addMaterial(family_material(rng)))
    elif kind == "soup":
        p, n, f = triangle_soup(rng, n_tris)
        mb.add(p, n, None, f, scene.addMaterial(family_material(rng)))
    else:
        raise ValueError(kind)
    return mb.finish(scene)


BATCH_CAMERAS = {
    "spheres": ((0.0, 3.0, 12.0), (0.0, 1.0, 0.0)),
    "terrain": ((0.0, 6.0, 14.0), (0.0, 0.0, 0.0)),
    "soup": ((0.0, 5.0, 18.0), (0.0, 5.0, 0.0)),
}


def batch_mix(seed=3, n_jobs=32, tri_range=(10_000, 2_000_000), side_range=(256, 1024), spp_range=(16, 256),
              ibl_size=(1024, 512), max_bounces=8, profile="webgl-full"):
    """The uneven batched mix: n_jobs scenes with log-uniform triangle counts
    in tri_range, a kind drawn from {spheres, terrain, soup}, a square
    resolution with log-uniform pixel count (side in side_range, a multiple
    of 16), log-uniform spp in spp_range and a seeded log-normal IBL each.
    Draw order per job: n_tris, kind, side, spp, scene seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = ["spheres", "terrain", "soup"]
    out = []
    for k in range(n_jobs):
        n_tris = int(round(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))))
        kind = kinds[int(rng.integers(0, 3))]
        side = math.exp(rng.uniform(math.log(side_range[0]), math.log(side_range[1])))
        side = int(min(side_range[1], max(side_range[0], 16 * round(side / 16))))
        spp = int(round(math.exp(rng.uniform(math.log(spp_range[0]), math.log(spp_range[1])))))
        sseed = int(rng.integers(0, 2**31 - 1))
        scene = batch_scene(sseed, n_tris, kind)
        pos, tgt = BATCH_CAMERAS[kind]
        cam = ViewportCamera(position=pos, target=tgt, fov=45, near=0.01)
        ibl = lognormal_ibl(*ibl_size, seed=sseed + 1)
        out.append(Workload(f"batch{k}_{kind}", scene, ibl, cam, side, side, spp, max_bounces, profile))
    return out


def _checker(w, h, n, c0, c1):
    y, x = np.mgrid[0:h, 0:w]
    m = ((x * n // w + y * n // h) % 2).astype(bool)
    img = np.where(m[..., None], np.asarray(c1, np.uint8), np.asarray(c0, np.uint8))
    return np.ascontiguousarray(img.astype(np.uint8))


def _noise_rgba(rng, w, h):
    return rng.integers(0, 256, (h, w, 4), dtype=np.uint8)


def _bump_normal_map(w, h, freq=4.0, amp=0.6):
    y, x = np.mgrid[0:h, 0:w] / np.array([h, w])[:, None, None]
    dx = amp * np.cos(2 * math.pi * freq * x) * np.sin(2 * math.pi * freq * y)
    dy = amp * np.sin(2 * math.pi * freq * x) * np.cos(2 * math.pi * freq * y)
    n = np.stack([-dx, -dy, np.ones_like(dx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    rgb = np.clip(np.round((n * 0.5 + 0.5) * 255), 0, 255).astype(np.uint8)
    return np.concatenate([rgb, np.full((h, w, 1), 255, np.uint8)], -1)


def textured_spheres(seed=6, n_spheres=16, width=256, height=256, spp=16, max_bounces=8, ibl_size=(512, 256),
                     profile="webgl-full"):
    """Sphere field whose materials use every texture slot of MaterialData
    (material.glsl:5-110): checker / noise / normal-map RGBA8 textures in
    three resolutions (so several texture arrays with several layers), with
    offsets and repeats; the ground repeats its textures across the quad."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = PathtracingSceneData()
    mb = MeshBuilder()
    checker = _checker(64, 64, 8, (230, 230, 230, 255), (40, 60, 160, 255))
    checker2 = _checker(64, 64, 4, (200, 40, 30, 255), (250, 220, 120, 200))
    normal = _bump_normal_map(64, 64)
    noise = _noise_rgba(rng, 32, 32)
    noise2 = _noise_rgba(rng, 32, 32)
    stripes = np.zeros((96, 48, 4), np.uint8)
    stripes[..., :] = 255
    stripes[::8, :, :3] = 30
    t_checker = scene.addTexture(checker, key="checker")
    t_checker_rep = scene.addTexture(checker, offset=(0.25, 0.1), repeat=(2, 3), key="checker")
    t_checker2 = scene.addTexture(checker2, key="checker2")
    t_normal = scene.addTexture(normal, key="normal")
    t_noise = scene.addTexture(noise, key="noise")
    t_noise2 = scene.addTexture(noise2, repeat=(3, 3), key="noise2")
    t_stripes = scene.addTexture(stripes, key="stripes")
    g = basic_material(rng)
    g.baseColorTextureId = t_checker_rep
    g.metallicRoughnessTextureId = t_noise
    g.normalTextureId = t_normal
    gm = scene.addMaterial(g)
    gp = np.array([[-10, 0, -10], [10, 0, -10], [10, 0, 10], [-10, 0, 10]], np.float64)
    mb.add(gp, np.tile([0, 1, 0], (4, 1)), [[0, 0], [4, 0], [4, 4], [0, 4]], [[0, 2, 1], [0, 3, 2]], gm)
    slots = [
        dict(baseColorTextureId=t_checker2, emissiveTextureId=t_stripes, emissiveFactor=[0.6, 0.5, 0.3]),
        dict(metallicRoughnessTextureId=t_noise2, normalTextureId=t_normal, baseColorTextureId=t_checker),
        dict(clearcoatFactor=0.8, clearcoatTextureId=t_noise, clearcoatRoughnessTextureId=t_noise2),
        dict(sheenColorFactor=[0.8, 0.7, 0.9], sheenRoughnessFactor=0.6, sheenColorTextureId=t_checker2,
             sheenRoughnessTextureId=t_noise),
        dict(specularFactor=0.8, specularTextureId=t_noise2, specularColorTextureId=t_checker),
        dict(transmissionFactor=0.9, transmissionTextureId=t_stripes, thinWalled=0, attenuationDistance=1.0),
        dict(diffuseTransmissionFactor=0.7, diffuseTransmissionTextureId=t_noise,
             diffuseTransmissionColorTextureId=t_checker2),
        dict(anisotropy=0.7, anisotropyTextureId=t_noise, roughnessFactor=0.4),
        dict(anisotropy=0.6, anisotropyTextureId=t_noise2, anisotropyDirectionTextureId=t_normal, roughnessFactor=0.3),
        dict(iridescenceFactor=0.8, iridescenceTextureId=t_noise, iridescenceThicknessTextureId=t_stripes),
        dict(normalTextureId=t_normal, normalScale=0.5, metallicFactor=1.0, roughnessFactor=0.2),
    ]
    sp, sn, suv, sf = uv_sphere(1.0)
    for k in range(n_spheres):
        c = np.array([rng.uniform(-5, 5), rng.uniform(0, 3), rng.uniform(-5, 5)])
        r = rng.uniform(0.3, 0.8)
        m = basic_material(rng)
        for name, v in slots[k % len(slots)].items():
            setattr(m, name, v)
        mb.add(sp * r + c, sn, suv, sf, scene.addMaterial(m))
    mb.finish(scene)
    cam = ViewportCamera(position=(0, 3, 12), target=(0, 1, 0), fov=45, near=0.01)
    return Workload("textured_spheres", scene, sky_ibl(*ibl_size), cam, width, height, spp, max_bounces, profile)


# ----------------------------------------------------------------- instancing
def _rock(n_tris, radius, rng):
    """Displaced UV sphere with about n_tris triangles (closed surface)."""
    hs = max(4, int(round(math.sqrt(n_tris / 4.0))))
    p, n, uv, f = uv_sphere(1.0, 2 * hs, hs)
    ph = rng.uniform(0, 2 * math.pi, 3)
    amp = rng.uniform(0.05, 0.2)
    disp = 1.0 + amp * np.sin(3 * p[:, 0] + ph[0]) * np.sin(3 * p[:, 1] + ph[1]) * np.sin(3 * p[:, 2] + ph[2])
    p = p * (radius * disp)[:, None]
    return p, vertex_normals(p, f), uv, f


def _torus(n_tris, radius, rng):
    """Torus (major radius `radius`, minor U[0.15, 0.4] of it) with about n_tris triangles."""
    v = max(4, int(round(math.sqrt(n_tris / 4.0))))
    u = 2 * v
    rr = radius * rng.uniform(0.15, 0.4)
    a = np.arange(u) / u * 2 * math.pi
    b = np.arange(v) / v * 2 * math.pi
    A, B = np.meshgrid(a, b, indexing="ij")
    ring = np.stack([np.cos(A), np.zeros_like(A), np.sin(A)], -1)
    n = ring * np.cos(B)[..., None] + np.array([0.0, 1.0, 0.0]) * np.sin(B)[..., None]
    p = ring * radius + n * rr
    ii = np.arange(u * v).reshape(u, v)
    i0, i1 = ii, np.roll(ii, -1, 0)
    j0, j1 = i0, np.roll(i0, -1, 1)
    k0, k1 = i1, np.roll(i1, -1, 1)
    f = np.concatenate([np.stack([j0, j1, k1], -1).reshape(-1, 3), np.stack([j0, k1, k0], -1).reshape(-1, 3)])
    uv = np.stack([A / (2 * math.pi), B / (2 * math.pi)], -1).reshape(-1, 2)
    return p.reshape(-1, 3), n.reshape(-1, 3), uv, f


def random_rotation(rng):
    """Uniform rotation matrix from a normalised Gaussian quaternion."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


class InstancedScene:
    """Object-space meshes placed by rigid instances (engine extension).

    scene         PathtracingSceneData: object-space vertex stream, u32
                  indices, materials (one per mesh here)
    mesh_first / mesh_count   i64 triangle range of each mesh
    mesh_vfirst / mesh_vcount vertex range of each mesh (for flatten)
    o2w           f32[I, 3, 4] object -> world affine of each instance
    inst_mesh     i32[I] mesh of each instance

    flatten() bakes it into one world-space scene the way the reference's
    loader bakes a glTF node hierarchy (scene_loader.ts:173-222, 244-345):
    positions by the world matrix, normals and tangents by its normal
    matrix (the inverse transpose) and renormalised, instance by instance;
    that is the oracle's view of an instanced scene.
    """

    def __init__(self):
        self.scene = PathtracingSceneData()
        self.mesh_first, self.mesh_count, self.mesh_vfirst, self.mesh_vcount = [], [], [], []
        self.o2w = np.zeros((0, 3, 4), np.float32)
        self.inst_mesh = np.zeros(0, np.int32)

    @property
    def num_instances(self):
        return len(self.inst_mesh)

    @property
    def instanced_triangles(self):
        return int(np.asarray(self.mesh_count, np.int64)[self.inst_mesh].sum())

    def world_boxes(self):
        """f64[I, 2, 3]: world box of each instance (its mesh's object box,
        corners transformed)."""
        p = self.scene.vertexBuffer.reshape(-1, VERTEX_STRIDE)[:, 0:3].astype(np.float64)
        out = np.zeros((len(self.inst_mesh), 2, 3))
        mb = []
        for m in range(len(self.mesh_vfirst)):
            q = p[self.mesh_vfirst[m]:self.mesh_vfirst[m] + self.mesh_vcount[m]]
            lo, hi = q.min(0), q.max(0)
            mb.append(np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]))
        for i, m in enumerate(self.inst_mesh):
            M = self.o2w[i].astype(np.float64)
            c = mb[m] @ M[:, :3].T + M[:, 3]
            out[i, 0], out[i, 1] = c.min(0), c.max(0)
        return out

    def engine_args(self):
        s = self.scene
        return (s.vertexBuffer, s.triangleIndexBuffer, np.asarray(self.mesh_first, np.int64),
                np.asarray(self.mesh_count, np.int64), self.o2w, self.inst_mesh, s.getFlatMaterialBuffer())

    def flatten(self, instances=None):
        """(PathtracingSceneData in world space, inst_of_tri i32[T'], mesh_tri_of_tri i32[T']).
        instances: ids of the instances to bake (default all), in that order."""
        s = self.scene
        vb = s.vertexBuffer.reshape(-1, VERTEX_STRIDE)
        tri = s.triangleIndexBuffer.reshape(-1, 3).astype(np.int64)
        out_v, out_t, inst_of, mtri_of = [], [], [], []
        nv = 0
        ids = range(len(self.inst_mesh)) if instances is None else [int(i) for i in instances]
        for i in ids:
            m = self.inst_mesh[i]
            M = self.o2w[i].astype(np.float64)
            R, tr = M[:, :3], M[:, 3]
            Nm = np.linalg.inv(R).T
            v0, vc = self.mesh_vfirst[m], self.mesh_vcount[m]
            t0, tc = self.mesh_first[m], self.mesh_count[m]
            v = vb[v0:v0 + vc].astype(np.float64).copy()
            v[:, 0:3] = v[:, 0:3] @ R.T + tr
            for a, b in ((4, 7), (12, 15)):
                d = v[:, a:b] @ Nm.T
                ln = np.linalg.norm(d, axis=1, keepdims=True)
                v[:, a:b] = np.where(ln > 1e-8, d / np.maximum(ln, 1e-30), d)
            out_v.append(v.astype(np.float32))
            out_t.append(tri[t0:t0 + tc] - v0 + nv)
            inst_of.append(np.full(tc, i, np.int32))
            mtri_of.append(np.arange(t0, t0 + tc, dtype=np.int32))
            nv += vc
        f = PathtracingSceneData()
        f.materials = list(s.materials)
        V = np.concatenate(out_v)
        f.vertexBuffer = V.reshape(-1)
        f.triangleIndexBuffer = np.concatenate(out_t).astype(np.uint32).reshape(-1)
        f.bvhPositionBuffer = np.ascontiguousarray(V[:, 0:3]).reshape(-1)
        f.bvhIndexBuffer = f.triangleIndexBuffer
        return f, np.concatenate(inst_of), np.concatenate(mtri_of)


def instanced_scene(seed=5, n_meshes=16, n_instances=4096, tri_range=(4_000, 64_000), extent=40.0, height=6.0,
                    width=3840, height_px=2160, spp=256, max_bounces=16, ibl_size=(8192, 4096),
                    families=("transmission", "dispersion")):
    """Large instanced scene: n_meshes base meshes (displaced spheres and
    tori, log-uniform triangle counts in tri_range, radius U[0.4, 1.6]) each
    with one transmission+volume or dispersion material, placed by
    n_instances seeded rigid transforms (uniform rotation, translation in
    [-extent, extent] x [0, height] x [-extent, extent]); log-normal IBL +
    sun. Draw order: per mesh (n_tris, kind, radius, shape draws, family,
    material), then per instance (mesh, rotation, translation), then the
    IBL seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = InstancedScene()
    s = out.scene
    mb = MeshBuilder()
    nt = 0
    for m in range(n_meshes):
        n_tris = int(round(math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))))
        kind = int(rng.integers(0, 2))
        radius = rng.uniform(0.4, 1.6)
        p, n, uv, f = (_rock if kind == 0 else _torus)(n_tris, radius, rng)
        fam = families[int(rng.integers(0, len(families)))]
        mi = s.addMaterial(family_material(rng, fam))
        out.mesh_vfirst.append(mb.nverts)
        out.mesh_vcount.append(len(p))
        out.mesh_first.append(nt)
        out.mesh_count.append(len(f))
        nt += len(f)
        mb.add(p, n, uv, f, mi)
    mb.finish(s)
    I = n_instances
    o2w = np.zeros((I, 3, 4), np.float64)
    im = np.zeros(I, np.int32)
    for i in range(I):
        im[i] = int(rng.integers(0, n_meshes))
        o2w[i, :, :3] = random_rotation(rng)
        o2w[i, :, 3] = [rng.uniform(-extent, extent), rng.uniform(0, height), rng.uniform(-extent, extent)]
    out.o2w = o2w.astype(np.float32)
    out.inst_mesh = im
    iseed = int(rng.integers(0, 2**31 - 1))
    cam = ViewportCamera(position=(0, 14, extent * 1.15), target=(0, 0, 0), fov=45, near=0.01)
    ibl = lognormal_ibl(*ibl_size, seed=iseed)
    return Workload("instanced", out, ibl, cam, width, height_px, spp, max_bounces, "webgl-full")
