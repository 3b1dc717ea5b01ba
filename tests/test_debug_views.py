"""Pins the oracle's debug program (SH/integrator/debug.glsl:1-83) before it
checks the engine's: the renderer runs it whenever debugMode is not 'None'
(renderer.ts:1020) with u_debug_mode = the mode's index in debugModes
(renderer.ts:181, 406; -1 for a name not in the list).

Every sample of one frame is restated here in numpy from the reference's
shaders, not from the oracle's code: the camera ray (the oracle's camera ray
and generator, pinned in test_oracle_pins.py / test_oracle.py), its closest
hit (the oracle's traversal, pinned against brute force), and then
  hits    the material factors of the triangle's material (material.glsl,
          untextured: the factor itself), the interpolated uv0, the shading
          normal after fix_normals, the tangent of get_onb(n) rotated by the
          anisotropy angle atan(0, 1) + pi, and the bitangent of
          calculateBasis (bvh.glsl:101-127, utils.glsl:88-123, 183-190);
  misses  the environment, 10 x the pdf table or the cdf table, each a GL
          LINEAR + REPEAT bilinear fetch at dir_to_uv of the rotated
          direction (lighting.glsl:56-63, utils.glsl:141-147), with alpha 1
          (10 for the pdf view), or vec4(0) when the background is off.
The environment is per-texel noise, so a wrong texel index, half-texel
offset, axis swap or rotation sign fails; every material field the views
show holds its own random value per material, so a field read from the
wrong slot fails.
"""
import math

import numpy as np
import pytest

import oracle_ref as O

W, H = 48, 36
MODES = ["None", "Albedo", "Metalness", "Roughness", "Normals", "Tangents", "Bitangents", "Transparency", "UV0",
         "Clearcoat", "IBL PDF", "IBL CDF", "Specular", "SpecularTint", "Translucency"]


def debug_workload():
    """Sphere field (12 spheres on a ground quad) with per-material random
    values in every field a debug view shows and a 64 x 32 noise IBL."""
    from dspt import scenes, build_bvh
    wl = scenes.sphere_field(n_spheres=12, width=W, height=H, ibl_size=(64, 32))
    rng = np.random.default_rng(11)
    for m in wl.scene.materials:
        m.metallicFactor = rng.uniform(0.0, 1.0)
        m.transmissionFactor = rng.uniform(0.0, 1.0)
        m.clearcoatFactor = rng.uniform(0.0, 1.0)
        m.specularFactor = rng.uniform(0.0, 1.0)
        m.specularColorFactor = rng.uniform(0.0, 1.0, 3)
        m.diffuseTransmissionFactor = rng.uniform(0.0, 1.0)
        m.diffuseTransmissionColorFactor = rng.uniform(0.0, 1.0, 3)
    iw, ih = 64, 32
    env = np.ones((ih, iw, 4), np.float32)
    env[..., :3] = rng.uniform(0.05, 2.0, (ih, iw, 3)).astype(np.float32)
    wl.ibl = {"image": {"data": env.reshape(-1), "width": iw, "height": ih}}
    wl.scene.bvhNodes, wl.scene.bvhTriangleIndices = build_bvh(wl.scene.bvhPositionBuffer, wl.scene.bvhIndexBuffer)
    return wl


def debug_block(wl, total, mode, show_background=True, rot=0.7):
    from dspt import pack_frame_block
    s, img = wl.scene, wl.ibl["image"]
    idx = MODES.index(mode) if mode in MODES else -1
    return pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                            ibl_total_sum=total, ibl_rotation_rad=rot, show_background=show_background,
                            scene_counts=(s.num_triangles, len(s.bvhNodes) // 16, len(s.bvhTriangleIndices), 0),
                            debug_mode=idx)


def oracle_scene(wl):
    s = wl.scene
    return O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices,
                         s.getFlatMaterialBuffer(), wl.ibl, 0)


@pytest.fixture(scope="module")
def dbg():
    wl = debug_workload()
    osc = oracle_scene(wl)
    blk = debug_block(wl, osc.total, "Albedo")
    seed = 1 * wl.max_bounces  # rng_init(frame_count * max_bounces, px, py) of frame 1, renderer.frag:146
    org = np.zeros((H, W, 3), np.float32)
    dirs = np.zeros((H, W, 3), np.float32)
    for py in range(H):
        for px in range(W):
            f = O.rng_seq(seed, px, py, 2)[3]
            org[py, px], dirs[py, px] = O.camera_ray(blk, W, H, px, py, f[0], f[1])
    rays = np.zeros((H * W, 8), np.float32)
    rays[:, 0:3] = org.reshape(-1, 3)
    rays[:, 3] = 1e30
    rays[:, 4:7] = dirs.reshape(-1, 3)
    tri, t, u, v, _ = osc.trace(rays)
    return {"wl": wl, "osc": osc, "dir": dirs.reshape(-1, 3).astype(np.float64), "tri": tri, "u": u, "v": v}


def _render(d, mode, show_background=True):
    blk = debug_block(d["wl"], d["osc"].total, mode, show_background)
    o = d["osc"].render(blk, W, H, 1, 1, threads=8)
    return o["samples"].reshape(-1, 4), o["prim"].reshape(-1)


def _vertex_attr(d, first, width):
    """(a0, a1, a2) of the hit triangles' vertices, attribute at float `first`."""
    vb = np.asarray(d["wl"].scene.vertexBuffer, np.float32).reshape(-1, 20)
    tris = np.asarray(d["wl"].scene.triangleIndexBuffer, np.int64).reshape(-1, 3)[d["tri"][d["tri"] >= 0]]
    return [vb[tris[:, k], first:first + width].astype(np.float64) for k in range(3)]


def _bary(d):
    hit = d["tri"] >= 0
    u = d["u"][hit].astype(np.float64)[:, None]
    v = d["v"][hit].astype(np.float64)[:, None]
    return 1.0 - u - v, u, v


def _normalize(x):
    return x / np.linalg.norm(x, axis=1, keepdims=True)


def _frames(d):
    """(n, t, b) of calculateBasis after fix_normals and the anisotropy rotation."""
    w0, u, v = _bary(d)
    n0, n1, n2 = _vertex_attr(d, 4, 3)
    n = _normalize(w0 * n0 + u * n1 + v * n2)           # compute_interpolated_normal
    p0, p1, p2 = _vertex_attr(d, 0, 3)
    ng = _normalize(np.cross(p1 - p0, p2 - p0))          # compute_triangle_normal
    # vertex tangents are zero: get_onb(n)[0], handedness 1 (bvh.glsl:118-127)
    c = np.cross(n, np.stack([-n[:, 2], n[:, 0], -n[:, 1]], 1))
    ok = np.linalg.norm(c, axis=1) > 1e-3
    t = _normalize(c)
    wi = -d["dir"][d["tri"] >= 0]
    ng = np.where((np.sum(wi * ng, 1) < 0)[:, None], -ng, ng)       # fix_normals
    n = np.where((np.sum(ng * n, 1) < 0)[:, None], -n, n)
    ang = math.atan2(0.0, 1.0) + math.pi                             # anisotropyDirection (1, 0, 0)
    b = np.cross(n, t)
    t = t * math.cos(ang) + b * math.sin(ang)                        # rotation_to_tangent
    return n, t, np.cross(n, t), ok


def _bilinear(tab, w, h, uv):
    """texture() with GL LINEAR + REPEAT, rows from v = 0."""
    x = uv[:, 0] * w - 0.5
    y = uv[:, 1] * h - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    a, b = (x - x0)[:, None], (y - y0)[:, None]
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)
    t = tab.reshape(h, w, -1).astype(np.float64)
    g = lambda yy, xx: t[yy % h, xx % w]
    return (g(y0, x0) * (1 - a) + g(y0, x0 + 1) * a) * (1 - b) + (g(y0 + 1, x0) * (1 - a) + g(y0 + 1, x0 + 1) * a) * b


def _miss_uv(d, rot=0.7):
    dd = d["dir"][d["tri"] < 0]
    c, s = math.cos(rot), math.sin(rot)
    r = np.stack([c * dd[:, 0] - s * dd[:, 2], dd[:, 1], s * dd[:, 0] + c * dd[:, 2]], 1)  # rotate_dir_phi
    theta = np.arccos(np.clip(r[:, 1], -1, 1))
    phi = np.arctan2(r[:, 2], r[:, 0])
    phi = np.where(phi < 0, phi + 2 * math.pi, phi)
    return np.stack([phi / (2 * math.pi), theta / math.pi], 1)


def test_debug_rays_hit_and_miss(dbg):
    """The workload shows both hits and background; the program's first hit
    is the traversal's closest hit."""
    _, prim = _render(dbg, "Albedo")
    np.testing.assert_array_equal(prim, dbg["tri"])
    assert 0.2 < (prim >= 0).mean() < 0.9


@pytest.mark.parametrize("mode,field", [("Albedo", "baseColorFactor"), ("Metalness", "metallicFactor"),
                                        ("Roughness", "roughnessFactor"), ("Transparency", "transmissionFactor"),
                                        ("Clearcoat", "clearcoatFactor"), ("Specular", "specularFactor"),
                                        ("SpecularTint", "specularColorFactor"), ("Translucency", None)])
def test_material_views(dbg, mode, field):
    L, prim = _render(dbg, mode)
    hit = prim >= 0
    vb = np.asarray(dbg["wl"].scene.vertexBuffer, np.float32).reshape(-1, 20)
    tris = np.asarray(dbg["wl"].scene.triangleIndexBuffer, np.int64).reshape(-1, 3)
    mats = dbg["wl"].scene.materials
    mi = vb[tris[prim[hit], 0], 3].astype(np.int64)
    if field is None:  # diffuseTransmissionFactor * diffuseTransmissionColorFactor in f32
        want = np.array([np.float32(mats[k].diffuseTransmissionFactor) *
                         np.asarray(mats[k].diffuseTransmissionColorFactor, np.float32) for k in mi])
    else:
        want = np.array([np.broadcast_to(np.asarray(getattr(mats[k], field), np.float32), 3) for k in mi])
    if field == "baseColorFactor":
        # times the interpolated vertex colour (all ones, but interpolated in f32:
        # bvh.glsl:129-134, material.glsl:13-14)
        u = dbg["u"][hit][:, None]
        v = dbg["v"][hit][:, None]
        c = [a.astype(np.float32) for a in _vertex_attr(dbg, 16, 3)]
        want = want * ((np.float32(1) - u - v) * c[0] + u * c[1] + v * c[2])
        assert want.dtype == np.float32
    np.testing.assert_array_equal(L[hit, :3], want)
    np.testing.assert_array_equal(L[hit, 3], 1.0)
    assert len(np.unique(want[:, 0])) > 5  # many materials on screen


def test_uv0_view(dbg):
    L, prim = _render(dbg, "UV0")
    hit = prim >= 0
    w0, u, v = _bary(dbg)
    a0, a1, a2 = _vertex_attr(dbg, 8, 2)
    np.testing.assert_allclose(L[hit, :2], w0 * a0 + u * a1 + v * a2, atol=2e-6)
    assert (L[hit, 2] == 0).all() and (L[hit, 3] == 1).all()


@pytest.mark.parametrize("mode,k", [("Normals", 0), ("Tangents", 1), ("Bitangents", 2)])
def test_frame_views(dbg, mode, k):
    L, prim = _render(dbg, mode)
    hit = prim >= 0
    fr = _frames(dbg)
    ok = fr[3]
    assert ok.mean() > 0.99
    np.testing.assert_allclose(L[hit][ok, :3], fr[k][ok], atol=2e-5)
    assert (L[hit, 3] == 1).all()


def test_environment_views(dbg):
    """Misses: the environment under every mode but the two table views
    (here Albedo and an unknown name), equal to the path integrator's own
    primary miss (mode None); 10 x the pdf table (alpha 10) and the cdf
    table; hits of the table views and of the unknown name show vec4(0, 0, 0, 1)."""
    osc = dbg["osc"]
    img = dbg["wl"].ibl["image"]
    w, h = img["width"], img["height"]
    miss = dbg["tri"] < 0
    uv = _miss_uv(dbg)
    env = _bilinear(np.asarray(img["data"], np.float32).reshape(-1, 4)[:, :3], w, h, uv)
    Lp, _ = _render(dbg, "None")
    for mode in ("Albedo", "Wireframe"):
        L, _ = _render(dbg, mode)
        np.testing.assert_allclose(L[miss, :3], env, rtol=0, atol=3e-5)  # f32 uv: ~1e-5 texel x texel steps <= 2
        np.testing.assert_array_equal(L[miss], Lp[miss])
        assert (L[miss, 3] == 1).all()
        if mode == "Wireframe":
            np.testing.assert_array_equal(L[~miss], np.tile([0, 0, 0, 1], ((~miss).sum(), 1)))
    pdf, cdf, _, _, _ = O.ibl_cdf(img["data"], w, h)
    L, _ = _render(dbg, "IBL PDF")
    np.testing.assert_allclose(L[miss, 0], 10 * _bilinear(pdf, w, h, uv)[:, 0], rtol=0, atol=1e-5)
    assert (L[miss, 1:3] == 0).all() and (L[miss, 3] == 10).all()
    np.testing.assert_array_equal(L[~miss], np.tile([0, 0, 0, 1], ((~miss).sum(), 1)))
    L, _ = _render(dbg, "IBL CDF")
    np.testing.assert_allclose(L[miss, 0], _bilinear(cdf, w, h, uv)[:, 0], rtol=0, atol=2e-6)
    assert (L[miss, 1:3] == 0).all() and (L[miss, 3] == 1).all()


def test_background_off(dbg):
    """u_background_from_ibl = 0: misses are vec4(0) in every view; hits are unchanged."""
    L1, _ = _render(dbg, "Albedo")
    for mode in ("Albedo", "IBL PDF", "IBL CDF"):
        L, _ = _render(dbg, mode, show_background=False)
        miss = dbg["tri"] < 0
        assert (L[miss] == 0).all()
        if mode == "Albedo":
            np.testing.assert_array_equal(L[~miss], L1[~miss])
