"""GPU parity with the oracle on the edge cases of the path, and at full size.

  * malformed and adversarial trees (SH/bvh.glsl:137-212): cycles, self-links,
    out-of-range children, leaf ranges past the index array, invalid triangle
    ids, the 4096-step cap, and comb-shaped trees whose stack runs past the
    engine's 16 LDS entries (global spill) and past the reference's 64-entry
    capacity (pushes dropped) -- hits, t/u/v bits, occlusion and the
    node/triangle counters, through both the ray API and the wavefront
    renderer;
  * rays the arithmetic has to survive: NaN origins, zero and denormal
    direction components, zero and negative tfar;
  * ragged and tiny images (1x1, 1xH, Wx1, odd sizes) and bounce limits 0-2;
  * the 256x256 sphere field (webgl-lean, depth 8) checked whole.
The full-size full-material frame is checked in test_gpu_converged.py.
"""
import numpy as np
import pytest

import oracle_ref as O
from test_gpu_parity import _camera_rays, _check_samples, _surface_rays

pytestmark = pytest.mark.gpu


def _toy_ibl(w=64, h=32, seed=3):
    rng = np.random.default_rng(seed)
    img = np.ones((h, w, 4), np.float32)
    img[..., :3] = rng.lognormal(0, 1.0, (h, w, 3)).astype(np.float32)
    return {"image": {"data": img.reshape(-1), "width": w, "height": h}}


def _tri_buffers(P):
    """Vertex stream (20 f32 per vertex, material 0, face normals) for explicit triangles P[T, 3, 3]."""
    from dspt import MaterialData
    T = len(P)
    verts = np.zeros((3 * T, 20), np.float32)
    verts[:, 0:3] = P.reshape(-1, 3)
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    verts[:, 4:7] = np.repeat(n, 3, axis=0)
    verts[:, 16:20] = 1
    return verts.reshape(-1), np.arange(3 * T, dtype=np.uint32), MaterialData().data


def _planes(T, rng, ascending=False):
    """T large triangles parallel to z = 0 covering x, y in [0, 1], at random
    depths (ascending: triangle k + 1 lies above triangle k)."""
    z = rng.uniform(0.0, 1.0, T).astype(np.float32)
    if ascending:
        z = np.sort(z)
    P = np.zeros((T, 3, 3), np.float32)
    P[:, 0] = np.stack([np.full(T, -1.0), np.full(T, -1.0), z], 1)
    P[:, 1] = np.stack([np.full(T, 3.0), np.full(T, -1.0), z], 1)
    P[:, 2] = np.stack([np.full(T, -1.0), np.full(T, 3.0), z], 1)
    return P


def _comb_tree(D, rng, boxes):
    """D inner nodes in a chain: node i has a one-triangle leaf (node D + i)
    on the left and node i + 1 on the right (the last node has the same leaf
    on both sides). Equal child boxes make the leaf the far child, pushed at
    every level, so the stack grows by one per level."""
    nodes = np.zeros((2 * D, 16), np.float32)
    for i in range(D):
        if boxes == "same":
            lo_l, hi_l = [-1, -1, -1], [3, 3, 3]
            lo_r, hi_r = lo_l, hi_l
        else:
            lo_l = rng.uniform(-1.0, 0.6, 3)
            hi_l = lo_l + rng.uniform(0.3, 3.0, 3)
            lo_r = rng.uniform(-1.0, 0.6, 3)
            hi_r = lo_r + rng.uniform(0.3, 3.0, 3)
        left = D + i
        right = i + 1 if i + 1 < D else D + i
        nodes[i] = [*lo_l, left, *hi_l, right, *lo_r, 0, *hi_r, 0]
        nodes[D + i, 11] = 1  # count
        nodes[D + i, 15] = i  # first
    return nodes.reshape(-1), np.arange(D, dtype=np.int32)


def _down_rays(n, rng, tfar=1e5):
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0] = rng.uniform(0, 1, n)
    rays[:, 1] = rng.uniform(0, 1, n)
    rays[:, 2] = 5.0
    d = np.stack([rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n), -np.ones(n)], 1)
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = tfar
    return rays


def _compare_trace(eng, osc, rays):
    eng.reset_counters()
    tri, t, u, v = eng.trace_closest(rays)
    otri, ot, ou, ov, ocnt = osc.trace(rays)
    np.testing.assert_array_equal(tri, otri)
    hit = tri >= 0
    for a, b in ((t, ot), (u, ou), (v, ov)):
        np.testing.assert_array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32))
    c = eng.counters()
    assert (c["node_visits"], c["tri_tests"]) == (ocnt[0], ocnt[1])
    occ = eng.trace_occluded(rays)
    np.testing.assert_array_equal(occ, (otri >= 0).astype(np.int32))
    return otri, ocnt


@pytest.mark.parametrize("boxes", ["same", "random"])
def test_comb_trees_spill_and_overflow_the_stack(boxes):
    from dspt import Engine
    rng = np.random.default_rng(11)
    for D in (10, 20, 70, 200):
        P = _planes(D, rng, ascending=True)  # the deepest leaf holds the nearest plane
        verts, tris, mats = _tri_buffers(P)
        nodes, bidx = _comb_tree(D, rng, boxes)
        eng = Engine(0)
        eng.upload_scene(verts, tris, nodes, bidx, mats)
        osc = O.OracleScene(verts, tris, nodes, bidx, mats)
        rays = _down_rays(4000, rng)
        otri, ocnt = _compare_trace(eng, osc, rays)
        if boxes == "same":
            hit = otri >= 0  # rays past the planes' hypotenuse miss
            assert hit.mean() > 0.9
            if D < 64:  # every leaf is tested (the last one twice) and the nearest plane wins
                assert (otri[hit] == D - 1).all() and ocnt[1] == len(rays) * (D + 1)
            else:  # the 64-entry stack dropped pushes: the deepest leaves were never reached
                assert ocnt[1] == len(rays) * 64 and (otri[hit] < D - 1).all()
        eng.close()


def test_malformed_trees_match_oracle():
    from dspt import Engine
    rng = np.random.default_rng(12)
    T = 6
    P = _planes(T, rng)
    verts, tris, mats = _tri_buffers(P)
    big = [-10, -10, -10, 0, 10, 10, 10, 0]

    def inner(left, right):
        n = np.zeros(16, np.float32)
        n[0:3], n[4:7], n[8:11], n[12:15] = big[0:3], big[4:7], big[0:3], big[4:7]
        n[3], n[7] = left, right
        return n

    def leaf(first, count):
        n = np.zeros(16, np.float32)
        n[11], n[15] = count, first
        return n

    cases = {
        # node 0 -> 1 and itself (skipped), node 1 -> 0 (cycle) and 99 (out of range): the step cap ends it
        "cycle": ([inner(1, 0), inner(0, 99)], np.arange(T, dtype=np.int32)),
        # children out of range on both sides, negative child
        "bad_children": ([inner(-1, 7), inner(2, 3)], np.arange(T, dtype=np.int32)),
        # leaves whose ranges run past the index array, start before it, or count above 64
        "leaf_ranges": ([inner(1, 2), inner(3, 4), inner(5, 6), leaf(4, 5), leaf(-3, 5), leaf(0, 200), leaf(50, 2),
                         leaf(2, 1)], np.arange(T, dtype=np.int32)),
        # index array entries that are not triangles
        "bad_tri_ids": ([inner(1, 2), leaf(0, 4), leaf(2, 4)], np.array([0, -5, T, 3, 1 << 20, 5], np.int32)),
        # a root that is a leaf, and a one-node tree of an inner node pointing nowhere
        "root_leaf": ([leaf(0, T)], np.arange(T, dtype=np.int32)),
        "lone_inner": ([inner(5, 9)], np.arange(T, dtype=np.int32)),
        # a leaf child reached both as a leaf code (well-formed) and a self-cycle deeper down
        "mixed": ([inner(1, 2), inner(3, 1), leaf(1, 2), leaf(0, 1)], np.arange(T, dtype=np.int32)),
        # inner nodes after leaves, so the engine's copy (inner nodes packed
        # first) relabels every node: a leaf read by index (range starts
        # before the array), an unreferenced leaf, child fields 2.7 and 4.9
        # (read as 2 and 4) and -0.5 (read as 0: a cycle back to the root)
        "permuted": ([inner(1, 2.7), leaf(0, 2), inner(3, 5), leaf(-3, 5), leaf(2, 3), inner(4.9, -0.5)],
                     np.arange(T, dtype=np.int32)),
    }
    for name, (nl, bidx) in cases.items():
        nodes = np.concatenate(nl).astype(np.float32)
        eng = Engine(0)
        eng.upload_scene(verts, tris, nodes, bidx, mats)
        osc = O.OracleScene(verts, tris, nodes, bidx, mats)
        rays = _down_rays(512, rng)
        _, ocnt = _compare_trace(eng, osc, rays)
        if name == "cycle":
            assert ocnt[0] == 4096 * len(rays)  # every ray ran into the step cap
        eng.close()


def _render_vs_oracle(wl_scene, ibl, cam, W, H, frames, profile, max_bounces):
    from dspt import Engine, pack_frame_block
    verts, tris, nodes, bidx, mats = wl_scene
    eng = Engine(profile)
    eng.upload_scene(verts, tris, nodes, bidx, mats)
    img = ibl["image"]
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    osc = O.OracleScene(verts, tris, nodes, bidx, mats, ibl, profile)
    assert tot == osc.total
    blk = pack_frame_block(cam, W, H, max_bounces=max_bounces, ibl_size=(img["width"], img["height"]), ibl_total_sum=tot)
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.set_record_samples(True)
    eng.reset_counters()
    eng.render(1, frames)
    g_s, g_sig, g_prim = eng.read_samples(frames)
    g_acc = eng.read_accum()
    c = eng.counters()
    o = osc.render(blk, W, H, 1, frames, threads=8)
    _check_samples(g_s, g_sig, g_prim, o)
    assert c["paths"] == W * H * frames and c["rays_primary"] == W * H * frames
    eng.close()
    return g_acc, o


def test_wavefront_render_over_comb_tree():
    """Camera rays through the streaming kernels over a 100-deep comb (stack
    spill to global memory and dropped pushes in k_trace_ext / k_trace_shadow)."""
    from dspt.scene import ViewportCamera
    rng = np.random.default_rng(13)
    D = 100
    P = _planes(D, rng)
    verts, tris, mats = _tri_buffers(P)
    nodes, bidx = _comb_tree(D, rng, "same")
    cam = ViewportCamera(position=(0.5, 0.5, 5.0), target=(0.5, 0.5, 0.0), fov=30.0)
    _render_vs_oracle((verts, tris, nodes, bidx, mats), _toy_ibl(), cam, 24, 16, 2, 0, 4)


def test_edge_rays_bit_exact(sphere_field_small):
    from dspt import Engine
    wl = sphere_field_small
    s = wl.scene
    eng = Engine(0)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer())
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer())
    rays = np.concatenate([_camera_rays(wl, 3000, 21), _surface_rays(wl, 3000, 22)])
    n = len(rays)
    k = n // 10
    rays[0:k, 0] = np.nan                                   # NaN origin
    rays[k:2 * k, 4:7] = 0.0                                # zero direction
    rays[2 * k:3 * k, 4] = 0.0                              # one zero component
    rays[2 * k:3 * k, 5] = np.float32(1e-40)                # denormal component
    rays[3 * k:4 * k, 3] = 0.0                              # tfar 0
    rays[4 * k:5 * k, 3] = -1.0                             # negative tfar
    rays[5 * k:6 * k, 3] = np.float32(3.0e38)               # huge tfar
    rays[6 * k:7 * k, 4:7] *= np.float32(1e-3)              # unnormalised direction
    _compare_trace(eng, osc, rays)
    eng.close()


@pytest.mark.parametrize("W,H", [(1, 1), (3, 7), (17, 5), (64, 1), (1, 33)])
def test_tiny_and_ragged_images(sphere_field_small, W, H):
    wl = sphere_field_small
    s = wl.scene
    _render_vs_oracle((s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer()),
                      wl.ibl, wl.camera, W, H, 3, 0, wl.max_bounces)


@pytest.mark.parametrize("max_bounces", [0, 1, 2])
def test_bounce_limits(sphere_field_small, max_bounces):
    wl = sphere_field_small
    s = wl.scene
    g_acc, o = _render_vs_oracle(
        (s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer()), wl.ibl,
        wl.camera, 32, 24, 2, 0, max_bounces)
    if max_bounces == 0:
        # no bounce is taken: a primary hit returns the zero radiance it starts with
        hit = o["prim"][0] >= 0
        assert hit.any() and (o["samples"][0][hit][:, :3] == 0).all()


def test_sphere_field_256_whole_frame():
    """The 256x256, 16-spp, depth-8 lean workload: two whole frames."""
    from dspt import build_bvh, scenes
    wl = scenes.sphere_field()
    s = wl.scene
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    assert (wl.width, wl.height, wl.max_bounces) == (256, 256, 8)
    g_acc, o = _render_vs_oracle((s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer()), wl.ibl,
                                 wl.camera, wl.width, wl.height, 2, 0, wl.max_bounces)
    rel = np.abs(g_acc - o["accum"]) / np.maximum(np.abs(o["accum"]), 1e-3)
    assert np.median(rel) < 1e-6


def test_fast_slab_test_equals_exact_form(sphere_field_small):
    """Scenes whose inner-node boxes are all finite run the closest-hit
    kernel's slab test on hardware min/max; one non-finite box anywhere in
    the array selects the exact GLSL select form for the whole scene. An
    unreachable inner node with a NaN corner switches the same scene to the
    exact form: samples, accumulation and counters must not change."""
    from dspt import Engine, pack_frame_block
    wl = sphere_field_small
    s = wl.scene
    nodes = np.asarray(s.bvhNodes, np.float32).reshape(-1, 16)
    bad = np.zeros((1, 16), np.float32)
    bad[0, 0] = np.nan
    bad[0, 3], bad[0, 7] = -1, -1  # an inner node without children, referenced by none
    img = wl.ibl["image"]
    out = []
    for nd in (nodes, np.concatenate([nodes, bad])):
        eng = Engine(0)
        eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nd.reshape(-1), s.bvhTriangleIndices,
                         s.getFlatMaterialBuffer())
        tot = eng.upload_ibl(img["data"], img["width"], img["height"])
        blk = pack_frame_block(wl.camera, 96, 64, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                               ibl_total_sum=tot)
        eng.resize(96, 64)
        eng.set_frame(blk)
        eng.set_record_samples(True)
        eng.reset_counters()
        eng.render(1, 3)
        c = eng.counters()
        out.append((eng.read_accum(), *eng.read_samples(3), c["node_visits"], c["tri_tests"], c["rays_extension"]))
        eng.close()
    for a, b in zip(out[0], out[1]):
        if isinstance(a, np.ndarray):
            np.testing.assert_array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                          b.view(np.uint32) if b.dtype == np.float32 else b)
        else:
            assert a == b
