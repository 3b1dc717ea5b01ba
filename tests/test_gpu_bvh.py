"""GPU BVH builds (linear BVH and PLOC over Morton codes, dspt_build_bvh_gpu)
against what a bvh_gpu tree must satisfy, and traversal over them.

Builder parity with the reference's TinyBVH is unpinned by design (the tree
is an input both traversals share); pinned here: the node array is a
well-formed bvh_gpu tree rooted at 0 (every triangle in exactly one leaf,
leaves of at most max_leaf triangles, every child box in a parent contains
the boxes of all triangles below that child, inner nodes reference valid
children); traversal over it is bit-exact with the oracle traversing the
same arrays; closest hits equal those over the host-built SAH tree (the
tree changes the visit order, not the nearest hit, except exact ties)."""
import numpy as np
import pytest

import oracle_ref as O

pytestmark = pytest.mark.gpu


def _check_tree(nodes, bidx, verts, tris, T, max_leaf):
    N = nodes.size // 16
    nd = nodes.reshape(N, 16)
    p = verts.reshape(-1, 20)[:, :3].astype(np.float64)
    tri = tris.reshape(-1, 3)
    tlo = np.minimum(np.minimum(p[tri[:, 0]], p[tri[:, 1]]), p[tri[:, 2]])
    thi = np.maximum(np.maximum(p[tri[:, 0]], p[tri[:, 1]]), p[tri[:, 2]])
    assert sorted(bidx.tolist()) == list(range(T))
    covered = np.zeros(T, np.int32)
    box = {}  # node -> (lo, hi) of its triangles
    order, stack, seen = [], [0], set()
    while stack:
        n = stack.pop()
        assert 0 <= n < N and n not in seen
        seen.add(n)
        order.append(n)
        cnt = int(nd[n, 11])
        if cnt > 0:
            continue
        stack += [int(nd[n, 3]), int(nd[n, 7])]
    for n in reversed(order):
        cnt = int(nd[n, 11])
        if cnt > 0:
            first = int(nd[n, 15])
            assert 1 <= cnt <= max_leaf and 0 <= first and first + cnt <= T
            covered[first:first + cnt] += 1
            ids = bidx[first:first + cnt]
            box[n] = (tlo[ids].min(0), thi[ids].max(0))
        else:
            l, r = int(nd[n, 3]), int(nd[n, 7])
            for c, lo_s, hi_s in ((l, slice(0, 3), slice(4, 7)), (r, slice(8, 11), slice(12, 15))):
                lo, hi = box[c]
                assert np.all(nd[n, lo_s] <= lo) and np.all(nd[n, hi_s] >= hi), f"node {n}: child {c} box"
            box[n] = (np.minimum(box[l][0], box[r][0]), np.maximum(box[l][1], box[r][1]))
    assert np.all(covered == 1)
    return N, len(order)


def _soup(n, seed, dup=0):
    from dspt.scene import VERTEX_STRIDE
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3, 3, (n, 3))
    if dup:
        c[: n // 2] = c[0]  # many triangles with one centroid (equal Morton keys)
    p = np.repeat(c, 3, axis=0) + rng.normal(scale=0.05, size=(3 * n, 3))
    v = np.zeros((3 * n, VERTEX_STRIDE), np.float32)
    v[:, :3] = p
    return v.reshape(-1), np.arange(3 * n, dtype=np.uint32)


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("case", ["sphere_field", "soup", "dups", "n1", "n2", "n7"])
def test_gpu_tree_is_well_formed(case, method, sphere_field_small):
    from dspt import Engine
    eng = Engine(0)
    if case == "sphere_field":
        s = sphere_field_small.scene
        verts, tris = s.vertexBuffer, s.triangleIndexBuffer
    else:
        n = {"soup": 20000, "dups": 5000, "n1": 1, "n2": 2, "n7": 7}[case]
        verts, tris = _soup(n, 3, dup=(case == "dups"))
    T = tris.size // 3
    for max_leaf in (1, 4):
        nodes, bidx, ms = eng.build_bvh_gpu(verts, tris, max_leaf, method)
        N, reached = _check_tree(nodes, bidx, verts, tris, T, max_leaf)
        assert reached == N  # no unreferenced slots
        assert ms > 0


def _sah_cost(nodes, ct=1.0, ci=1.0):
    """Surface-area cost of a bvh_gpu tree relative to its root box."""
    nd = nodes.reshape(-1, 16).astype(np.float64)

    def area(lo, hi):
        d = np.maximum(hi - lo, 0)
        return 2 * (d[..., 0] * d[..., 1] + d[..., 1] * d[..., 2] + d[..., 2] * d[..., 0])
    inner = nd[:, 11] == 0
    # every child box is stored in its parent: sum child areas over inner nodes
    la, ra = area(nd[:, 0:3], nd[:, 4:7]), area(nd[:, 8:11], nd[:, 12:15])
    cnt = nd[:, 11]
    leaf_area = area(nd[:, 0:3], nd[:, 4:7])
    root = area(np.minimum(nd[0, 0:3], nd[0, 8:11]), np.maximum(nd[0, 4:7], nd[0, 12:15])) if inner[0] else leaf_area[0]
    return (ct * (la[inner] + ra[inner]).sum() + ci * (leaf_area[~inner] * cnt[~inner]).sum()) / root


def test_ploc_tree_beats_linear_tree(sphere_field_small):
    """Clustering by surface area gives a lower SAH cost than the Morton
    split of the linear BVH (Meister & Bittner's result), close to the host
    binned SAH build."""
    from dspt import Engine
    s = sphere_field_small.scene
    eng = Engine(0)
    lb, _, _ = eng.build_bvh_gpu(s.vertexBuffer, s.triangleIndexBuffer, 4, 0)
    pl, _, _ = eng.build_bvh_gpu(s.vertexBuffer, s.triangleIndexBuffer, 4, 1)
    c_lb, c_pl, c_host = _sah_cost(lb), _sah_cost(pl), _sah_cost(s.bvhNodes)
    print(f"SAH cost: linear {c_lb:.1f}, ploc {c_pl:.1f}, host {c_host:.1f}")
    assert c_pl < c_lb and c_pl < 1.25 * c_host


@pytest.mark.parametrize("method", [0, 1])
def test_traversal_over_gpu_tree(sphere_field_small, method):
    from dspt import Engine, build_bvh
    from test_gpu_parity import _camera_rays, _surface_rays
    wl = sphere_field_small
    s = wl.scene
    eng = Engine(0)
    nodes, bidx, _ = eng.build_bvh_gpu(s.vertexBuffer, s.triangleIndexBuffer, 4, method)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    rays = np.concatenate([_camera_rays(wl, 20000, 1), _surface_rays(wl, 20000, 2)])
    tri, t, u, v = eng.trace_closest(rays)
    osc = O.OracleScene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    otri, ot, ou, ov = osc.trace(rays)[:4]
    np.testing.assert_array_equal(tri, otri)
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))
    # the same nearest hits as over the host-built SAH tree
    eng2 = Engine(0)
    eng2.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, s.bvhNodes, s.bvhTriangleIndices, s.getFlatMaterialBuffer())
    tri2, t2, _, _ = eng2.trace_closest(rays)
    assert (tri == tri2).mean() >= 0.9999
    same = tri == tri2
    np.testing.assert_array_equal(t[same], t2[same])
    # scene 0 built on the device: the build is deterministic, so the same tree and the same hits
    eng3 = Engine(0)
    ms = eng3.upload_scene_gpu_bvh(s.vertexBuffer, s.triangleIndexBuffer, s.getFlatMaterialBuffer(), 4, method)
    assert ms > 0
    tri3, t3, _, _ = eng3.trace_closest(rays)
    np.testing.assert_array_equal(tri3, tri)
    np.testing.assert_array_equal(t3.view(np.uint32), t.view(np.uint32))
    sid, ms = eng3.scene_create_gpu_bvh(s.vertexBuffer, s.triangleIndexBuffer, s.getFlatMaterialBuffer(), 4, method)
    assert sid >= 1 and ms > 0
