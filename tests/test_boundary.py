"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
declared in include/dspt.h, reports errors through dspt_last_error(), and the
engine-side BVH builder emits a well-formed bvh_gpu layout."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dspt.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dspt_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    from dspt import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), f"missing export {s}"
    # the Python binding declares exactly the header's entry points
    assert sorted(_lib.SIGNATURES) == syms


def test_create_without_device_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from dspt import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.dspt_create(0, 0, ctypes.byref(h))
    assert rc == -2  # DSPT_ERR_DEVICE
    assert b"device" in L.dspt_last_error()
    assert L.dspt_create(7, 0, ctypes.byref(h)) == -1


def _check_bvh(nodes, bidx, pos, ind, max_leaf):
    N = len(nodes) // 16
    nd = nodes.reshape(N, 16)
    T = len(ind) // 3
    seen = np.zeros(T, np.int64)
    tri_pos = pos.reshape(-1, 3)[ind.reshape(-1, 3)]
    stack = [(0, None)]
    visited = np.zeros(N, bool)
    while stack:
        n, box = stack.pop()
        assert 0 <= n < N and not visited[n], "cycle or dangling child"
        visited[n] = True
        cnt, first = int(nd[n, 11]), int(nd[n, 15])
        if cnt > 0:
            assert cnt <= max_leaf
            tris = bidx[first:first + cnt]
            seen[tris] += 1
            if box is not None:
                tp = tri_pos[tris].reshape(-1, 3)
                assert (tp >= box[0] - 1e-6).all() and (tp <= box[1] + 1e-6).all(), "leaf outside parent box"
        else:
            l, r = int(nd[n, 3]), int(nd[n, 7])
            stack.append((l, (nd[n, 0:3], nd[n, 4:7])))
            stack.append((r, (nd[n, 8:11], nd[n, 12:15])))
    assert visited.all()
    assert (seen == 1).all(), "every triangle in exactly one leaf"
    assert sorted(bidx.tolist()) == list(range(T))


def test_bvh_builder_layout():
    from dspt import build_bvh, scenes
    wl = scenes.sphere_field(n_spheres=8)
    pos, ind = wl.scene.bvhPositionBuffer, wl.scene.bvhIndexBuffer
    for ml in (1, 4, 8):
        nodes, bidx = build_bvh(pos, ind, max_leaf=ml)
        _check_bvh(nodes, bidx, pos, ind, ml)


def test_bvh_builder_degenerate_and_single():
    from dspt import build_bvh
    pos = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    nodes, bidx = build_bvh(pos, np.array([0, 1, 2], np.uint32))
    assert len(nodes) == 16 and nodes[11] == 1 and list(bidx) == [0]  # root is a leaf
    # many coincident triangles: builder must still terminate with bounded leaves
    ind = np.tile(np.array([0, 1, 2], np.uint32), 100)
    nodes, bidx = build_bvh(pos, ind, max_leaf=4)
    _check_bvh(nodes, bidx, pos, ind, 4)


def test_bvh_builder_rejects_bad_index():
    from dspt import build_bvh, DsptError
    pos = np.zeros(9, np.float32)
    with pytest.raises(DsptError, match="out of range"):
        build_bvh(pos, np.array([0, 1, 5], np.uint32))
