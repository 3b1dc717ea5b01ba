"""The reference's tiled progressive schedule (renderer.ts:567-714) through
dspt_render_region: each frame is rendered as scissored tile passes in the
reference's tile order (settling 2x2 for the first frame, then tileRes x
tileRes), every pass folding its frame into the pixels of its rectangle. The
accumulation must equal, bit for bit, the oracle's rectangle renders of the
same sequence, overlapping ceil-rounded rectangles of odd sizes included
(an overlapped pixel takes its frame twice, as the reference's scissored
draw + copy does); a tiling without overlaps must equal whole-frame
rendering. The Python and Node drop-ins with tileSchedule 'tiles' make the
same calls and the reference's callbacks (one per tile, one per frame)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_ref as O  # noqa: F401
from dspt.renderer import tile_order, tile_rect
from test_gpu_parity import _oracle_scene, _upload

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _schedule(W, H, frames, tile_res):
    """(frame, rect) of the reference's loop: frame 1 settles in 2x2 tiles,
    later frames accumulate in tile_res x tile_res tiles (renderer.ts:567-571)."""
    out = []
    for f in range(1, frames + 1):
        tr = 2 if f == 1 else max(2, tile_res)
        for t in tile_order(tr):
            r = tile_rect(t, tr, (W, H))
            if r is not None:
                out.append((f, r))
    return out


def _block(wl, W, H, total):
    from dspt import pack_frame_block
    s = wl.scene
    img = wl.ibl["image"]
    return pack_frame_block(wl.camera, W, H, max_bounces=wl.max_bounces, ibl_size=(img["width"], img["height"]),
                            ibl_total_sum=total,
                            scene_counts=(s.num_triangles, len(s.bvhNodes) // 16, len(s.bvhTriangleIndices), 0))


def _engine_tiled(wl, W, H, frames, tile_res):
    from dspt import Engine
    eng = Engine(0)
    _upload(eng, wl)
    img = wl.ibl["image"]
    total = eng.upload_ibl(img["data"], img["width"], img["height"])
    blk = _block(wl, W, H, total)
    eng.resize(W, H)
    eng.set_frame(blk)
    eng.clear_accum()  # no render target yet: nothing to clear
    for f, (x, y, w, h) in _schedule(W, H, frames, tile_res):
        eng.render_region(f, 1, x, y, w, h)
    acc = eng.read_accum()
    return eng, blk, acc


@pytest.mark.parametrize("W,H,tile_res", [(64, 48, 4), (67, 45, 4), (33, 7, 3), (5, 3, 4)])
def test_tiled_schedule_matches_oracle(sphere_field_small, W, H, tile_res):
    wl = sphere_field_small
    eng, blk, acc = _engine_tiled(wl, W, H, 3, tile_res)
    eng.close()
    osc = _oracle_scene(wl, 0)
    ref = np.zeros(W * H * 4, np.float32)
    for f, (x, y, w, h) in _schedule(W, H, 3, tile_res):
        osc.render(blk, W, H, f, 1, rect=(x, y, x + w, y + h), accum=ref, record=False)
    ref = ref.reshape(H, W, 4)
    print(f"{W}x{H} tileRes {tile_res}: {len(_schedule(W, H, 3, tile_res))} tile passes")
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_tiles_without_overlap_equal_whole_frames(sphere_field_small):
    wl = sphere_field_small
    W, H, F = 64, 48, 3  # divisible by 2 and 4: the rectangles tile the image exactly
    eng, blk, acc = _engine_tiled(wl, W, H, F, 4)
    eng.render(1, F)  # whole frames restart the accumulation
    whole = eng.read_accum()
    eng.close()
    np.testing.assert_array_equal(acc.view(np.uint32), whole.view(np.uint32))


def test_region_arguments_and_clear(sphere_field_small):
    from dspt import DsptError
    wl = sphere_field_small
    W, H = 32, 24
    eng, blk, acc = _engine_tiled(wl, W, H, 1, 2)
    assert np.abs(acc).max() > 0
    for bad in [(-1, 0, 4, 4), (0, -1, 4, 4), (0, 0, 0, 4), (0, 0, 4, 0), (30, 0, 4, 4), (0, 22, 4, 4)]:
        with pytest.raises(DsptError):
            eng.render_region(2, 1, *bad)
    # a region keeps the pixels outside it
    before = eng.read_accum()
    eng.render_region(2, 1, 8, 4, 10, 6)
    after = eng.read_accum()
    outside = np.ones((H, W), bool)
    outside[4:10, 8:18] = False
    np.testing.assert_array_equal(after[outside].view(np.uint32), before[outside].view(np.uint32))
    assert (after[~outside] != before[~outside]).any()
    eng.clear_accum()
    assert not eng.read_accum().any()
    eng.close()


def test_python_renderer_tile_schedule(sphere_field_small):
    """PathtracingRenderer(tileSchedule='tiles'): one tileFinishedCB per tile
    (4 while settling, tileRes^2 after), frameFinishedCB after each frame's
    last tile, and the engine-level schedule's accumulation."""
    from dspt.renderer import PathtracingRenderer
    wl = sphere_field_small
    W, H, N = 67, 45, 4
    r = PathtracingRenderer("webgl-lean", width=W, height=H, tileSchedule="tiles")
    r.maxBounces = wl.max_bounces
    r.setScene(wl.scene)
    r.setIBL(wl.ibl)
    tiles, per_frame, frames = [0], [], []

    def on_tile():
        tiles[0] += 1

    def on_frame(n):
        frames.append(n)
        per_frame.append(tiles[0])
        tiles[0] = 0
    r.render(wl.camera, N, on_tile, on_frame)
    assert frames == [2, 3, 4] and per_frame == [4, 16, 16]
    acc = r.readAccumulation()
    eng, blk, ref = _engine_tiled(wl, W, H, N - 1, 4)
    eng.close()
    np.testing.assert_array_equal(acc.view(np.uint32), ref.view(np.uint32))


def test_node_tile_schedule(sphere_field_small, tmp_path):
    addon = os.path.join(ROOT, "addon", "build", "dspt.node")
    if not os.path.exists(addon):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "addon")], check=True)
    wl = sphere_field_small
    s = wl.scene
    W, H, N = 67, 45, 4
    img = wl.ibl["image"]
    np.asarray(s.vertexBuffer, np.float32).tofile(tmp_path / "verts.f32")
    np.asarray(s.triangleIndexBuffer, np.uint32).tofile(tmp_path / "tris.u32")
    np.asarray(s.bvhPositionBuffer, np.float32).tofile(tmp_path / "bvhpos.f32")
    s.getFlatMaterialBuffer().astype(np.float32).tofile(tmp_path / "mats.f32")
    np.asarray(img["data"], np.float32).tofile(tmp_path / "ibl.f32")
    meta = {"profile": "webgl-lean", "width": W, "height": H, "maxBounces": wl.max_bounces, "iblWidth": img["width"],
            "iblHeight": img["height"], "fov": wl.camera.fov, "near": wl.camera.near,
            "matrixWorld": [float(x) for x in wl.camera.matrixWorld], "position": dict(zip("xyz", wl.camera.position)),
            "numSamples": N, "tileSchedule": "tiles", "tileRes": 4}
    (tmp_path / "meta.json").write_text(json.dumps(meta))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node_render.js"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["sampleCount"] == N - 1 and info["tilesPerFrame"] == [4, 16, 16]
    node_acc = np.fromfile(tmp_path / "accum.f32", np.float32).reshape(H, W, 4)
    # the scene's own BVH is built by the addon with the same host builder
    eng, blk, ref = _engine_tiled(wl, W, H, N - 1, 4)
    eng.close()
    np.testing.assert_array_equal(node_acc.view(np.uint32), ref.view(np.uint32))
