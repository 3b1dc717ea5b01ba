"""The N-API addon + Node PathtracingRenderer drop-in (addon/): loads on the
CPU and reports engine errors as JS Errors; on the GPU it renders the same
image as the engine driven from Python (same BVH builder, same frames)."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "addon", "build", "dspt.node")


def _node(script, *args, timeout=120):
    return subprocess.run(["node", "-e", script, *args], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@pytest.fixture(scope="module")
def addon_built():
    if not os.path.exists(ADDON):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "addon")], check=True)
    return ADDON


def test_addon_exports_and_error_channel(addon_built):
    r = _node("const d=require('./addon/build/dspt.node'); console.log(JSON.stringify(Object.keys(d)));"
              "try { d.create(9, 0) } catch (e) { console.log(e.message) }")
    assert r.returncode == 0, r.stderr
    keys = json.loads(r.stdout.splitlines()[0])
    assert {"create", "uploadScene", "uploadIbl", "resize", "setFrame", "render", "readAccum", "buildBvh"} <= set(keys)
    assert "dspt error -" in r.stdout.splitlines()[1]


def test_addon_bvh_builder_matches_python(addon_built, tmp_path):
    from dspt import build_bvh, scenes
    wl = scenes.sphere_field(n_spheres=4)
    pos = np.asarray(wl.scene.bvhPositionBuffer, np.float32)
    ind = np.asarray(wl.scene.bvhIndexBuffer, np.uint32)
    pos.tofile(tmp_path / "p.f32")
    ind.tofile(tmp_path / "i.u32")
    r = _node("const fs=require('fs'); const d=require('./addon/build/dspt.node');"
              "const rd=(f,T)=>{const b=fs.readFileSync(f); return new T(b.buffer.slice(b.byteOffset,b.byteOffset+b.byteLength))};"
              "const o=d.buildBvh(rd(process.argv[1],Float32Array), rd(process.argv[2],Uint32Array), 4);"
              "fs.writeFileSync(process.argv[3], Buffer.from(o.nodeData.buffer));"
              "fs.writeFileSync(process.argv[4], Buffer.from(o.triangleIndices.buffer));",
              str(tmp_path / "p.f32"), str(tmp_path / "i.u32"), str(tmp_path / "n.f32"), str(tmp_path / "b.i32"))
    assert r.returncode == 0, r.stderr
    nodes, bidx = build_bvh(pos, ind)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "n.f32", np.float32), nodes)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "b.i32", np.int32), bidx)


@pytest.mark.gpu
def test_node_render_matches_python_engine(addon_built, tmp_path):
    from dspt import Engine, build_bvh, pack_frame_block, scenes
    wl = scenes.sphere_field(n_spheres=16)
    s = wl.scene
    W, H, F = 64, 48, 3
    img = wl.ibl["image"]
    np.asarray(s.vertexBuffer, np.float32).tofile(tmp_path / "verts.f32")
    np.asarray(s.triangleIndexBuffer, np.uint32).tofile(tmp_path / "tris.u32")
    np.asarray(s.bvhPositionBuffer, np.float32).tofile(tmp_path / "bvhpos.f32")
    s.getFlatMaterialBuffer().astype(np.float32).tofile(tmp_path / "mats.f32")
    np.asarray(img["data"], np.float32).tofile(tmp_path / "ibl.f32")
    meta = {"profile": "webgl-lean", "width": W, "height": H, "maxBounces": 8, "iblWidth": img["width"],
            "iblHeight": img["height"], "fov": wl.camera.fov, "near": wl.camera.near,
            "matrixWorld": [float(x) for x in wl.camera.matrixWorld],
            "position": dict(zip("xyz", wl.camera.position)), "numSamples": F + 1, "framesPerCall": 2}
    (tmp_path / "meta.json").write_text(json.dumps(meta))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node_render.js"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["sampleCount"] == F  # frames 1 .. numSamples-1, as the reference's loop
    node_acc = np.fromfile(tmp_path / "accum.f32", np.float32).reshape(H, W, 4)
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    eng = Engine(0)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    eng.resize(W, H)
    eng.set_frame(pack_frame_block(wl.camera, W, H, max_bounces=8, ibl_size=(img["width"], img["height"]),
                                   ibl_total_sum=tot, scene_counts=(s.num_triangles, len(nodes) // 16, len(bidx), 0)))
    eng.render(1, F)
    np.testing.assert_array_equal(eng.read_accum().view(np.uint32), node_acc.view(np.uint32))
    # readDisplay: ACES + gamma + FXAA at the display size; saveImage: the same canvas, top row first
    from dspt.image import read_png
    disp = np.fromfile(tmp_path / "display.u8", np.uint8).reshape(H, W, 4)
    np.testing.assert_array_equal(disp, eng.postprocess(3, 1.0, True, True))
    np.testing.assert_array_equal(read_png(str(tmp_path / "display.png")), disp[::-1])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["plain", "instanced", "gpu_bvh"])
def test_node_render_server_matches_python_jobs(addon_built, tmp_path, kind):
    """Requests served by the Node RenderServer (continuous batching, staggered
    arrivals) produce the same images as the same jobs run from Python, for a
    one-level and a two-level (instanced) scene."""
    from dspt import Engine, PROFILE_LEAN, build_bvh, pack_frame_block, scenes
    inst = None
    if kind in ("plain", "gpu_bvh"):
        wl = scenes.sphere_field(n_spheres=16)
        s = wl.scene
    else:
        wl = scenes.instanced_scene(seed=7, n_meshes=3, n_instances=40, tri_range=(300, 2000), extent=6.0, height=3.0,
                                    ibl_size=(128, 64))
        inst, s = wl.scene, wl.scene.scene
        np.asarray(inst.mesh_first, np.int32).tofile(tmp_path / "meshfirst.i32")
        np.asarray(inst.mesh_count, np.int32).tofile(tmp_path / "meshcount.i32")
        np.asarray(inst.o2w, np.float32).tofile(tmp_path / "o2w.f32")
        np.asarray(inst.inst_mesh, np.int32).tofile(tmp_path / "instmesh.i32")
    img = wl.ibl["image"]
    np.asarray(s.vertexBuffer, np.float32).tofile(tmp_path / "verts.f32")
    np.asarray(s.triangleIndexBuffer, np.uint32).tofile(tmp_path / "tris.u32")
    np.asarray(s.bvhPositionBuffer, np.float32).tofile(tmp_path / "bvhpos.f32")
    s.getFlatMaterialBuffer().astype(np.float32).tofile(tmp_path / "mats.f32")
    np.asarray(img["data"], np.float32).tofile(tmp_path / "ibl.f32")
    reqs = [{"width": 64, "height": 48, "spp": 5, "delayMs": 0}, {"width": 40, "height": 40, "spp": 3, "delayMs": 5},
            {"width": 96, "height": 32, "spp": 4, "delayMs": 10}]
    meta = {"profile": "webgl-lean", "policy": "rr", "maxBounces": 6, "iblWidth": img["width"], "iblHeight": img["height"],
            "fov": wl.camera.fov, "near": wl.camera.near, "matrixWorld": [float(x) for x in wl.camera.matrixWorld],
            "position": dict(zip("xyz", wl.camera.position)), "requests": reqs, "instanced": inst is not None,
            "gpuBvh": kind == "gpu_bvh"}
    (tmp_path / "meta.json").write_text(json.dumps(meta))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node_serve.js"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["frames"] == [q["spp"] for q in reqs] and info["progress"] == info["frames"]
    eng = Engine(PROFILE_LEAN)
    if kind == "gpu_bvh":
        sid, _ = eng.scene_create_gpu_bvh(s.vertexBuffer, s.triangleIndexBuffer, s.getFlatMaterialBuffer())
        counts = (s.num_triangles, 0, s.num_triangles, 0)
    elif inst is None:
        nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
        sid = eng.scene_create(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
        counts = (s.num_triangles, len(nodes) // 16, len(bidx), 0)
    else:
        sid = eng.scene_create_instanced(*inst.engine_args())
        counts = (s.num_triangles, 0, 0, 0)
    iid, tot = eng.ibl_create(img["data"], img["width"], img["height"])
    for i, q in enumerate(reqs):
        blk = pack_frame_block(wl.camera, q["width"], q["height"], max_bounces=6, ibl_size=(img["width"], img["height"]),
                               ibl_total_sum=tot, scene_counts=counts)
        j = eng.job_create(sid, iid, blk, q["width"], q["height"], q["spp"])
        eng.jobs_run([j])
        ref = eng.job_read_accum(j)
        got = np.fromfile(tmp_path / f"accum{i}.f32", np.float32).reshape(q["height"], q["width"], 4)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.gpu
def test_node_viewport_call_sequence(addon_built, tmp_path):
    """The viewport's calls through the Node drop-in (tests/node_viewport.js):
    start() with no limit renders until stop() and no frame callback follows
    stop(); a camera move + resetAccumulation() restarts at frame 1 with the
    new view; start() during a render stops it; a sample limit ends the loop
    with renderingFinishedCB; interaction mode gives one tile per frame; a
    debugMode renders the debug program with u_debug_mode = its index in
    debugModes, -1 for an unknown name. Every accumulation equals the Python
    engine's render of the same frames with the same camera and mode, bit for
    bit."""
    from dspt import Engine, build_bvh, pack_frame_block, scenes
    from dspt.scene import ViewportCamera
    wl = scenes.sphere_field(n_spheres=16)
    s = wl.scene
    W, H = 64, 48
    img = wl.ibl["image"]
    np.asarray(s.vertexBuffer, np.float32).tofile(tmp_path / "verts.f32")
    np.asarray(s.triangleIndexBuffer, np.uint32).tofile(tmp_path / "tris.u32")
    np.asarray(s.bvhPositionBuffer, np.float32).tofile(tmp_path / "bvhpos.f32")
    s.getFlatMaterialBuffer().astype(np.float32).tofile(tmp_path / "mats.f32")
    np.asarray(img["data"], np.float32).tofile(tmp_path / "ibl.f32")
    cam2 = ViewportCamera(position=(4.0, 2.5, 9.0), target=(0.0, 0.5, 0.0), fov=wl.camera.fov)
    cams = [wl.camera, cam2]
    meta = {"profile": "webgl-lean", "width": W, "height": H, "maxBounces": 6, "iblWidth": img["width"],
            "iblHeight": img["height"], "framesBeforeStop": 12, "sampleLimit": 5,
            "cameras": [{"fov": c.fov, "near": c.near, "matrixWorld": [float(x) for x in c.matrixWorld],
                         "position": dict(zip("xyz", map(float, c.position)))} for c in cams]}
    (tmp_path / "meta.json").write_text(json.dumps(meta))
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "node_viewport.js"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    print(info)
    un, mv, lim = info["unbounded"], info["moved"], info["limited"]
    assert un["sampleCount"] >= 12 and un["mode"] == "stopped" and un["finished"] == []
    assert un["frameEvents"] == list(range(2, len(un["frameEvents"]) + 2))
    assert mv["sampleCount"] >= 4
    assert lim["sampleCount"] == 4 and lim["finished"] == [4] and lim["frameEvents"] == [2, 3, 4, 5]
    assert info["interactive"]["frames"] == 5 and info["interactive"]["tiles"] == 5
    for name in ("debug_normals", "debug_unknown"):
        assert info[name]["sampleCount"] == 4 and info[name]["finished"] == [4], name
    assert info["lists"]["debugModes"] == 15
    assert info["lists"]["tonemappingModes"] == ["None", "Reinhard", "Cineon", "AcesFilm"]
    nodes, bidx = build_bvh(s.bvhPositionBuffer, s.bvhIndexBuffer)
    eng = Engine(0)
    eng.upload_scene(s.vertexBuffer, s.triangleIndexBuffer, nodes, bidx, s.getFlatMaterialBuffer())
    tot = eng.upload_ibl(img["data"], img["width"], img["height"])
    eng.resize(W, H)
    for name, cam, n, dbg in (("unbounded", cams[0], un["sampleCount"], 0), ("moved", cams[1], mv["sampleCount"], 0),
                              ("limited", cams[1], lim["sampleCount"], 0), ("debug_normals", cams[1], 4, 4),
                              ("debug_unknown", cams[1], 4, -1)):
        eng.set_frame(pack_frame_block(cam, W, H, max_bounces=6, ibl_size=(img["width"], img["height"]),
                                       ibl_total_sum=tot, scene_counts=(s.num_triangles, len(nodes) // 16, len(bidx), 0),
                                       debug_mode=dbg))
        eng.render(1, n)
        got = np.fromfile(tmp_path / f"{name}.f32", np.float32).reshape(H, W, 4)
        np.testing.assert_array_equal(eng.read_accum().view(np.uint32), got.view(np.uint32), err_msg=name)


def test_tile_order_and_rects(addon_built):
    """The shim's tile order (renderer.ts:573-593) and scissor rectangles
    (renderer.ts:632-643) equal the Python mirror's; the order is a
    permutation of the grid, coarse strides first; the rectangles cover the
    render target, overlapping only where ceil rounding makes them."""
    from dspt.renderer import tile_order, tile_rect
    cases = [(1, (7, 5)), (2, (64, 48)), (3, (33, 7)), (4, (1001, 601)), (4, (1, 1)), (5, (13, 1)), (8, (1920, 1080))]
    r = _node("const {tileOrder, tileRect} = require('./addon/pathtracing_renderer.js');"
              "const cases = JSON.parse(process.argv[1]);"
              "console.log(JSON.stringify(cases.map(([tr, res]) => ({order: tileOrder(tr),"
              " rects: tileOrder(tr).map((t) => tileRect(t, tr, res))}))));", json.dumps(cases))
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert tile_order(4) == [0, 2, 8, 10, 1, 3, 4, 5, 6, 7, 9, 11, 12, 13, 14, 15]
    assert tile_order(2) == [0, 1, 2, 3] and tile_order(1) == [0]
    for (tr, res), o in zip(cases, out):
        order = tile_order(tr)
        assert o["order"] == order and sorted(order) == list(range(tr * tr))
        rects = [tile_rect(t, tr, res) for t in order]
        assert [None if x is None else list(x) for x in rects] == o["rects"]
        W, H = res
        cover = np.zeros((H, W), np.int32)
        for rc in rects:
            if rc is not None:
                x, y, w, h = rc
                assert 0 <= x and 0 <= y and x + w <= W and y + h <= H
                cover[y:y + h, x:x + w] += 1
        assert cover.min() >= 1
        if W % tr == 0 and H % tr == 0:
            assert cover.max() == 1
    # ceil rounding: 1001 / 4 columns overlap at column 751 (tiles 2 and 3)
    assert tile_rect(2, 4, (1001, 601))[0] + tile_rect(2, 4, (1001, 601))[2] == 752
    assert tile_rect(3, 4, (1001, 601))[0] == 751


# The reference renderer's public members (renderer.ts:160-306, 327-365,
# 512, 840, 914): settings properties, mode lists, methods.
REFERENCE_RENDERER_MEMBERS = [
    "exposure", "debugModes", "debugMode", "tonemappingModes", "tonemapping", "maxBounces", "useIBL",
    "showBackground", "forceIBLEval", "enableGamma", "iblRotation", "pixelRatio", "backgroundColor", "rayEps",
    "enableFxaa", "tileRes", "clampThreshold", "resetAccumulation", "stopRendering", "setInteractionMode", "resize",
    "render", "setIBL", "setScene",
]


def test_renderer_surface_matches_reference(addon_built):
    """Both drop-ins expose every public member of the reference renderer,
    checked without a GPU: the Python class's attributes, and the Node
    class's prototype plus the fields its constructor assigns."""
    from dspt.renderer import PathtracingRenderer
    missing = [m for m in REFERENCE_RENDERER_MEMBERS if not hasattr(PathtracingRenderer, m)]
    assert not missing, missing
    assert PathtracingRenderer.debugModes[10:12] == ["IBL PDF", "IBL CDF"] and len(PathtracingRenderer.debugModes) == 15
    r = _node("const {PathtracingRenderer} = require('./addon/pathtracing_renderer.js');"
              "const names = JSON.parse(process.argv[1]);"
              "const proto = PathtracingRenderer.prototype;"
              "const src = PathtracingRenderer.toString();"
              "console.log(JSON.stringify(names.filter((n) => !(n in proto) && !src.includes('this.' + n + ' '))));",
              json.dumps(REFERENCE_RENDERER_MEMBERS))
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout.strip().splitlines()[-1]) == []
