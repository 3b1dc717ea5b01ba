"""Host-side mirror of the reference's renderer surface over the HIP engine.

PathtracingRenderer   packages/lib/renderer.ts:116-1109 (the object the
                      engine replaces): setScene, setIBL, resize, render,
                      stopRendering, resetAccumulation, setInteractionMode and
                      the settings properties; the GL render loop is replaced
                      by dspt_render() on the MI355X.
PathtracingViewport   packages/lib/pathtracing_viewport.ts:100-316 (public
                      embedding API: setScene/setIBL/setLookAtCamera/start/
                      stop/on/off/setSettings).
readAccumulation() returns the linear accumulation; readDisplay() runs the
display chain (tonemap, gamma, FXAA, display resample) on the GPU. Any
debugMode but 'None' renders the debug program (integrator/debug.glsl): one
attribute of each sample's first hit, as u_debug_mode = the mode's index in
debugModes (-1 for a name not in the list), renderer.ts:406.
"""
import math
import time

import numpy as np

from . import _lib
from .scene import PathtracingSceneData, ViewportCamera


def pack_frame_block(camera, width, height, *, frame_count=1, max_bounces=8, use_ibl=True, ibl_rotation_rad=0.0,
                     show_background=True, background_color=(0, 0, 0, 1), ibl_size=(0, 0), ray_eps=1e-4,
                     clamp_threshold=3.0, ibl_total_sum=0.0, scene_counts=(0, 0, 0, 0), light=None, debug_mode=0):
    """The 52-float uniform block of renderer.ts:376-438 (JS doubles rounded into a Float32Array)."""
    film_height = math.tan(camera.fov * 0.5 * math.pi / 180.0) * camera.near
    vals = []
    vals += [float(x) for x in np.asarray(camera.matrixWorld, np.float64).reshape(16)]
    vals += [float(x) for x in background_color]
    vals += [camera.position[0], camera.position[1], camera.position[2], film_height]
    vals += [1.0 / width, 1.0 / height]
    vals += [float(ibl_size[0]), float(ibl_size[1])]
    vals += [float(frame_count), float(debug_mode), 0.0, 1.0 if use_ibl else 0.0]
    vals += [float(ibl_rotation_rad), 1.0 if show_background else 0.0, float(max_bounces), camera.near]
    vals += [float(ibl_total_sum), float(ray_eps), float(clamp_threshold), 0.0]
    vals += [float(x) for x in scene_counts]
    if light is not None:
        vals += [light.position[0], light.position[1], light.position[2], 1.0]
        vals += [light.emission[0], light.emission[1], light.emission[2], 1.0]
    else:
        vals += [0.0] * 8
    block = np.array(vals, dtype=np.float64).astype(np.float32)
    assert block.size == _lib.FRAME_FLOATS
    return block


def tile_order(tile_res):
    """renderer.ts:573-593: the tiles of a tile_res x tile_res grid, coarse
    strides first (every stride-th tile of every stride-th row, stride
    halving down to 1)."""
    tiles, used = [], set()
    stride = tile_res
    while stride >= 1:
        for y in range(0, tile_res, stride):
            for x in range(0, tile_res, stride):
                t = y * tile_res + x
                if t not in used:
                    used.add(t)
                    tiles.append(t)
        if stride == 1:
            break
        stride //= 2
    return tiles


def tile_rect(tile, tile_res, render_res):
    """renderer.ts:632-643: the scissor rectangle of a tile in GL window
    coordinates (row 0 at the bottom; tile rows counted from the top),
    clipped to the render target as the scissor test clips it:
    (x0, y0, width, height), or None when nothing of it is inside."""
    W, H = render_res
    tx, ty = tile % tile_res, tile // tile_res
    x0 = -((-tx * W) // tile_res)  # ceil of the JS double quotient (exact for integers)
    y0 = -((-(tile_res - ty - 1) * H) // tile_res)
    x1 = min(x0 - ((-W) // tile_res), W)
    y1 = min(y0 - ((-H) // tile_res), H)
    return (x0, y0, x1 - x0, y1 - y0) if x1 > x0 and y1 > y0 else None


class PathtracingRenderer:
    """MI355X replacement of the reference PathtracingRenderer (renderer.ts:116)."""

    # renderer.ts:181
    debugModes = ["None", "Albedo", "Metalness", "Roughness", "Normals", "Tangents", "Bitangents", "Transparency", "UV0",
                  "Clearcoat", "IBL PDF", "IBL CDF", "Specular", "SpecularTint", "Translucency"]
    tonemappingModes = ["None", "Reinhard", "Cineon", "AcesFilm"]

    def __init__(self, materialProfile="webgl-lean", device=0, width=1, height=1, tileSchedule="frames"):
        if materialProfile not in ("webgl-lean", "webgl-full"):
            raise ValueError(f"unknown material profile {materialProfile!r}")
        if tileSchedule not in ("frames", "tiles"):
            raise ValueError(f"unknown tile schedule {tileSchedule!r}")
        # 'frames': whole frames per engine call; 'tiles': the reference's
        # scissored tile passes (renderer.ts:567-643), one engine call per tile
        self._tileSchedule = tileSchedule
        self._tileIndex = 0
        self.materialProfile = materialProfile
        self._engine = _lib.Engine(_lib.PROFILE_FULL if materialProfile == "webgl-full" else _lib.PROFILE_LEAN, device)
        # settings with the reference defaults (renderer.ts:172-306)
        self._exposure = 1.0
        self._tonemapping = "None"
        self._maxBounces = 8
        self._useIBL = True
        self._showBackground = True
        self._enableGamma = True
        self._iblRotation = 0.0  # radians
        self._pixelRatio = 1.0
        self._backgroundColor = [0.0, 0.0, 0.0, 1.0]
        self._rayEps = 0.0001
        self._enableFxaa = False
        self._tileRes = 4
        self._clampThreshold = 3.0
        self._debugMode = "None"
        self._forceIBLEval = False
        self._frameCount = 1
        self._isRendering = False
        self._reset_pending = False
        self.schedulerMode = "stopped"
        self._scene = None
        self._bvh_counts = (0, 0)
        self._ibl = None
        self._ibl_total = 0.0
        self.displayRes = [0, 0]
        self.renderRes = [0, 0]
        self.diagnostics = {"backend": "hip-gfx950", "sampleCount": 0, "mode": "stopped", "profiling": {}}
        self.resize(width, height)

    # ---- settings (renderer.ts:172-306); every physics setter resets accumulation
    def _setting(name, reset=True, conv=None):
        def get(self):
            return getattr(self, "_" + name)

        def set_(self, v):
            setattr(self, "_" + name, conv(v) if conv else v)
            if reset:
                self.resetAccumulation()
        return property(get, set_)

    exposure = _setting("exposure")
    tonemapping = _setting("tonemapping", reset=False)
    maxBounces = _setting("maxBounces", conv=int)
    useIBL = _setting("useIBL", conv=bool)
    showBackground = _setting("showBackground", conv=bool)
    enableGamma = _setting("enableGamma")
    backgroundColor = _setting("backgroundColor", conv=list)
    rayEps = _setting("rayEps", conv=float)
    enableFxaa = _setting("enableFxaa", reset=False)
    clampThreshold = _setting("clampThreshold", conv=float)
    # stored and resets the accumulation; no shader reads it (renderer.ts:229-236)
    forceIBLEval = _setting("forceIBLEval", conv=bool)
    del _setting

    @property
    def iblRotation(self):
        return self._iblRotation / math.pi * 180.0

    @iblRotation.setter
    def iblRotation(self, deg):
        self._iblRotation = deg / 180.0 * math.pi
        self.resetAccumulation()

    @property
    def pixelRatio(self):
        return self._pixelRatio

    @pixelRatio.setter
    def pixelRatio(self, v):
        self._pixelRatio = v
        self.resize(*self.displayRes)

    @property
    def tileRes(self):
        return self._tileRes

    @tileRes.setter
    def tileRes(self, v):
        self._tileRes = max(2, int(v) or 2)

    @property
    def tileSchedule(self):
        return self._tileSchedule

    @tileSchedule.setter
    def tileSchedule(self, v):
        if v not in ("frames", "tiles"):
            raise ValueError(f"unknown tile schedule {v!r}")
        self._tileSchedule = v
        self.resetAccumulation()

    @property
    def debugMode(self):
        return self._debugMode

    @debugMode.setter
    def debugMode(self, v):
        self._debugMode = v
        self.resetAccumulation()

    # ---- scene / IBL / resolution
    def setScene(self, scene: PathtracingSceneData, progress=None):
        t0 = time.perf_counter()
        self.stopRendering()
        from .scenes import InstancedScene
        if isinstance(scene, InstancedScene):
            # engine extension: object-space meshes + instances, both BVH levels built by the engine
            self._engine.upload_scene_instanced(*scene.engine_args())
            inner = scene.scene
            if getattr(inner, "num_textures", 0):
                dims, texels = inner.getTextureArrays()
                self._engine.scene_set_textures(0, inner.getFlatTextureInfoBuffer(), dims, texels)
            self._scene = inner
            self._bvh_counts = (0, 0)
            self.diagnostics["scene"] = {"triangles": inner.num_triangles, "instances": scene.num_instances,
                                         "instancedTriangles": scene.instanced_triangles,
                                         "materials": inner.num_materials, "lights": len(inner.lights)}
            self.diagnostics["profiling"]["rendererSetupMs"] = (time.perf_counter() - t0) * 1e3
            self.resetAccumulation()
            return
        if scene.bvhNodes is not None:
            nodes, bidx = scene.bvhNodes, scene.bvhTriangleIndices
        else:
            pos = scene.bvhPositionBuffer if scene.bvhPositionBuffer is not None else scene.getPositionBuffer()
            ind = scene.bvhIndexBuffer if scene.bvhIndexBuffer is not None else scene.triangleIndexBuffer
            t1 = time.perf_counter()
            nodes, bidx = _lib.build_bvh(pos, ind)
            self.diagnostics["profiling"]["bvhBuildMs"] = (time.perf_counter() - t1) * 1e3
        self._engine.upload_scene(scene.vertexBuffer, scene.triangleIndexBuffer, nodes, bidx,
                                  scene.getFlatMaterialBuffer())
        if getattr(scene, "num_textures", 0):
            # scene_data_webgl2.ts:233-363: texture arrays + TexInfo records
            dims, texels = scene.getTextureArrays()
            self._engine.scene_set_textures(0, scene.getFlatTextureInfoBuffer(), dims, texels)
        self._scene = scene
        self._bvh_counts = (len(nodes) // 16, len(bidx))
        self.diagnostics["scene"] = {"triangles": scene.num_triangles, "materials": scene.num_materials,
                                     "lights": len(scene.lights)}
        self.diagnostics["bvh"] = {"nodes": self._bvh_counts[0], "indices": self._bvh_counts[1]}
        self.diagnostics["profiling"]["rendererSetupMs"] = (time.perf_counter() - t0) * 1e3
        self.resetAccumulation()

    def setIBL(self, texture):
        """texture: {'image': {'data': f32[4wh], 'width': w, 'height': h}} (ibl_loader.ts:4-14)."""
        img = texture["image"] if isinstance(texture, dict) else texture.image
        data, w, h = img["data"], int(img["width"]), int(img["height"])
        self._ibl_total = self._engine.upload_ibl(np.asarray(data, np.float32), w, h)
        self._ibl = (w, h)
        self.resetAccumulation()

    def resize(self, width, height):
        self.displayRes = [int(width), int(height)]
        self.renderRes = [int(math.ceil(width * self._pixelRatio)), int(math.ceil(height * self._pixelRatio))]
        self._engine.resize(*self.renderRes)
        self.resetAccumulation()

    # ---- render loop
    def resetAccumulation(self):
        """renderer.ts:327-336: the next frame restarts at frame 1 (with the
        camera and settings of that moment)."""
        self._frameCount = 1
        self._tileIndex = 0
        self._reset_pending = True
        if self._isRendering and self.schedulerMode == "accumulating":
            self.schedulerMode = "settling"
        self._update_diagnostics()

    def stopRendering(self):
        """renderer.ts:338-352: no frame is submitted after this call."""
        self._isRendering = False
        self.schedulerMode = "stopped"
        self._update_diagnostics()

    def setInteractionMode(self, flag):
        """renderer.ts:354-363."""
        if flag:
            self.schedulerMode = "interactive"
        elif self.schedulerMode != "stopped":
            self.schedulerMode = "settling"
        self.resetAccumulation()

    def _active_tile_res(self):
        # renderer.ts:567-571
        if self.schedulerMode == "interactive":
            return 1
        if self.schedulerMode == "settling":
            return 2
        return max(2, self._tileRes or 2)

    def _update_diagnostics(self):
        # renderer.ts:972-981
        d = self.diagnostics
        d["sampleCount"] = max(0, self._frameCount - 1)
        d["mode"] = self.schedulerMode
        d["renderResolution"] = f"{self.renderRes[0]}x{self.renderRes[1]}"
        d["displayResolution"] = f"{self.displayRes[0]}x{self.displayRes[1]}"
        d["tileResolution"] = self._active_tile_res()

    def frame_block(self, camera, frame_count=None):
        light = self._scene.lights[0] if (self._scene is not None and self._scene.lights) else None
        return pack_frame_block(
            camera, self.renderRes[0], self.renderRes[1],
            frame_count=self._frameCount if frame_count is None else frame_count,
            max_bounces=self._maxBounces, use_ibl=self._useIBL, ibl_rotation_rad=self._iblRotation,
            show_background=self._showBackground, background_color=self._backgroundColor,
            ibl_size=self._ibl or (0, 0), ray_eps=self._rayEps, clamp_threshold=self._clampThreshold,
            ibl_total_sum=self._ibl_total,
            scene_counts=(self._scene.num_triangles if self._scene else 0, self._bvh_counts[0], self._bvh_counts[1], 0),
            light=light, debug_mode=self._debug_index())

    def _debug_index(self):
        """debugModes.indexOf(debugMode) (renderer.ts:406): -1 for a name not in the list."""
        return self.debugModes.index(self._debugMode) if self._debugMode in self.debugModes else -1

    PATHS_PER_CALL = 1 << 24

    def render(self, camera, num_samples, tileFinishedCB=None, frameFinishedCB=None, renderingFinishedCB=None,
               frames_per_call=None):
        """renderer.ts:512-565: stops a running render, resets the
        accumulation and renders frames 1, 2, ... until the frame counter
        reaches num_samples (frames 1 .. num_samples-1), or, for
        num_samples == -1, until stopRendering() is called (from a callback
        or another thread). A bounded render goes to the engine in one call
        (or frames_per_call at a time); an unbounded one in calls of growing
        size (1, 2, 4, ... up to about one path pool). The camera is read
        again for every call, and resetAccumulation() restarts at frame 1. Per
        frame, tileFinishedCB runs once per tile of the active tile order and
        frameFinishedCB(frameCount) once. Frames of a call that was on the GPU
        when stopRendering() came are counted without callbacks. With
        tileSchedule 'tiles' every tile of the active tile order is its own
        engine call over its scissor rectangle, followed by tileFinishedCB,
        and frameFinishedCB follows the frame's last tile, as in the
        reference (renderer.ts:595-714)."""
        self.stopRendering()
        self.schedulerMode = "settling"
        self._isRendering = True
        self.resetAccumulation()
        ramp = 1
        while self._isRendering:
            tiles = self._tileSchedule == "tiles"
            if self._reset_pending:
                self._reset_pending, self._frameCount, self._tileIndex, ramp = False, 1, 0, 1
                if tiles:
                    self._engine.clear_accum()  # a tile pass never clears: the reset does (renderer.ts:604-621)
            if num_samples != -1 and self._frameCount >= num_samples:
                self.stopRendering()
                if renderingFinishedCB:
                    renderingFinishedCB()
                return
            if tiles:
                self._tile_pass(camera, tileFinishedCB, frameFinishedCB)
                continue
            first = self._frameCount
            cap = max(1, self.PATHS_PER_CALL // (self.renderRes[0] * self.renderRes[1]))
            if frames_per_call:
                n = max(1, frames_per_call)
            else:  # a bounded render goes to the engine in one call; an unbounded one ramps up
                n = min(ramp, cap) if num_samples == -1 else num_samples - first
            if num_samples != -1:
                n = min(n, num_samples - first)
            self._engine.set_frame(self.frame_block(camera, first))
            self._engine.render(first, n)
            self._engine.synchronize()
            for _ in range(n):
                if self._reset_pending:
                    break  # frames of this call after a reset are dropped
                self._frameCount += 1
                if not self._isRendering:
                    continue  # stopped from a callback: counted, no callbacks
                for _t in range(self._active_tile_res() ** 2):
                    if tileFinishedCB:
                        tileFinishedCB()
                if frameFinishedCB:
                    frameFinishedCB(self._frameCount)
                if self.schedulerMode == "settling":
                    self.schedulerMode = "accumulating"
            self._update_diagnostics()
            ramp = min(ramp * 2, cap)

    def _tile_pass(self, camera, tileFinishedCB, frameFinishedCB):
        """One tile of the active tile order at the current frame (renderer.ts:595-714)."""
        tr = self._active_tile_res()
        order = tile_order(tr)
        rect = tile_rect(order[self._tileIndex % len(order)], tr, self.renderRes)
        frame = self._frameCount
        if rect is not None:
            self._engine.set_frame(self.frame_block(camera, frame))
            self._engine.render_region(frame, 1, *rect)
            self._engine.synchronize()
        if self._reset_pending:
            return  # reset from another thread meanwhile: the next pass restarts
        self._tileIndex += 1
        if self._isRendering and tileFinishedCB:
            tileFinishedCB()
        if self._tileIndex >= tr * tr:
            self._tileIndex = 0
            self._frameCount += 1
            if self._isRendering and frameFinishedCB:
                frameFinishedCB(self._frameCount)
            if self.schedulerMode == "settling":
                self.schedulerMode = "accumulating"
        self._update_diagnostics()

    def readAccumulation(self):
        """Linear RGBA f32 accumulation buffer, shape (H, W, 4), row 0 = bottom."""
        return self._engine.read_accum()

    def readDisplay(self, want_float=False):
        """What the reference's display chain puts on the canvas
        (renderer.ts:665-700): tonemapping (an unknown name applies none, as
        indexOf() == -1 does), exposure, gamma, FXAA, resampled to the display
        resolution. RGBA8 u8[displayH, displayW, 4], row 0 = bottom; with
        want_float also the float display buffer."""
        mode = self.tonemappingModes.index(self._tonemapping) if self._tonemapping in self.tonemappingModes else -1
        return self._engine.postprocess(mode, self._exposure, bool(self._enableGamma), bool(self._enableFxaa),
                                        (self.displayRes[0], self.displayRes[1]), want_float)

    def saveImage(self, path):
        """The displayed canvas as a PNG file (top row first)."""
        from .image import write_png
        write_png(path, self.readDisplay()[::-1])

    @property
    def engine(self):
        return self._engine


class PathtracingViewport:
    """Mirror of pathtracing_viewport.ts:100-316 without a canvas."""

    def __init__(self, width=1, height=1, materialProfile="webgl-lean", settings=None, device=0):
        self.renderer = PathtracingRenderer(materialProfile=materialProfile, device=device, width=width, height=height)
        self.camera = ViewportCamera()
        self._listeners = {}
        self.running = False
        self.sampleLimit = -1
        if settings:
            self.setSettings(settings)

    @property
    def diagnostics(self):
        return self.renderer.diagnostics

    def on(self, event, listener):
        self._listeners.setdefault(event, []).append(listener)
        return lambda: self.off(event, listener)

    def off(self, event, listener):
        if listener in self._listeners.get(event, []):
            self._listeners[event].remove(listener)

    def _emit(self, event, payload):
        for l in list(self._listeners.get(event, [])):
            l(payload)

    _SETTINGS = {
        "exposure": "exposure", "toneMapping": "tonemapping", "pixelRatio": "pixelRatio", "maxBounces": "maxBounces",
        "useIBL": "useIBL", "showBackground": "showBackground", "enableGamma": "enableGamma",
        "enableFxaa": "enableFxaa", "iblRotationDegrees": "iblRotation", "backgroundColor": "backgroundColor",
        "rayEps": "rayEps", "clampThreshold": "clampThreshold", "tileResolution": "tileRes", "debugMode": "debugMode",
    }

    def setSettings(self, settings):
        for k, v in settings.items():
            if k in self._SETTINGS:
                setattr(self.renderer, self._SETTINGS[k], v)
        self._emit("diagnosticsChanged", {"diagnostics": self.diagnostics})

    def setScene(self, scene):
        self.renderer.setScene(scene)
        self._emit("sceneLoaded", {"scene": scene})

    def setIBL(self, ibl):
        self.renderer.setIBL(ibl)
        self._emit("iblLoaded", {"ibl": ibl})

    # pathtracing_viewport.ts:172-234 over local files (URLs are not fetched)
    def loadSceneFromFile(self, path, progress=None):
        from .loaders import loadSceneFromFile
        try:
            scene, profile = loadSceneFromFile(path, progress=progress)
            self.renderer.setScene(scene, progress)
        except Exception as e:
            self._emit("sceneLoadFailed", {"error": e})
            raise
        self.renderer.diagnostics["profiling"] = {**profile, **self.renderer.diagnostics.get("profiling", {})}
        self._emit("sceneLoaded", {"scene": scene, "loaded": {"scene": scene, "profile": profile}})
        return {"scene": scene, "profile": profile}

    def loadIblFromFile(self, path):
        from .loaders import loadIblFromFile
        try:
            ibl = loadIblFromFile(path)
            self.setIBL(ibl)
        except Exception as e:
            self._emit("iblLoadFailed", {"error": e})
            raise
        return ibl

    def setCamera(self, camera):
        self.camera = camera
        self.renderer.resetAccumulation()

    def setLookAtCamera(self, position, target, up=(0, 1, 0), fov=None, near=None):
        if fov is not None:
            self.camera.fov = fov
        if near is not None:
            self.camera.near = near
        self.camera.setLookAt(position, target, up)
        self.setCamera(self.camera)

    def resize(self, width, height):
        self.renderer.resize(max(1, int(width)), max(1, int(height)))

    def start(self, sampleLimit=-1):
        self.running = True
        self.sampleLimit = sampleLimit
        self.renderer.render(
            self.camera, sampleLimit, None,
            lambda n: self._emit("frameRendered", {"frameCount": n}),
            lambda: self._emit("renderFinished", {"frameCount": self.diagnostics["sampleCount"]}))
        self.running = False

    def stop(self):
        self.running = False
        self.renderer.stopRendering()

    def resetAccumulation(self):
        self.renderer.resetAccumulation()

    def setInteractionMode(self, active):
        self.renderer.setInteractionMode(active)

    def readAccumulation(self):
        return self.renderer.readAccumulation()

    def readDisplay(self, want_float=False):
        return self.renderer.readDisplay(want_float)

    def saveImage(self, path):
        self.renderer.saveImage(path)
