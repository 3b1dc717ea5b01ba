'use strict';
// Node-side drop-in for the reference's PathtracingRenderer
// (packages/lib/renderer.ts:116-1109) backed by the MI355X engine through the
// N-API addon. Same method names, argument meaning and error behaviour:
// setScene(scene) / setIBL(texture) / resize(w, h) / render(camera, numSamples,
// tileCB, frameCB, finishedCB) / stopRendering() / resetAccumulation() and the
// settings properties. readAccumulation() returns the linear RGBA32F buffer;
// readDisplay() runs the display chain (tonemap, gamma, FXAA, display resample,
// renderer.ts:665-700) on the GPU and returns the RGBA8 canvas; saveImage(path)
// writes that canvas as a PNG.
const path = require('path');
const native = require(path.join(__dirname, 'build', 'dspt.node'));

const FRAME_FLOATS = 52;

// renderer.ts:181
const DEBUG_MODES = ['None', 'Albedo', 'Metalness', 'Roughness', 'Normals', 'Tangents', 'Bitangents', 'Transparency',
                     'UV0', 'Clearcoat', 'IBL PDF', 'IBL CDF', 'Specular', 'SpecularTint', 'Translucency'];

function packFrameBlock(r, camera, frameCount) {
  // renderer.ts:376-438: JS numbers rounded into a Float32Array in key order
  const filmHeight = Math.tan(camera.fov * 0.5 * Math.PI / 180.0) * camera.near;
  const light = r._scene && r._scene.lights && r._scene.lights.length ? r._scene.lights[0] : null;
  const v = [];
  v.push(...Array.from(camera.matrixWorld.elements));
  v.push(...r._backgroundColor);
  v.push(camera.position.x, camera.position.y, camera.position.z, filmHeight);
  v.push(1.0 / r.renderRes[0], 1.0 / r.renderRes[1]);
  v.push(r._iblSize[0], r._iblSize[1]);
  v.push(frameCount, DEBUG_MODES.indexOf(r._debugMode === undefined ? 'None' : r._debugMode), 0, r._useIBL ? 1 : 0);  // u_debug_mode, renderer.ts:406
  v.push(r._iblRotation, r._showBackground ? 1 : 0, r._maxBounces, camera.near);
  v.push(r._iblTotalSum, r._rayEps, r._clampThreshold, 0);
  v.push(r._sceneCounts[0], r._sceneCounts[1], r._sceneCounts[2], 0);
  if (light) v.push(light.position[0], light.position[1], light.position[2], 1, light.emission[0], light.emission[1], light.emission[2], 1);
  else v.push(0, 0, 0, 0, 0, 0, 0, 0);
  const block = new Float32Array(v);
  if (block.length !== FRAME_FLOATS) throw new Error('frame block layout');
  return block;
}

// Paths per native render call once the loop has ramped up: about one path
// pool, so stopRendering(), resetAccumulation() and camera moves take effect
// within a few milliseconds of GPU work.
const PATHS_PER_CALL = 1 << 24;

class PathtracingRenderer {
  constructor(parameters = {}) {
    this.materialProfile = parameters.materialProfile || 'webgl-lean';
    if (this.materialProfile !== 'webgl-lean' && this.materialProfile !== 'webgl-full')
      throw new Error(`unknown material profile ${this.materialProfile}`);
    this._h = native.create(this.materialProfile === 'webgl-full' ? 1 : 0, parameters.device || 0);
    this._maxBounces = 8; this._useIBL = true; this._showBackground = true; this._iblRotation = 0;
    this._backgroundColor = [0, 0, 0, 1]; this._rayEps = 0.0001; this._clampThreshold = 3.0;
    this._exposure = 1.0; this._tonemapping = 'None'; this._enableGamma = true; this._enableFxaa = false;
    this._pixelRatio = 1.0; this._tileRes = 4; this._debugMode = 'None'; this._forceIBLEval = false;
    // 'frames': whole frames per native call (a 1080p frame takes milliseconds
    // here, shorter than one of the reference's tiles); 'tiles': the
    // reference's scissored tile passes, one native call per tile
    this._tileSchedule = parameters.tileSchedule || 'frames';
    if (this._tileSchedule !== 'frames' && this._tileSchedule !== 'tiles')
      throw new Error(`unknown tile schedule ${this._tileSchedule}`);
    this._tileIndex = 0;
    // renderer.ts:181,192
    this.debugModes = DEBUG_MODES.slice();
    this.tonemappingModes = TONEMAPPING_MODES.slice();
    this._frameCount = 1; this._isRendering = false; this.schedulerMode = 'stopped';
    this._gen = 0;               // bumped by render() and stopRendering(): a loop of an older generation exits
    this._inflight = null;       // promise of the native call on the GPU, if any
    this._resetPending = false;  // accumulation restarts at frame 1 before the next call
    this._pending = [];          // native updates requested while a call was in flight
    this._iblSize = [0, 0]; this._iblTotalSum = 0; this._sceneCounts = [0, 0, 0];
    this._scene = null;
    this.diagnostics = { backend: 'hip-gfx950', sampleCount: 0, mode: 'stopped', renderResolution: '',
                         displayResolution: '', tileResolution: 2, framePending: false, profiling: {},
                         shaderPrograms: [] };
    this.displayRes = [0, 0]; this.renderRes = [0, 0];
    this.resize(parameters.width || 1, parameters.height || 1);
  }

  get maxBounces() { return this._maxBounces; } set maxBounces(v) { this._maxBounces = v; this.resetAccumulation(); }
  get useIBL() { return this._useIBL; } set useIBL(v) { this._useIBL = v; this.resetAccumulation(); }
  get showBackground() { return this._showBackground; } set showBackground(v) { this._showBackground = v; this.resetAccumulation(); }
  get forceIBLEval() { return this._forceIBLEval; } set forceIBLEval(v) { this._forceIBLEval = v; this.resetAccumulation(); }
  get iblRotation() { return this._iblRotation / Math.PI * 180.0; } set iblRotation(v) { this._iblRotation = v / 180.0 * Math.PI; this.resetAccumulation(); }
  get backgroundColor() { return this._backgroundColor; } set backgroundColor(v) { this._backgroundColor = v; this.resetAccumulation(); }
  get rayEps() { return this._rayEps; } set rayEps(v) { this._rayEps = v; this.resetAccumulation(); }
  get clampThreshold() { return this._clampThreshold; } set clampThreshold(v) { this._clampThreshold = v; this.resetAccumulation(); }
  get exposure() { return this._exposure; } set exposure(v) { this._exposure = v; this.resetAccumulation(); }
  get tonemapping() { return this._tonemapping; } set tonemapping(v) { this._tonemapping = v; }
  get enableGamma() { return this._enableGamma; } set enableGamma(v) { this._enableGamma = v; this.resetAccumulation(); }
  get enableFxaa() { return this._enableFxaa; } set enableFxaa(v) { this._enableFxaa = v; }
  get pixelRatio() { return this._pixelRatio; } set pixelRatio(v) { this._pixelRatio = v; this.resize(this.displayRes[0], this.displayRes[1]); }
  get tileRes() { return this._tileRes; } set tileRes(v) { this._tileRes = Math.max(2, Math.floor(v) || 2); }
  get tileSchedule() { return this._tileSchedule; }
  set tileSchedule(v) {
    if (v !== 'frames' && v !== 'tiles') throw new Error(`unknown tile schedule ${v}`);
    this._tileSchedule = v; this.resetAccumulation();
  }
  // renderer.ts:182-190. Any mode but 'None' renders the debug program (one
  // attribute of the first hit per sample); a name not in debugModes packs
  // -1, which shows the hit as black and the background as in 'Albedo'.
  get debugMode() { return this._debugMode; } set debugMode(v) { this._debugMode = v; this.resetAccumulation(); }

  // resolves once no native call is in flight (readAccumulation, readDisplay
  // and saveImage need an idle engine)
  async idle() {
    while (this._inflight) await this._inflight.catch(() => {});
    this._applyPending();
  }

  // renderer.ts:914-958; the BVH is built by the engine when the scene carries no prebuilt one
  async setScene(scene, progress) {
    this.stopRendering();
    await this.idle();
    if (progress) progress({ phase: 'Building BVH' });
    let nodes = scene.bvhNodes, idx = scene.bvhTriangleIndices;
    if (!nodes) {
      const t0 = Date.now();
      const r = native.buildBvh(scene.bvhPositionBuffer, scene.bvhIndexBuffer, 4);
      nodes = r.nodeData; idx = r.triangleIndices;
      this.diagnostics.profiling.bvhBuildMs = Date.now() - t0;
    }
    if (progress) progress({ phase: 'Uploading scene' });
    const mats = new Float32Array(scene.materials.length * 68);
    scene.materials.forEach((m, i) => mats.set(m.data, i * 68));
    native.uploadScene(this._h, scene.vertexBuffer, scene.triangleIndexBuffer, nodes, idx, mats);
    // scene_data_webgl2.ts:233-363: texture arrays (RGBA8, row 0 first) + TexInfo records
    if (scene.textures) native.sceneSetTextures(this._h, 0, scene.textures.texInfos, scene.textures.arrayDims, scene.textures.texels);
    this._scene = scene;
    this._sceneCounts = [scene.triangleIndexBuffer.length / 3, nodes.length / 16, idx.length];
    this.diagnostics.bvh = { nodes: nodes.length / 16, indices: idx.length };
    if (progress) progress({ phase: 'Ready' });
    this.resetAccumulation();
  }

  // renderer.ts:840-865 (applied before the next frame when a call is in flight)
  setIBL(texture) {
    this._defer(() => {
      const img = texture.image;
      this._iblTotalSum = native.uploadIbl(this._h, img.data, img.width, img.height);
      this._iblSize = [img.width, img.height];
    });
    this.resetAccumulation();
  }

  // renderer.ts:365-372
  resize(width, height) {
    this.displayRes = [width, height];
    this.renderRes = [Math.ceil(width * this._pixelRatio), Math.ceil(height * this._pixelRatio)];
    const [w, h] = this.renderRes;
    this._defer(() => native.resize(this._h, w, h));
    this.resetAccumulation();
  }

  // renderer.ts:327-336: the next frame restarts the accumulation at frame 1
  // (with the camera and settings of that moment)
  resetAccumulation() {
    this._frameCount = 1;
    this._tileIndex = 0;
    this._resetPending = true;
    if (this._isRendering && this.schedulerMode === 'accumulating') this.schedulerMode = 'settling';
    this._updateDiagnostics();
  }

  // renderer.ts:338-352: no frame is submitted after this call; frames of a
  // call already on the GPU finish into the accumulation (and are counted)
  stopRendering() {
    this._gen++;
    this._isRendering = false;
    this.schedulerMode = 'stopped';
    this._updateDiagnostics();
  }

  // renderer.ts:354-363
  setInteractionMode(flag) {
    if (flag) this.schedulerMode = 'interactive';
    else if (this.schedulerMode !== 'stopped') this.schedulerMode = 'settling';
    this.resetAccumulation();
    this._updateDiagnostics();
  }

  // renderer.ts:512-565: render(camera, numSamples, tileFinishedCB,
  // frameFinishedCB, renderingFinishedCB) stops a running render, then renders
  // frames 1, 2, ... until the frame counter reaches numSamples (frames
  // 1 .. numSamples-1), or until stopRendering() when numSamples is -1. The
  // camera object is read again for every call, so a caller that moves it and
  // calls resetAccumulation() (as the viewport does) restarts with the new
  // view. With tileSchedule 'frames', frames are submitted in calls of
  // growing size (1, 2, 4, ... frames, up to about one path pool); per frame,
  // tileFinishedCB runs once per tile of the active tile order and
  // frameFinishedCB(frameCount) once. With 'tiles', every tile of the active
  // tile order (renderer.ts:567-593) is its own native call over the
  // scissor rectangle of renderer.ts:632-643, followed by tileFinishedCB,
  // and frameFinishedCB follows the frame's last tile, as in the reference.
  // Returns a promise that settles when the loop ends; errors stop the loop
  // and reject it.
  render(camera, numSamples, tileFinishedCB, frameFinishedCB, renderingFinishedCB, framesPerCall) {
    this.stopRendering();
    this._camera = camera;
    this._numSamples = numSamples;
    this._cb = { tile: tileFinishedCB, frame: frameFinishedCB, done: renderingFinishedCB };
    this._framesPerCall = framesPerCall || 0;
    this.schedulerMode = 'settling';
    this._isRendering = true;
    this.resetAccumulation();
    const gen = ++this._gen;
    const p = this._loop(gen);
    p.catch(() => {});  // the caller may ignore the promise, as the reference's render() returns nothing
    return p;
  }

  async _loop(gen) {
    let ramp = 1;
    try {
      for (;;) {
        while (this._inflight) await this._inflight.catch(() => {});  // a call of an earlier loop
        if (gen !== this._gen || !this._isRendering) return;
        this._applyPending();
        const tiles = this._tileSchedule === 'tiles';
        if (this._resetPending) {
          this._resetPending = false; this._frameCount = 1; this._tileIndex = 0; ramp = 1;
          if (tiles) native.clearAccum(this._h);  // a tile pass never clears: the reset does (renderer.ts:604-621)
        }
        if (this._numSamples !== -1 && this._frameCount >= this._numSamples) {
          this.stopRendering();
          if (this._cb.done) this._cb.done();
          return;
        }
        if (tiles) {
          await this._tilePass(gen);
          continue;
        }
        const first = this._frameCount;
        const cap = Math.max(1, Math.floor(PATHS_PER_CALL / (this.renderRes[0] * this.renderRes[1])));
        let n = this._framesPerCall || Math.min(ramp, cap);
        if (this._numSamples !== -1) n = Math.min(n, this._numSamples - first);
        native.setFrame(this._h, packFrameBlock(this, this._camera, first));
        const call = native.render(this._h, first, n);
        this._inflight = call;
        this.diagnostics.framePending = true;
        try {
          await call;
        } finally {
          if (this._inflight === call) this._inflight = null;
          this.diagnostics.framePending = false;
        }
        if (this._resetPending) continue;  // reset while on the GPU: those frames are dropped
        const live = gen === this._gen && this._isRendering;
        for (let f = 0; f < n; f++) {
          this._frameCount++;
          if (!live) continue;  // stopped meanwhile: counted, no callbacks
          const tiles = this._activeTileRes() ** 2;
          for (let t = 0; t < tiles; t++) if (this._cb.tile) this._cb.tile();
          if (this._cb.frame) this._cb.frame(this._frameCount);
          if (this.schedulerMode === 'settling') this.schedulerMode = 'accumulating';
        }
        this._updateDiagnostics();
        if (!live) return;
        ramp = Math.min(ramp * 2, cap);
      }
    } catch (e) {
      if (gen === this._gen) this.stopRendering();
      throw e;
    }
  }

  // One tile of the active tile order at the current frame (renderer.ts:595-714)
  async _tilePass(gen) {
    const tr = this._activeTileRes();
    const order = tileOrder(tr);
    const tile = order[this._tileIndex % order.length];
    const rect = tileRect(tile, tr, this.renderRes);
    const frame = this._frameCount;
    if (rect) {
      native.setFrame(this._h, packFrameBlock(this, this._camera, frame));
      const call = native.renderRegion(this._h, frame, 1, rect[0], rect[1], rect[2], rect[3]);
      this._inflight = call;
      this.diagnostics.framePending = true;
      try {
        await call;
      } finally {
        if (this._inflight === call) this._inflight = null;
        this.diagnostics.framePending = false;
      }
    }
    if (this._resetPending) return;  // reset while on the GPU: the next pass restarts
    const live = gen === this._gen && this._isRendering;
    this._tileIndex++;
    if (live && this._cb.tile) this._cb.tile();
    if (this._tileIndex >= tr * tr) {
      this._tileIndex = 0;
      this._frameCount++;
      if (live && this._cb.frame) this._cb.frame(this._frameCount);
      if (this.schedulerMode === 'settling') this.schedulerMode = 'accumulating';
    }
    this._updateDiagnostics();
  }

  // renderer.ts:567-571
  _activeTileRes() {
    if (this.schedulerMode === 'interactive') return 1;
    if (this.schedulerMode === 'settling') return 2;
    return Math.max(2, this._tileRes || 2);
  }

  // renderer.ts:972-981
  _updateDiagnostics() {
    const d = this.diagnostics;
    d.sampleCount = Math.max(0, this._frameCount - 1);
    d.mode = this.schedulerMode;
    d.renderResolution = `${this.renderRes[0]}x${this.renderRes[1]}`;
    d.displayResolution = `${this.displayRes[0]}x${this.displayRes[1]}`;
    d.tileResolution = this._activeTileRes();
    d.framePending = !!this._inflight;
  }

  // runs fn now when no call is in flight, else before the next call
  _defer(fn) {
    this._pending.push(fn);
    if (!this._inflight) this._applyPending();
    else this._inflight.catch(() => {}).then(() => { if (!this._inflight) this._applyPending(); });
  }

  _applyPending() {
    while (this._pending.length) this._pending.shift()();
  }

  _checkIdle() {
    if (this._inflight) throw new Error('a render call is on the GPU: await renderer.idle() first');
    this._applyPending();
  }

  readAccumulation() {
    this._checkIdle();
    const out = new Float32Array(this.renderRes[0] * this.renderRes[1] * 4);
    native.readAccum(this._h, out);
    return out;
  }

  // renderer.ts:192 lists the operators; an unknown name gives indexOf() == -1: no operator
  readDisplay() {
    this._checkIdle();
    const mode = TONEMAPPING_MODES.indexOf(this._tonemapping);
    return native.postprocess(this._h, -1, mode, this._exposure, !!this._enableGamma, !!this._enableFxaa,
                              this.displayRes[0], this.displayRes[1]);
  }

  saveImage(file) {
    require('fs').writeFileSync(file, encodePng(this.readDisplay(), this.displayRes[0], this.displayRes[1]));
  }

  counters() { return native.counters(this._h); }
}

const TONEMAPPING_MODES = ['None', 'Reinhard', 'Cineon', 'AcesFilm'];

// renderer.ts:573-593: the tiles of a tileRes x tileRes grid, coarse strides
// first (every stride-th tile of every stride-th row, stride halving down to 1)
const TILE_ORDERS = new Map();
function tileOrder(tileRes) {
  if (TILE_ORDERS.has(tileRes)) return TILE_ORDERS.get(tileRes);
  const tiles = [];
  const used = new Set();
  for (let stride = tileRes; stride >= 1; stride = Math.floor(stride / 2)) {
    for (let y = 0; y < tileRes; y += stride)
      for (let x = 0; x < tileRes; x += stride) {
        const t = y * tileRes + x;
        if (!used.has(t)) { used.add(t); tiles.push(t); }
      }
    if (stride === 1) break;
  }
  TILE_ORDERS.set(tileRes, tiles);
  return tiles;
}

// renderer.ts:632-643: the scissor rectangle of a tile in GL window
// coordinates (row 0 at the bottom; tile rows counted from the top), clipped
// to the render target as the scissor test clips it: [x0, y0, width, height],
// or null when nothing of it is inside
function tileRect(tile, tileRes, renderRes) {
  const tilex = tile % tileRes, tiley = Math.floor(tile / tileRes);
  const x0 = Math.ceil(tilex * renderRes[0] / tileRes);
  const y0 = Math.ceil((tileRes - tiley - 1) * renderRes[1] / tileRes);
  const x1 = Math.min(x0 + Math.ceil(renderRes[0] / tileRes), renderRes[0]);
  const y1 = Math.min(y0 + Math.ceil(renderRes[1] / tileRes), renderRes[1]);
  return x1 > x0 && y1 > y0 ? [x0, y0, x1 - x0, y1 - y0] : null;
}

// PNG of an RGBA8 canvas given bottom row first (GL order): rows are written top first
const CRC_TABLE = (() => {
  const t = new Uint32Array(256);
  for (let n = 0; n < 256; n++) {
    let c = n;
    for (let k = 0; k < 8; k++) c = c & 1 ? 0xedb88320 ^ (c >>> 1) : c >>> 1;
    t[n] = c >>> 0;
  }
  return t;
})();
function crc32(buf) {
  let c = 0xffffffff;
  for (let i = 0; i < buf.length; i++) c = CRC_TABLE[(c ^ buf[i]) & 0xff] ^ (c >>> 8);
  return (c ^ 0xffffffff) >>> 0;
}
function pngChunk(tag, data) {
  const len = Buffer.alloc(4); len.writeUInt32BE(data.length, 0);
  const td = Buffer.concat([Buffer.from(tag, 'ascii'), data]);
  const crc = Buffer.alloc(4); crc.writeUInt32BE(crc32(td), 0);
  return Buffer.concat([len, td, crc]);
}
function encodePng(rgba, w, h) {
  const raw = Buffer.alloc((w * 4 + 1) * h);
  for (let y = 0; y < h; y++) {
    const src = (h - 1 - y) * w * 4;
    raw[y * (w * 4 + 1)] = 0;
    Buffer.from(rgba.buffer, rgba.byteOffset + src, w * 4).copy(raw, y * (w * 4 + 1) + 1);
  }
  const ihdr = Buffer.alloc(13);
  ihdr.writeUInt32BE(w, 0); ihdr.writeUInt32BE(h, 4);
  ihdr[8] = 8; ihdr[9] = 6; ihdr[10] = 0; ihdr[11] = 0; ihdr[12] = 0;
  return Buffer.concat([Buffer.from([0x89, 0x50, 0x4e, 0x47, 0x0d, 0x0a, 0x1a, 0x0a]), pngChunk('IHDR', ihdr),
                        pngChunk('IDAT', require('zlib').deflateSync(raw)), pngChunk('IEND', Buffer.alloc(0))]);
}

const POLICIES = { rr: 0, fifo: 1, srpt: 2 };
const ERR_MEMORY = -4;

function errorCode(e) {
  const m = /dspt error (-?\d+)/.exec(e && e.message ? e.message : '');
  return m ? parseInt(m[1], 10) : 0;
}

// Continuous-batching render server over scenes kept resident in HBM (no
// counterpart in the reference, which renders one scene at a time). Requests
// are admitted while the engine's memory budget allows (a refused request
// waits for a running one to finish), share the engine's path pool with the
// requests already running, report per-frame progress and resolve with the
// linear RGBA32F accumulation of frames 1..spp. One loop drives the engine;
// the JS thread stays free while the GPU works (jobsStep is asynchronous).
class RenderServer {
  constructor({ materialProfile = 'webgl-full', device = 0, memoryBudget = 0, policy = 'rr', gpuBvh = false } = {}) {
    this._h = native.create(materialProfile === 'webgl-full' ? 1 : 0, device);
    this._gpuBvh = !!gpuBvh;  // build scene BVHs on the GPU instead of the host
    if (memoryBudget) native.setMemoryBudget(this._h, memoryBudget);
    if (!(policy in POLICIES)) throw new Error(`unknown policy ${policy}`);
    native.setPolicy(this._h, POLICIES[policy]);
    this._scenes = new Map();
    this._ibls = new Map();
    this._pending = [];
    this._live = new Map();
    this._loop = null;
  }

  addScene(scene) {
    if (scene.instances) {
      // two-level scene: object-space meshes placed by instances (engine extension)
      const mats = new Float32Array(scene.materials.length * 68);
      scene.materials.forEach((m, i) => mats.set(m.data, i * 68));
      const inst = scene.instances;
      const id = native.sceneCreateInstanced(this._h, scene.vertexBuffer, scene.triangleIndexBuffer, inst.meshFirst,
                                             inst.meshCount, inst.objectToWorld, inst.mesh, mats);
      if (scene.textures) native.sceneSetTextures(this._h, id, scene.textures.texInfos, scene.textures.arrayDims, scene.textures.texels);
      this._scenes.set(id, { counts: [scene.triangleIndexBuffer.length / 3, 0, 0], lights: scene.lights || [] });
      return id;
    }
    let nodes = scene.bvhNodes, idx = scene.bvhTriangleIndices;
    const mats = new Float32Array(scene.materials.length * 68);
    scene.materials.forEach((m, i) => mats.set(m.data, i * 68));
    if (!nodes && this._gpuBvh) {
      // BVH built on the GPU (PLOC): milliseconds, for admitting scenes while serving
      const r = native.sceneCreateGpuBvh(this._h, scene.vertexBuffer, scene.triangleIndexBuffer, mats, 4, 1);
      if (scene.textures) native.sceneSetTextures(this._h, r.id, scene.textures.texInfos, scene.textures.arrayDims, scene.textures.texels);
      const nt = scene.triangleIndexBuffer.length / 3;
      this._scenes.set(r.id, { counts: [nt, 0, nt], lights: scene.lights || [] });
      return r.id;
    }
    if (!nodes) {
      const r = native.buildBvh(scene.bvhPositionBuffer, scene.bvhIndexBuffer, 4);
      nodes = r.nodeData; idx = r.triangleIndices;
    }
    const id = native.sceneCreate(this._h, scene.vertexBuffer, scene.triangleIndexBuffer, nodes, idx, mats);
    if (scene.textures) native.sceneSetTextures(this._h, id, scene.textures.texInfos, scene.textures.arrayDims, scene.textures.texels);
    this._scenes.set(id, { counts: [scene.triangleIndexBuffer.length / 3, nodes.length / 16, idx.length],
                           lights: scene.lights || [] });
    return id;
  }

  addIbl(texture) {
    const img = texture.image;
    const r = native.iblCreate(this._h, img.data, img.width, img.height);
    this._ibls.set(r.id, { size: [img.width, img.height], totalSum: r.totalSum });
    return r.id;
  }

  // req: { scene, ibl (id or -1), camera, width, height, spp, maxBounces, useIBL, showBackground,
  //        backgroundColor, iblRotation (degrees), rayEps, clampThreshold, onProgress(done, total) }
  submit(req) {
    return new Promise((resolve, reject) => {
      this._pending.push({ req, resolve, reject, arrival: Date.now() });
      if (!this._loop) this._loop = this._run().finally(() => { this._loop = null; });
    });
  }

  _block(req) {
    const sc = this._scenes.get(req.scene);
    const ib = req.ibl >= 0 ? this._ibls.get(req.ibl) : { size: [0, 0], totalSum: 0 };
    const useIBL = req.useIBL !== undefined ? req.useIBL : req.ibl >= 0;
    const settings = {
      _backgroundColor: req.backgroundColor || [0, 0, 0, 1], renderRes: [req.width, req.height],
      _iblSize: ib.size, _useIBL: useIBL,
      _iblRotation: (req.iblRotation || 0) / 180.0 * Math.PI,
      _showBackground: req.showBackground !== undefined ? req.showBackground : useIBL,
      _maxBounces: req.maxBounces || 8, _iblTotalSum: ib.totalSum,
      _rayEps: req.rayEps !== undefined ? req.rayEps : 0.0001,
      _clampThreshold: req.clampThreshold !== undefined ? req.clampThreshold : 3.0,
      _sceneCounts: sc.counts, _scene: { lights: sc.lights }, _debugMode: req.debugMode,
    };
    return packFrameBlock(settings, req.camera, 1);
  }

  async _run() {
    while (this._pending.length || this._live.size) {
      // admission in arrival order while memory allows
      while (this._pending.length) {
        const p = this._pending[0];
        let id;
        try {
          id = native.jobCreate(this._h, p.req.scene, p.req.ibl, this._block(p.req), p.req.width, p.req.height, p.req.spp);
        } catch (e) {
          if (errorCode(e) === ERR_MEMORY && this._live.size) break;
          this._pending.shift();
          p.reject(e);
          continue;
        }
        this._pending.shift();
        this._live.set(id, p);
      }
      if (!this._live.size) continue;
      await native.jobsStep(this._h, Int32Array.from(this._live.keys()));
      for (const [id, p] of Array.from(this._live)) {
        const [done, total] = native.jobStatus(this._h, id);
        if (p.req.onProgress) p.req.onProgress(done, total);
        if (done >= total) {
          const out = new Float32Array(p.req.width * p.req.height * 4);
          native.jobReadAccum(this._h, id, out);
          native.jobDestroy(this._h, id);
          this._live.delete(id);
          p.resolve({ accumulation: out, width: p.req.width, height: p.req.height, frames: total,
                      latencyMs: Date.now() - p.arrival });
        }
      }
    }
  }

  memoryUsage() { return native.memoryUsage(this._h); }
}

module.exports = { PathtracingRenderer, RenderServer, packFrameBlock, tileOrder, tileRect, native };
