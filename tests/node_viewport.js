'use strict';
// Drives the Node PathtracingRenderer drop-in with the call sequence of the
// reference's PathtracingViewport (packages/lib/pathtracing_viewport.ts:262-298):
// start(sampleLimit = -1) calls renderer.render(camera, sampleLimit, tileCB,
// frameCB, finishedCB); stop() calls renderer.stopRendering();
// resetAccumulation() after a camera move; setInteractionMode(). Writes each
// phase's accumulation and frame count for tests/test_node_addon.py.
const fs = require('fs');
const path = require('path');
const { PathtracingRenderer } = require(path.join(__dirname, '..', 'addon', 'pathtracing_renderer.js'));

function load(dir, name, T) {
  const b = fs.readFileSync(path.join(dir, name));
  return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}
const sleep = (ms) => new Promise((r) => setTimeout(r, ms));

// the viewport's own start/stop/reset logic, restated (VP:262-298)
class Viewport {
  constructor(renderer, camera) {
    this.renderer = renderer; this.camera = camera; this.running = false;
    this.events = { frameRendered: [], renderFinished: [], diagnostics: 0 };
  }
  emitDiagnostics() { this.events.diagnostics++; }
  start(sampleLimit = -1) {
    if (!this.camera) throw new Error('Cannot start rendering without a camera.');
    this.running = true;
    this.sampleLimit = sampleLimit;
    return this.renderer.render(
      this.camera, sampleLimit,
      () => this.emitDiagnostics(),
      (frameCount) => { this.events.frameRendered.push(frameCount); this.emitDiagnostics(); },
      () => { this.running = false; this.events.renderFinished.push(this.renderer.diagnostics.sampleCount); },
    );
  }
  stop() { this.running = false; this.renderer.stopRendering(); this.emitDiagnostics(); }
  resetAccumulation() { this.renderer.resetAccumulation(); this.emitDiagnostics(); }
  setInteractionMode(active) { this.renderer.setInteractionMode(active); this.emitDiagnostics(); }
}

async function waitFor(pred, ms = 20000) {
  const t0 = Date.now();
  while (!pred()) {
    if (Date.now() - t0 > ms) throw new Error('timed out waiting for frames');
    await sleep(1);
  }
}

(async () => {
  const dir = process.argv[2];
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const r = new PathtracingRenderer({ materialProfile: meta.profile, width: meta.width, height: meta.height });
  r.maxBounces = meta.maxBounces;
  const mats = load(dir, 'mats.f32', Float32Array);
  await r.setScene({
    vertexBuffer: load(dir, 'verts.f32', Float32Array),
    triangleIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    bvhPositionBuffer: load(dir, 'bvhpos.f32', Float32Array),
    bvhIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    materials: Array.from({ length: mats.length / 68 }, (_, i) => ({ data: mats.subarray(i * 68, (i + 1) * 68) })),
    lights: [],
  });
  r.setIBL({ image: { data: load(dir, 'ibl.f32', Float32Array), width: meta.iblWidth, height: meta.iblHeight } });
  const cams = meta.cameras.map((c) => ({ fov: c.fov, near: c.near, matrixWorld: { elements: c.matrixWorld.slice() },
                                          position: Object.assign({}, c.position) }));
  const camera = cams[0];
  const vp = new Viewport(r, camera);
  const out = {};
  const save = (name) => {
    fs.writeFileSync(path.join(dir, `${name}.f32`), Buffer.from(r.readAccumulation().buffer));
    out[name] = { sampleCount: r.diagnostics.sampleCount, frameEvents: vp.events.frameRendered.slice(),
                  finished: vp.events.renderFinished.slice(), mode: r.diagnostics.mode };
  };

  // 1. the default start(): renders until stopped
  vp.start();
  await waitFor(() => vp.events.frameRendered.length >= meta.framesBeforeStop);
  vp.stop();
  const atStop = vp.events.frameRendered.length;
  await sleep(100);
  if (vp.events.frameRendered.length !== atStop) throw new Error('frame callbacks after stop()');
  await r.idle();
  save('unbounded');
  vp.events.frameRendered = [];

  // 2. a camera move during the render: the viewport resets the accumulation
  vp.start();
  await waitFor(() => vp.events.frameRendered.length >= 3);
  camera.matrixWorld.elements = cams[1].matrixWorld.elements.slice();
  Object.assign(camera.position, cams[1].position);
  vp.resetAccumulation();
  const mark = vp.events.frameRendered.length;
  await waitFor(() => vp.events.frameRendered.slice(mark).filter((f) => f > 1).length >= 4);
  vp.stop();
  await r.idle();
  save('moved');
  vp.events.frameRendered = [];

  // 3. start() while a render runs stops it; a sample limit ends the loop
  vp.start();
  await waitFor(() => vp.events.frameRendered.length >= 2);
  vp.events.frameRendered = [];
  await vp.start(meta.sampleLimit);
  await r.idle();
  save('limited');

  // 4. interaction mode during a render: restarts, one tile per frame
  let tiles = 0, frames = 0;
  const p = r.render(camera, 6, () => { tiles++; }, () => { frames++; }, null);
  vp.setInteractionMode(true);
  await p;
  out.interactive = { tiles, frames, sampleCount: r.diagnostics.sampleCount, mode: r.diagnostics.mode };
  vp.setInteractionMode(false);

  // 5. debug views: the Normals view, then a name not in debugModes (packs -1)
  for (const [mode, name] of [['Normals', 'debug_normals'], ['Wireframe', 'debug_unknown']]) {
    vp.events.frameRendered = []; vp.events.renderFinished = [];
    r.debugMode = mode;
    await vp.start(meta.sampleLimit);
    await r.idle();
    save(name);
  }
  r.debugMode = 'None';
  out.lists = { debugModes: r.debugModes.length, tonemappingModes: r.tonemappingModes };
  console.log(JSON.stringify(out));
})().catch((e) => { console.error(e.stack || e.message); process.exit(1); });
