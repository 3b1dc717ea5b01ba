'use strict';
// Renders a scene dumped by tests/test_node_addon.py through the Node
// PathtracingRenderer drop-in and writes the accumulation buffer.
const fs = require('fs');
const path = require('path');
const { PathtracingRenderer } = require(path.join(__dirname, '..', 'addon', 'pathtracing_renderer.js'));

function load(dir, name, T) {
  const b = fs.readFileSync(path.join(dir, name));
  return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

(async () => {
  const dir = process.argv[2];
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const r = new PathtracingRenderer({ materialProfile: meta.profile, width: meta.width, height: meta.height,
                                      tileSchedule: meta.tileSchedule || 'frames' });
  if (meta.tileRes) r.tileRes = meta.tileRes;
  r.maxBounces = meta.maxBounces;
  const mats = load(dir, 'mats.f32', Float32Array);
  const scene = {
    vertexBuffer: load(dir, 'verts.f32', Float32Array),
    triangleIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    bvhPositionBuffer: load(dir, 'bvhpos.f32', Float32Array),
    bvhIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    materials: Array.from({ length: mats.length / 68 }, (_, i) => ({ data: mats.subarray(i * 68, (i + 1) * 68) })),
    lights: [],
  };
  await r.setScene(scene);
  r.setIBL({ image: { data: load(dir, 'ibl.f32', Float32Array), width: meta.iblWidth, height: meta.iblHeight } });
  const camera = { fov: meta.fov, near: meta.near, matrixWorld: { elements: meta.matrixWorld }, position: meta.position };
  let frames = 0, tiles = 0;
  const tilesPerFrame = [];
  await r.render(camera, meta.numSamples, () => { tiles++; }, (n) => { frames = n; tilesPerFrame.push(tiles); tiles = 0; },
                 null, meta.framesPerCall);
  fs.writeFileSync(path.join(dir, 'accum.f32'), Buffer.from(r.readAccumulation().buffer));
  // the display chain with the reference's settings names
  r.tonemapping = 'AcesFilm'; r.enableFxaa = true;
  fs.writeFileSync(path.join(dir, 'display.u8'), Buffer.from(r.readDisplay().buffer));
  r.saveImage(path.join(dir, 'display.png'));
  console.log(JSON.stringify({ frames, tilesPerFrame, sampleCount: r.diagnostics.sampleCount, bvh: r.diagnostics.bvh }));
})().catch((e) => { console.error(e.message); process.exit(1); });
