'use strict';
// Serves requests over a scene dumped by tests/test_node_addon.py through the
// Node RenderServer and writes each request's accumulation buffer.
const fs = require('fs');
const path = require('path');
const { RenderServer } = require(path.join(__dirname, '..', 'addon', 'pathtracing_renderer.js'));

function load(dir, name, T) {
  const b = fs.readFileSync(path.join(dir, name));
  return new T(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

(async () => {
  const dir = process.argv[2];
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const server = new RenderServer({ materialProfile: meta.profile, policy: meta.policy, gpuBvh: !!meta.gpuBvh });
  const mats = load(dir, 'mats.f32', Float32Array);
  const sceneData = {
    vertexBuffer: load(dir, 'verts.f32', Float32Array),
    triangleIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    bvhPositionBuffer: load(dir, 'bvhpos.f32', Float32Array),
    bvhIndexBuffer: load(dir, 'tris.u32', Uint32Array),
    materials: Array.from({ length: mats.length / 68 }, (_, i) => ({ data: mats.subarray(i * 68, (i + 1) * 68) })),
    lights: [],
  };
  if (meta.instanced) {
    sceneData.instances = {
      meshFirst: load(dir, 'meshfirst.i32', Int32Array), meshCount: load(dir, 'meshcount.i32', Int32Array),
      objectToWorld: load(dir, 'o2w.f32', Float32Array), mesh: load(dir, 'instmesh.i32', Int32Array),
    };
  }
  const scene = server.addScene(sceneData);
  const ibl = server.addIbl({ image: { data: load(dir, 'ibl.f32', Float32Array), width: meta.iblWidth, height: meta.iblHeight } });
  const camera = { fov: meta.fov, near: meta.near, matrixWorld: { elements: meta.matrixWorld }, position: meta.position };
  const progress = meta.requests.map(() => 0);
  const done = await Promise.all(meta.requests.map((q, i) => new Promise((res) => setTimeout(res, q.delayMs)).then(() =>
    server.submit({ scene, ibl, camera, width: q.width, height: q.height, spp: q.spp, maxBounces: meta.maxBounces,
                    onProgress: (d) => { progress[i] = d; } }))));
  done.forEach((r, i) => fs.writeFileSync(path.join(dir, `accum${i}.f32`), Buffer.from(r.accumulation.buffer)));
  console.log(JSON.stringify({ frames: done.map((r) => r.frames), progress, latencyMs: done.map((r) => r.latencyMs) }));
})().catch((e) => { console.error(e.message); process.exit(1); });
