// Shading-point set-up shared by the shading kernel and the CPU build of
// the shading code (tests/devhost): render state from the hit triangle
// (renderer.frag:72-91, bvh.glsl:91-134) and the glTF material closure set-up
// (material.glsl:5-110) for untextured materials.
#pragma once
#include "dev_scene.hpp"

struct Hitpoint {
  f3 pos, ng_raw, n, t;
  float tw;
  f3 vc;
  float vcw;
  uint32_t mat;
};

// fillRenderState (renderer.frag:72-91) minus the material.
DI Hitpoint fetch_hitpoint(const SceneView& S, f3 org, f3 dir, int tri, float t, float u, float v) {
  Hitpoint hp;
  hp.pos = org + dir * t;
  uint32_t vi[3];
  for (int k = 0; k < 3; k++) vi[k] = S.tris[(size_t)tri * 3 + k];
  const float4* v0 = S.verts + (size_t)vi[0] * 5;
  const float4* v1 = S.verts + (size_t)vi[1] * 5;
  const float4* v2 = S.verts + (size_t)vi[2] * 5;
  float4 p0 = v0[0], p1 = v1[0], p2 = v2[0];
  hp.ng_raw = nrm3(cross3(xyz(p1) - xyz(p0), xyz(p2) - xyz(p0)));
  float w0 = 1.0f - u - v;
  hp.n = nrm3(w0 * xyz(v0[1]) + u * xyz(v1[1]) + v * xyz(v2[1]));
  float4 t0 = v0[3], t1 = v1[3], t2 = v2[3];
  float hand = (t0.w == t1.w && t0.w == t2.w) ? t0.w : 0.0f;
  f3 tn = nrm3(w0 * xyz(t0) + u * xyz(t1) + v * xyz(t2));
  if (len3(tn) > 0.99f && fabsf(hand) > 0.99f) {
    hp.t = tn; hp.tw = -hand;
  } else {
    // get_onb(n)[0], utils.glsl:104-109
    hp.t = nrm3(cross3(hp.n, f3{-hp.n.z, hp.n.x, -hp.n.y}));
    hp.tw = 1.0f;
  }
  float4 c0 = v0[4], c1 = v1[4], c2 = v2[4];
  hp.vc = w0 * xyz(c0) + u * xyz(c1) + v * xyz(c2);
  hp.vcw = w0 * c0.w + u * c1.w + v * c2.w;
  hp.mat = (uint32_t)p0.w;
  return hp;
}

struct Shading {
  bsdf::GltfInputs in;
  f3 n, ng;          // flipped toward wi (fix_normals)
  f3 tanA;           // anisotropy tangent
  bool backside, thin;
  float cutout;
  f3 emissive, attColor;
  float attDist;
};

// configure_gltf_material (material.glsl:5-110) for untextured materials.
DI Shading configure(const SceneView& S, const Hitpoint& hp, f3 wi) {
  const float4* m = S.mats + (size_t)hp.mat * 17;
  float4 m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7], m8 = m[8], m9 = m[9],
         m10 = m[10], m15 = m[15], m16 = m[16];
  Shading sh;
  bsdf::GltfInputs& g = sh.in;
  g.baseColor = xyz(m0);
  float opacity = 1.0f;
  float l4 = sqrtf(hp.vc.x * hp.vc.x + hp.vc.y * hp.vc.y + hp.vc.z * hp.vc.z + hp.vcw * hp.vcw);
  if (l4 > 0.0f) { g.baseColor = g.baseColor * hp.vc; opacity = opacity * hp.vcw; }
  sh.cutout = m2.x * opacity;
  if (m6.w > 0.0f) sh.cutout = (sh.cutout < m6.w) ? 0.0f : 1.0f;
  if (m6.w == 1.0f) sh.cutout = 1.0f;
  g.metallic = m0.w * 1.0f;
  g.roughness = m1.x * 1.0f;
  g.anisotropy = m1.y * 1.0f;
  g.transmission = m1.w * 1.0f;
  g.ior = m2.w;
  g.specularColor = xyz(m3);
  g.specular = m3.w * 1.0f;
  g.sheenColor = xyz(m4);
  g.sheenRoughness = m4.w * 1.0f;
  g.clearcoat = m6.x * 1.0f;
  g.clearcoatRoughness = m6.y * 1.0f;
  g.diffuseTransmission = m6.z * 1.0f;
  g.diffuseTransmissionColor = xyz(m15);
  g.iridescence = m10.x * 1.0f;
  g.iridescenceIor = m10.y;
  g.iridescenceThickness = gmix(m10.z, m10.w, 1.0f);
  g.dispersion = m16.x;
  sh.thin = m8.w > 0.5f;
  g.thin = sh.thin;
  sh.emissive = xyz(m5);
  sh.attColor = xyz(m7);
  sh.attDist = m7.w;
  // fix_normals (utils.glsl:88-98)
  f3 n = hp.n, ng = hp.ng_raw;
  sh.backside = false;
  if (dot3(wi, ng) < 0.0f) { ng = -ng; sh.backside = true; }
  if (dot3(ng, n) < 0.0f) n = -n;
  sh.n = n; sh.ng = ng;
  // anisotropy direction -> tangent rotation by atan2 + pi (material.glsl:74-83)
  float rot = atan2f(m9.y, m9.x);
  float angle = rot + kPI;
  f3 t = hp.t;
  if (angle > 0.0f) {
    f3 b = cross3(n, hp.t) * hp.tw;
    t = hp.t * cosf(angle) + b * sinf(angle);
  }
  sh.tanA = t;
  return sh;
}

