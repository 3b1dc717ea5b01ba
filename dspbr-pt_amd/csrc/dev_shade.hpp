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
  float2 uv0;  // compute_interpolated_uv(set 0), bvh.glsl:108-115
  uint32_t mat;
};

// fillRenderState (renderer.frag:72-91) minus the material. For a hit on an
// instance of a two-level scene (inst >= 0) the mesh's object-space
// attributes are taken to world space: positions with object->world,
// normals with the transpose of world->object, tangents with object->world;
// the hit position is on the world ray as in the one-level case.
DI Hitpoint fetch_hitpoint(const SceneView& S, f3 org, f3 dir, int tri, float t, float u, float v, int inst = -1) {
  Hitpoint hp;
  hp.pos = org + dir * t;
  const float4 *v0, *v1, *v2;
  if (S.tri_verts) {  // the same records, stored per triangle
    v0 = S.tri_verts + (size_t)tri * 15;
    v1 = v0 + 5;
    v2 = v0 + 10;
  } else {
    uint32_t vi[3];
    for (int k = 0; k < 3; k++) vi[k] = S.tris[(size_t)tri * 3 + k];
    v0 = S.verts + (size_t)vi[0] * 5;
    v1 = S.verts + (size_t)vi[1] * 5;
    v2 = S.verts + (size_t)vi[2] * 5;
  }
  float4 p0 = v0[0], p1 = v1[0], p2 = v2[0];
  const InstanceDesc* D =
      (inst >= 0 && S.inst_base >= 0) ? reinterpret_cast<const InstanceDesc*>(S.nodes + (size_t)(S.inst_base + 2 * inst) * 4)
                                      : nullptr;
  float4 o2w[3], w2o[3];
  if (D) {
    for (int k = 0; k < 3; k++) { o2w[k] = D->o2w[k]; w2o[k] = D->w2o[k]; }
  }
  f3 q0 = xyz(p0), q1 = xyz(p1), q2 = xyz(p2);
  if (D) { q0 = xf_point(o2w, q0); q1 = xf_point(o2w, q1); q2 = xf_point(o2w, q2); }
  hp.ng_raw = nrm3(cross3(q1 - q0, q2 - q0));
  float w0 = 1.0f - u - v;
  f3 ni = w0 * xyz(v0[1]) + u * xyz(v1[1]) + v * xyz(v2[1]);
  if (D) ni = xf_tvec(w2o, ni);
  hp.n = nrm3(ni);
  float4 t0 = v0[3], t1 = v1[3], t2 = v2[3];
  float hand = (t0.w == t1.w && t0.w == t2.w) ? t0.w : 0.0f;
  f3 ti = w0 * xyz(t0) + u * xyz(t1) + v * xyz(t2);
  if (D) ti = xf_tvec(w2o, ti);  // tangents by the normal matrix too, as scene_loader.ts:327-336 bakes them
  f3 tn = nrm3(ti);
  if (len3(tn) > 0.99f && fabsf(hand) > 0.99f) {
    hp.t = tn; hp.tw = -hand;
  } else {
    // get_onb(n)[0], utils.glsl:104-109
    hp.t = nrm3(cross3(hp.n, f3{-hp.n.z, hp.n.x, -hp.n.y}));
    hp.tw = 1.0f;
  }
  float4 a0 = v0[2], a1 = v1[2], a2 = v2[2];
  hp.uv0 = make_float2(w0 * a0.x + u * a1.x + v * a2.x, w0 * a0.y + u * a1.y + v * a2.y);
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

DI f3 to_linear(float4 t) { return f3{pm::ppow(t.x, 2.2f), pm::ppow(t.y, 2.2f), pm::ppow(t.z, 2.2f)}; }

// Material texture slots in the order configure_gltf_material reads them.
enum TexSlot {
  TS_BASE = 0, TS_TRANSMISSION, TS_DIFFTRANS, TS_DIFFTRANS_COLOR, TS_METAL_ROUGH, TS_ANISO, TS_SPEC_COLOR, TS_SPEC,
  TS_SHEEN_COLOR, TS_SHEEN_ROUGH, TS_NORMAL, TS_ANISO_DIR, TS_EMISSIVE, TS_CLEARCOAT, TS_CLEARCOAT_ROUGH, TS_IRID,
  TS_IRID_THICK, TS_COUNT
};
DI float tex_id(float4 m9, float4 m11, float4 m12, float4 m13, float4 m14, float4 m15, int k) {
  switch (k) {
    case TS_BASE: return m11.x;
    case TS_TRANSMISSION: return m12.z;
    case TS_DIFFTRANS: return m9.w;
    case TS_DIFFTRANS_COLOR: return m15.w;
    case TS_METAL_ROUGH: return m11.y;
    case TS_ANISO: return m14.x;
    case TS_SPEC_COLOR: return m12.y;
    case TS_SPEC: return m12.x;
    case TS_SHEEN_COLOR: return m13.z;
    case TS_SHEEN_ROUGH: return m13.w;
    case TS_NORMAL: return m11.z;
    case TS_ANISO_DIR: return m14.y;
    case TS_EMISSIVE: return m11.w;
    case TS_CLEARCOAT: return m12.w;
    case TS_CLEARCOAT_ROUGH: return m13.x;
    case TS_IRID: return m14.z;
    default: return m14.w;  // TS_IRID_THICK
  }
}

// configure_gltf_material (material.glsl:5-110). Every factor starts as its
// untextured value (texture value vec4(1): to_linear_rgb(1) = 1, x * 1 = x,
// mix(a, b, 1) = b); a material with textures then replaces the factors its
// textures modulate, through one sampling site for all slots.
DI Shading configure(const SceneView& S, const Hitpoint& hp, f3 wi) {
  const float4* m = S.mats + (size_t)hp.mat * 17;
  float4 m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7], m8 = m[8], m9 = m[9],
         m10 = m[10], m15 = m[15], m16 = m[16];
  Shading sh;
  bsdf::GltfInputs& g = sh.in;
  g.baseColor = xyz(m0);
  float opacity = 1.0f;
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
  sh.emissive = xyz(m5);
  f3 n = hp.n, t = hp.t;
  bool aniso_tex = false, aniso_dir_tex = false;
  float2 aniso_rg = make_float2(0.0f, 0.0f);
  f3 aniso_dir = sp3(0.0f);
  if (S.num_tex > 0) {
    const float4 m11 = m[11], m12 = m[12], m13 = m[13], m14 = m[14];
    for (int k = 0; k < TS_COUNT; k++) {
      const float id = tex_id(m9, m11, m12, m13, m14, m15, k);
      if (!(id >= 0.0f)) continue;
      const float4 tv = tex_value(S, id, hp.uv0);
      switch (k) {
        case TS_BASE: g.baseColor = xyz(m0) * to_linear(tv); opacity = tv.w; break;
        case TS_TRANSMISSION: g.transmission = m1.w * tv.x; break;
        case TS_DIFFTRANS: g.diffuseTransmission = m6.z * tv.x; break;
        case TS_DIFFTRANS_COLOR: g.diffuseTransmissionColor = xyz(m15) * to_linear(tv); break;
        case TS_METAL_ROUGH: g.metallic = m0.w * tv.z; g.roughness = m1.x * tv.y; break;
        case TS_ANISO: g.anisotropy = m1.y * tv.z; aniso_tex = true; aniso_rg = make_float2(tv.x, tv.y); break;
        case TS_SPEC_COLOR: g.specularColor = xyz(m3) * to_linear(tv); break;
        case TS_SPEC: g.specular = m3.w * tv.w; break;
        case TS_SHEEN_COLOR: g.sheenColor = xyz(m4) * to_linear(tv); break;
        case TS_SHEEN_ROUGH: g.sheenRoughness = m4.w * tv.x; break;
        case TS_NORMAL: {
          // to_world = get_onb(n, t) (utils.glsl:111-115); n' = normalize(tex * 2 - 1),
          // scaled in xy; then the tangent is re-orthonormalised
          const f3 bb = nrm3(cross3(n, t));
          const f3 tt = cross3(bb, n);
          f3 nt = nrm3(f3{tv.x * 2.0f - 1.0f, tv.y * 2.0f - 1.0f, tv.z * 2.0f - 1.0f});
          nt = nrm3(nt * f3{m2.z, m2.z, 1.0f});
          n = tt * nt.x + bb * nt.y + n * nt.z;
          const f3 b2 = nrm3(cross3(n, t)) * hp.tw;
          t = cross3(b2, n);
          break;
        }
        case TS_ANISO_DIR: aniso_dir_tex = true; aniso_dir = f3{tv.x * 2.0f - 1.0f, tv.y * 2.0f - 1.0f, tv.z * 2.0f - 1.0f}; break;
        case TS_EMISSIVE: sh.emissive = xyz(m5) * to_linear(tv); break;
        case TS_CLEARCOAT: g.clearcoat = m6.x * tv.x; break;
        case TS_CLEARCOAT_ROUGH: g.clearcoatRoughness = m6.y * tv.x; break;
        case TS_IRID: g.iridescence = m10.x * tv.x; break;
        default: g.iridescenceThickness = gmix(m10.z, m10.w, tv.y); break;  // TS_IRID_THICK
      }
    }
  }
  float l4 = sqrtf(hp.vc.x * hp.vc.x + hp.vc.y * hp.vc.y + hp.vc.z * hp.vc.z + hp.vcw * hp.vcw);
  if (l4 > 0.0f) { g.baseColor = g.baseColor * hp.vc; opacity = opacity * hp.vcw; }
  sh.cutout = m2.x * opacity;
  if (m6.w > 0.0f) sh.cutout = (sh.cutout < m6.w) ? 0.0f : 1.0f;
  if (m6.w == 1.0f) sh.cutout = 1.0f;
  sh.thin = m8.w > 0.5f;
  g.thin = sh.thin;
  sh.attColor = xyz(m7);
  sh.attDist = m7.w;
  // fix_normals (utils.glsl:88-98)
  f3 ng = hp.ng_raw;
  sh.backside = false;
  if (dot3(wi, ng) < 0.0f) { ng = -ng; sh.backside = true; }
  if (dot3(ng, n) < 0.0f) n = -n;
  sh.n = n; sh.ng = ng;
  // anisotropy direction -> tangent rotation by atan2 + pi (material.glsl:74-83)
  float ax = m9.x, ay = m9.y;
  if (aniso_dir_tex) { ax = aniso_dir.x; ay = aniso_dir.y; }
  else if (aniso_tex) { ax = aniso_rg.x * 2.0f - 1.0f; ay = aniso_rg.y * 2.0f - 1.0f; }
  float rot = pm::patan2(ay, ax);
  float angle = rot + kPI;
  if (angle > 0.0f) {
    f3 b = cross3(n, t) * hp.tw;
    float sa, ca;
    pm::psincos(angle, sa, ca);
    t = t * ca + b * sa;
  }
  sh.tanA = t;
  return sh;
}
