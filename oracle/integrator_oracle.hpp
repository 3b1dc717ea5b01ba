// CPU ORACLE — test infrastructure only (see pt_oracle.hpp).
//
// Restatement of the per-pixel path sampler: material decode and glTF
// closure set-up (SH/material.glsl, SH/pbr_material_adapter.glsl,
// packages/lib/scene_data_webgl2.ts:163-210), render state
// (SH/renderer.frag:72-91), BSDF bounce, Russian roulette, camera ray,
// next-event estimation and the `misptdl` trace loop
// (SH/integrator/misptdl.glsl:2-93), and the running-mean accumulation
// (SH/renderer.frag:146-157).
#pragma once
#include "bsdf_oracle.hpp"

namespace orc {

// Decoded MaterialData (SH/structs.glsl:3-69), get_material accessor of
// packages/lib/scene_data_webgl2.ts:173-207. Texture ids are kept to detect
// textured materials (not supported by the oracle: textures return 1.0 only
// when id < 0).
struct MatData {
  float baseColor[3], metallic, roughness, anisotropy, anisotropyRotation, transmission;
  float cutoutOpacity; bool doubleSided; float normalScale, ior;
  float specularColor[3], specular;
  float sheenColor[3], sheenRoughness;
  float emissive[3], clearcoatNormalScale;
  float clearcoat, clearcoatRoughness, diffuseTransmission, alphaCutoff;
  float attenuationColor[3], attenuationDistance;
  float multiscatterColor[3]; bool thinWalled;
  float anisotropyDirection[3], diffuseTransmissionTextureId;
  float iridescence, iridescenceIor, iridescenceThicknessMin, iridescenceThicknessMax;
  float texIds[16];
  float diffuseTransmissionColor[3], diffuseTransmissionColorTextureId;
  float dispersion;
};
inline MatData decode_material(const float* m) {
  MatData d;
  for (int i = 0; i < 3; i++) {
    d.baseColor[i] = m[i]; d.specularColor[i] = m[12 + i]; d.sheenColor[i] = m[16 + i];
    d.emissive[i] = m[20 + i]; d.attenuationColor[i] = m[28 + i]; d.multiscatterColor[i] = m[32 + i];
    d.anisotropyDirection[i] = m[36 + i]; d.diffuseTransmissionColor[i] = m[60 + i];
  }
  d.metallic = m[3]; d.roughness = m[4]; d.anisotropy = m[5]; d.anisotropyRotation = m[6]; d.transmission = m[7];
  d.cutoutOpacity = m[8]; d.doubleSided = m[9] > 0.5f; d.normalScale = m[10]; d.ior = m[11];
  d.specular = m[15]; d.sheenRoughness = m[19]; d.clearcoatNormalScale = m[23];
  d.clearcoat = m[24]; d.clearcoatRoughness = m[25]; d.diffuseTransmission = m[26]; d.alphaCutoff = m[27];
  d.attenuationDistance = m[31]; d.thinWalled = m[35] > 0.5f; d.diffuseTransmissionTextureId = m[39];
  d.iridescence = m[40]; d.iridescenceIor = m[41]; d.iridescenceThicknessMin = m[42]; d.iridescenceThicknessMax = m[43];
  for (int i = 0; i < 16; i++) d.texIds[i] = m[44 + i];
  d.diffuseTransmissionColorTextureId = m[63];
  d.dispersion = m[64];
  return d;
}

template <class R> struct MClosure {  // MaterialClosure, SH/structs.glsl:80-91
  PbrMat<R> material;
  R cutout_opacity;
  bool thin_walled, double_sided, backside;
  V3<R> n, ng;
  V3<R> t; R tw;
  V3<R> anisotropyTangent;
};
template <class R> struct RState {  // RenderState, SH/structs.glsl:93-103
  V3<R> hitPos, n, ng;
  V3<R> tangent; R tangent_w;
  R uv0x = R(0), uv0y = R(0);  // compute_interpolated_uv(set 0)
  V3<R> wo, wi;
  MClosure<R> closure;
};

// get_texture_value(id, uv0, uv1), scene_data_webgl2.ts:343-357
template <class R> inline V4<R> gtv(const Scene& s, float id, const RState<R>& rs) {
  if (id < 0.0f) return V4<R>{R(1), R(1), R(1), R(1)};
  return tex_value<R>(s, id, rs.uv0x, rs.uv0y);
}
template <class R> inline V3<R> xyz4(const V4<R>& t) { return V3<R>(t.x, t.y, t.z); }

// SH/material.glsl:5-110. TexInfo ids: MatData.texIds[0..15] = slots 44..59
// (baseColor, metallicRoughness, normal, emissive, specular, specularColor,
// transmission, clearcoat, clearcoatRoughness, clearcoatNormal, sheenColor,
// sheenRoughness, anisotropy, anisotropyDirection, iridescence,
// iridescenceThickness), plus diffuseTransmission (slot 39) and
// diffuseTransmissionColor (slot 63).
template <class R> MClosure<R> configure_material(const Scene& s, uint32_t matIdx, const RState<R>& rs, V3<R> vcol, R vcol_w) {
  MatData md = decode_material(s.mats + (size_t)matIdx * 68);
  MClosure<R> c;
  PbrMat<R>& m = c.material;
  const R E22 = R(2.2f);
  const V4<R> bct = gtv(s, md.texIds[0], rs);
  m.baseColor = V3<R>(R(md.baseColor[0]), R(md.baseColor[1]), R(md.baseColor[2])) * vpow(xyz4(bct), E22);
  R opacity = bct.w;
  // length(vertexColor) > 0.0 on the vec4
  R l4 = std::sqrt(vcol.x * vcol.x + vcol.y * vcol.y + vcol.z * vcol.z + vcol_w * vcol_w);
  if (l4 > R(0)) { m.baseColor = m.baseColor * vcol; opacity = opacity * vcol_w; }
  m.opacity = opacity;
  c.cutout_opacity = R(md.cutoutOpacity) * opacity;
  if (md.alphaCutoff > 0.0f) c.cutout_opacity = (c.cutout_opacity < R(md.alphaCutoff)) ? R(0) : R(1);
  if (md.alphaCutoff == 1.0f) c.cutout_opacity = R(1);
  m.transmission = R(md.transmission) * gtv(s, md.texIds[6], rs).x;
  m.diffuseTransmission = R(md.diffuseTransmission) * gtv(s, md.diffuseTransmissionTextureId, rs).x;
  m.diffuseTransmissionColor = V3<R>(R(md.diffuseTransmissionColor[0]), R(md.diffuseTransmissionColor[1]), R(md.diffuseTransmissionColor[2])) *
                               vpow(xyz4(gtv(s, md.diffuseTransmissionColorTextureId, rs)), E22);
  c.thin_walled = md.thinWalled;
  m.ior = R(md.ior);
  c.double_sided = md.doubleSided;
  const V4<R> orm = gtv(s, md.texIds[1], rs);
  m.metallic = R(md.metallic) * orm.z;
  m.roughness = R(md.roughness) * orm.y;
  const V4<R> anis = gtv(s, md.texIds[12], rs);
  m.anisotropyStrength = R(md.anisotropy) * anis.z;
  m.specularColor = V3<R>(R(md.specularColor[0]), R(md.specularColor[1]), R(md.specularColor[2])) * vpow(xyz4(gtv(s, md.texIds[5], rs)), E22);
  m.specular = R(md.specular) * gtv(s, md.texIds[4], rs).w;
  const V4<R> shc = gtv(s, md.texIds[10], rs);
  const V4<R> shr = gtv(s, md.texIds[11], rs);
  m.sheenRoughness = R(md.sheenRoughness) * shr.x;
  m.sheenColor = V3<R>(R(md.sheenColor[0]), R(md.sheenColor[1]), R(md.sheenColor[2])) * vpow(xyz4(shc), E22);
  c.n = rs.n;
  c.ng = rs.ng;
  c.t = rs.tangent; c.tw = rs.tangent_w;
  if (md.texIds[2] >= 0.0f) {
    // to_world = get_onb(n, t), SH/utils.glsl:111-115
    const V3<R> b = normalize(cross(c.n, c.t));
    const V3<R> tt = cross(b, c.n);
    const V4<R> nm = gtv(s, md.texIds[2], rs);
    V3<R> nt = normalize(V3<R>(nm.x * R(2) - R(1), nm.y * R(2) - R(1), nm.z * R(2) - R(1)));
    nt = normalize(nt * V3<R>(R(md.normalScale), R(md.normalScale), R(1)));
    c.n = tt * nt.x + b * nt.y + c.n * nt.z;
    const V3<R> b2 = normalize(cross(c.n, c.t)) * c.tw;
    c.t = cross(b2, c.n);
  }
  c.backside = fix_normals(c.n, c.ng, rs.wi);
  V3<R> adir(R(md.anisotropyDirection[0]), R(md.anisotropyDirection[1]), R(0));
  if (md.texIds[13] >= 0.0f) {
    const V4<R> ad = gtv(s, md.texIds[13], rs);
    adir = V3<R>(ad.x * R(2) - R(1), ad.y * R(2) - R(1), R(0));
  } else if (md.texIds[12] >= 0.0f) {
    adir = V3<R>(anis.x * R(2) - R(1), anis.y * R(2) - R(1), R(0));
  }
  m.anisotropyRotation = oatan2(adir.y, adir.x);
  // rotation_to_tangent(angle, n, t), SH/utils.glsl:183-190
  R angle = m.anisotropyRotation + K<R>::PI;
  if (angle > R(0)) {
    V3<R> gt = c.t;
    V3<R> gb = cross(c.n, c.t) * c.tw;
    c.t = gt * ocos(angle) + gb * osin(angle);
  }
  c.anisotropyTangent = c.t;
  m.emissive = V3<R>(R(md.emissive[0]), R(md.emissive[1]), R(md.emissive[2])) * vpow(xyz4(gtv(s, md.texIds[3], rs)), E22);
  m.emissiveStrength = R(1);
  m.clearcoat = R(md.clearcoat) * gtv(s, md.texIds[7], rs).x;
  m.clearcoatRoughness = R(md.clearcoatRoughness) * gtv(s, md.texIds[8], rs).x;
  m.clearcoatNormalScale = R(md.clearcoatNormalScale);
  m.attenuationColor = V3<R>(R(md.attenuationColor[0]), R(md.attenuationColor[1]), R(md.attenuationColor[2]));
  m.attenuationDistance = R(md.attenuationDistance);
  m.thickness = c.thin_walled ? R(0) : R(1);
  m.multiscatterColor = V3<R>(R(0));
  m.scatterAnisotropy = R(0);
  m.iridescence = R(md.iridescence) * gtv(s, md.texIds[14], rs).x;
  m.iridescenceIor = R(md.iridescenceIor);
  m.iridescenceThickness = gmix(R(md.iridescenceThicknessMin), R(md.iridescenceThicknessMax), gtv(s, md.texIds[15], rs).y);
  m.dispersion = R(md.dispersion);
  m.normalScale = R(md.normalScale);
  return c;
}

// SH/pbr_material_adapter.glsl:1-40
template <class R> BState<R> make_state(const MClosure<R>& c, int profile) {
  BState<R> s;
  s.c = buildClosure(c.material, c.anisotropyTangent, profile);
  s.currentMediumIor = gmax(R(1), R(1));
  return s;
}

// SH/renderer.frag:72-91 + SH/bvh.glsl:91-134
template <class R> void fill_render_state(const Scene& s, const Ray<R>& r, const Hit<R>& h, RState<R>& rs) {
  rs.hitPos = r.org + r.dir * h.t;
  uint32_t tri = (uint32_t)h.tri;
  rs.wi = -r.dir;
  V3<R> p0 = vtx_attr<R>(s, tri, 0, 0), p1 = vtx_attr<R>(s, tri, 1, 0), p2 = vtx_attr<R>(s, tri, 2, 0);
  rs.ng = normalize(cross(p1 - p0, p2 - p0));  // compute_triangle_normal: cross(e1=p1-p0, e0=p2-p0)
  R u = h.u, v = h.v;
  R w0 = R(1) - u - v;
  V3<R> n0 = vtx_attr<R>(s, tri, 0, 1), n1 = vtx_attr<R>(s, tri, 1, 1), n2 = vtx_attr<R>(s, tri, 2, 1);
  rs.n = normalize(w0 * n0 + u * n1 + v * n2);
  R t0w, t1w, t2w;
  V3<R> t0 = vtx_attr<R>(s, tri, 0, 3, &t0w), t1 = vtx_attr<R>(s, tri, 1, 3, &t1w), t2 = vtx_attr<R>(s, tri, 2, 3, &t2w);
  R hand = (t0w == t1w && t0w == t2w) ? t0w : R(0);
  V3<R> tan = normalize(w0 * t0 + u * t1 + v * t2);
  if (length(tan) > R(0.99f) && std::fabs(hand) > R(0.99f)) { rs.tangent = tan; rs.tangent_w = -hand; }
  else { rs.tangent = onb_tangent(rs.n); rs.tangent_w = R(1); }
  // compute_interpolated_uv(tri, uv, 0), bvh.glsl:108-115 (attribute 2, .xy)
  {
    V3<R> a0 = vtx_attr<R>(s, tri, 0, 2), a1 = vtx_attr<R>(s, tri, 1, 2), a2 = vtx_attr<R>(s, tri, 2, 2);
    rs.uv0x = w0 * a0.x + u * a1.x + v * a2.x;
    rs.uv0y = w0 * a0.y + u * a1.y + v * a2.y;
  }
  R c0w, c1w, c2w;
  V3<R> c0 = vtx_attr<R>(s, tri, 0, 4, &c0w), c1 = vtx_attr<R>(s, tri, 1, 4, &c1w), c2 = vtx_attr<R>(s, tri, 2, 4, &c2w);
  V3<R> vc = w0 * c0 + u * c1 + v * c2;
  R vcw = w0 * c0w + u * c1w + v * c2w;
  R mi;
  vtx_attr<R>(s, tri, 0, 0, &mi);
  uint32_t matIdx = (uint32_t)mi;
  rs.closure = configure_material(s, matIdx, rs, vc, vcw);
}

// Per-sample path signature (tests only): FNV-1a over the branch decisions
// and hit triangle ids. Both engines mix the same events in the same order.
struct Sig {
  uint32_t h = 2166136261u;
  void mix(uint32_t v) { h = (h ^ v) * 16777619u; }
};

template <class R> struct PathOut { V3<R> L; R alpha; uint32_t sig; int primary_tri; };

// SH/integrator/misptdl.glsl:2-32 + SH/lighting.glsl:2-22
template <class R>
V3<R> eval_direct_light(const Scene& s, const Frame& f, const RState<R>& rs, float r0, float r1, Counters* cnt) {
  V3<R> L(R(0));
  V3<R> n = rs.closure.backside ? -rs.closure.n : rs.closure.n;
  if (f.use_ibl > 0.0f) {
    R ipdf;
    V3<R> dir = ibl_sample_direction(s, f, r0, r1, ipdf);
    R cosNL = sat(dot(dir, n));
    if (cosNL > K<R>::EPS_COS && ipdf > K<R>::EPS_PDF) {
      Ray<R> sr = make_ray(normalize(dir), rs.hitPos, K<R>::TFAR_MAX);
      Hit<R> sh;
      if (!intersect_nearest(s, sr, sh, cnt)) {
        BState<R> st = make_state(rs.closure, s.profile);
        V3<R> cn = rs.closure.n;
        V3<R> bsdf = evalMixture(st, cn, cn, rs.wi, dir) / gmax(std::fabs(dot(rs.closure.n, dir)), K<R>::EPS);
        R bpdf = pdfMixture(st, cn, cn, rs.wi, dir);
        L = ibl_eval(s, f, dir) * bsdf * cosNL / ipdf;
        if (bpdf > K<R>::EPS_PDF) L = L * (ipdf / (ipdf + bpdf));
      }
    }
  }
  if (f.light_pos[3] > 0.5f) {
    V3<R> lp(R(f.light_pos[0]), R(f.light_pos[1]), R(f.light_pos[2]));
    V3<R> le(R(f.light_emission[0]), R(f.light_emission[1]), R(f.light_emission[2]));
    V3<R> ld = lp - rs.hitPos;
    R dist2 = dot(ld, ld);
    ld = normalize(ld);
    V3<R> nn = rs.closure.backside ? -rs.closure.n : rs.closure.n;
    R cosNL = sat(dot(ld, nn));
    V3<R> d = lp - rs.hitPos;
    Ray<R> vr = make_ray(normalize(d), rs.hitPos, length(d));
    Hit<R> vh;
    bool visible = !intersect_nearest(s, vr, vh, cnt);
    V3<R> Lp(R(0));
    if (cosNL > K<R>::EPS_COS && visible) {
      BState<R> st = make_state(rs.closure, s.profile);
      V3<R> cn = rs.closure.n;
      V3<R> e = evalMixture(st, cn, cn, rs.wi, ld) / gmax(std::fabs(dot(rs.closure.n, ld)), K<R>::EPS);
      Lp = e * (le / dist2) * cosNL;
    }
    L = L + Lp;
  }
  return L;
}

// SH/renderer.frag:129-142
template <class R> Ray<R> camera_ray(const Frame& f, int W, int H, int px, int py, float r0, float r1) {
  float vux = (2.0f * (float(px) + 0.5f)) / float(W) - 1.0f;
  float vuy = (2.0f * (float(py) + 0.5f)) / float(H) - 1.0f;
  R offx = R(r0) * R(f.inv_res[0]), offy = R(r1) * R(f.inv_res[1]);
  R aspect = R(f.inv_res[1]) / R(f.inv_res[0]);
  R ux = (R(vux) * aspect + offx) * R(f.cam[3]);
  R uy = (R(vuy) * R(1) + offy) * R(f.cam[3]);
  V3<R> d = normalize(V3<R>(ux, uy, -R(f.focal_length)));
  const float* m = f.view;
  V3<R> wd(R(m[0]) * d.x + R(m[4]) * d.y + R(m[8]) * d.z,
           R(m[1]) * d.x + R(m[5]) * d.y + R(m[9]) * d.z,
           R(m[2]) * d.x + R(m[6]) * d.y + R(m[10]) * d.z);
  return make_ray(wd, V3<R>(R(f.cam[0]), R(f.cam[1]), R(f.cam[2])), K<R>::TFAR_MAX);
}

// SH/integrator/debug.glsl:1-83, the trace() of the debug program, which the
// renderer selects whenever debugMode is not 'None' (renderer.ts:1020) with
// u_debug_mode = the mode's index in debugModes (renderer.ts:181, 406; -1
// for a name not in the list): the camera ray's first hit shows one
// attribute of its render state; a miss shows the environment, or its
// pdf (x 10) or cdf table, when the background comes from the IBL, and
// vec4(0) otherwise.
template <class R> PathOut<R> debug_trace(const Scene& s, const Frame& f, const Ray<R>& ray, Counters* cnt, Sig& sig) {
  PathOut<R> out;
  const int mode = int(f.debug_mode);
  Hit<R> hit;
  out.primary_tri = -1;
  out.alpha = R(0);
  out.L = V3<R>(R(0));
  if (!intersect_nearest(s, ray, hit, cnt)) {
    sig.mix(0xFFFFFFFFu);
    if (f.bg_from_ibl > 0.0f) {
      R pdf;
      const V2<R> uv = dir_to_uv(rotate_dir_phi(ray.dir, R(f.ibl_rotation), false), pdf);
      const bool have = s.env && s.ew > 0 && s.eh > 0;
      if (mode == 10) {  // vec4(texture(pdf).xyz, 1.0) * 10.0
        const R p = have ? table_bilinear(s.pdf, s.ew, s.eh, uv) : R(0);
        out.L = V3<R>(p * R(10), R(0) * R(10), R(0) * R(10));
        out.alpha = R(1) * R(10);
      } else if (mode == 11) {
        out.L = V3<R>(have ? table_bilinear(s.cdf, s.ew, s.eh, uv) : R(0), R(0), R(0));
        out.alpha = R(1);
      } else {
        out.L = have ? env_bilinear(s, uv) : V3<R>(R(0));
        out.alpha = R(1);
      }
    }
    out.sig = sig.h;
    return out;
  }
  out.primary_tri = hit.tri;
  sig.mix((uint32_t)hit.tri);
  RState<R> rs;
  fill_render_state(s, ray, hit, rs);
  const MClosure<R>& c = rs.closure;
  const PbrMat<R>& m = c.material;
  V3<R> con(R(0));
  if (mode == 1) con = m.baseColor;                       // D_ALBEDO
  if (mode == 2) con = V3<R>(m.metallic);                 // D_METAL
  if (mode == 3) con = V3<R>(m.roughness);                // D_ROUGHNESS
  if (mode == 4) con = c.n;                               // D_NORMAL
  if (mode == 5) con = c.t;                               // D_TANGENT (after the anisotropy rotation)
  if (mode == 6) con = cross(c.n, c.t) * c.tw;            // D_BITANGENT: calculateBasis(c.n, c.t).b
  if (mode == 7) con = V3<R>(m.transmission);             // D_TRANSPARENCY
  if (mode == 14) con = m.diffuseTransmission * m.diffuseTransmissionColor;  // D_TRANSLUCENCY
  if (mode == 8) con = V3<R>(rs.uv0x, rs.uv0y, R(0));     // D_UV0
  if (mode == 9) con = V3<R>(m.clearcoat);                // D_CLEARCOAT
  if (mode == 12) con = V3<R>(m.specular);                // D_SPECULAR
  if (mode == 13) con = m.specularColor;                  // D_SPECULAR_TINT
  out.L = con;
  out.alpha = R(1);
  out.sig = sig.h;
  return out;
}

// SH/renderer.frag:146-157 + SH/integrator/misptdl.glsl:34-93
template <class R>
PathOut<R> trace_path(const Scene& s, const Frame& f, int W, int H, int px, int py, Counters* cnt) {
  PathOut<R> out;
  Sig sig;
  Rng rng = rng_init(int(f.frame_count) * int(f.max_bounces), px, py);
  float r0 = rng_float(rng), r1 = rng_float(rng);
  Ray<R> ray = camera_ray<R>(f, W, H, px, py, r0, r1);
  if (f.debug_mode != 0.0f) return debug_trace(s, f, ray, cnt, sig);
  Hit<R> hit;
  V3<R> pw(R(1));
  V3<R> L(R(0));
  R last_pdf = R(0);
  out.primary_tri = -1;
  if (!intersect_nearest(s, ray, hit, cnt)) {
    sig.mix(0xFFFFFFFFu);
    out.sig = sig.h;
    if (f.bg_from_ibl > 0.0f) { out.L = ibl_eval(s, f, ray.dir) * pw; out.alpha = R(1); }
    else {
      out.L = V3<R>(opow(R(f.bg[0]), R(2.2f)), opow(R(f.bg[1]), R(2.2f)), opow(R(f.bg[2]), R(2.2f)));
      out.alpha = R(f.bg[3]);
    }
    return out;
  }
  out.primary_tri = hit.tri;
  sig.mix((uint32_t)hit.tri);
  RState<R> rs;
  fill_render_state(s, ray, hit, rs);
  int bounce = 0;
  const int maxB = int(f.max_bounces);
  while (bounce < maxB) {
    // RR, SH/renderer.frag:119-123 (draw only when bounce > 2)
    pw = pw * (R(1) / (R(1) - R(RR_TERMINATION_PROB)));
    if (bounce > RR_START_DEPTH && rng_float(rng) <= RR_TERMINATION_PROB) { sig.mix(1u); break; }
    if (rs.closure.backside && !rs.closure.thin_walled) {
      V3<R> sigma = -vlog(rs.closure.material.attenuationColor) / rs.closure.material.attenuationDistance;
      pw = pw * vexp(-sigma * hit.t);
    }
    L = L + rs.closure.material.emissive * rs.closure.material.emissiveStrength * pw;
    // sample_bsdf_bounce, SH/renderer.frag:94-116
    V3<R> bw(R(1));
    R pdf = R(1);
    if (R(rng_float(rng)) > rs.closure.cutout_opacity) {
      rs.wo = -rs.wi;
      sig.mix(2u);
    } else {
      float lx = rng_float(rng), ly = rng_float(rng), bz = rng_float(rng);
      float comp = rng_float(rng);
      BState<R> st = make_state(rs.closure, s.profile);
      V3<R> cn = rs.closure.n;
      BSample<R> bs = sampleMixture(st, rs.closure.ng, cn, cn, cn, rs.wi, R(comp), R(lx), R(ly), R(bz));
      pdf = bs.pdf;
      bw = bs.bsdfOverPdf;
      rs.wo = bs.dir;
      sig.mix(3u + (uint32_t)(int)bs.specular);
      if (pdf < K<R>::EPS_PDF) { sig.mix(7u); out.L = L; out.alpha = R(1); out.sig = sig.h; return out; }
    }
    last_pdf = pdf;
    rs.hitPos = rs.hitPos + fix_normal(rs.ng, rs.wo) * R(f.ray_eps);
    {
      float n0 = rng_float(rng), n1 = rng_float(rng);
      L = L + eval_direct_light(s, f, rs, n0, n1, cnt) * pw;
    }
    pw = pw * bw;
    ray = make_ray(rs.wo, rs.hitPos, K<R>::TFAR_MAX);
    if (intersect_nearest(s, ray, hit, cnt)) {
      sig.mix((uint32_t)hit.tri);
      fill_render_state(s, ray, hit, rs);
      bounce++;
    } else {
      sig.mix(0xFFFFFFFFu);
      R ipdf = ibl_pdf(s, f, ray.dir);
      R mis = last_pdf / (last_pdf + ipdf);
      L = L + ibl_eval(s, f, ray.dir) * pw * mis * R(f.use_ibl);
      bounce = 1337;
    }
  }
  if (f.clamp_threshold > 0.0f) L = vclamp(L, V3<R>(R(0)), V3<R>(R(f.clamp_threshold)));
  out.L = L; out.alpha = R(1); out.sig = sig.h;
  return out;
}

}  // namespace orc
