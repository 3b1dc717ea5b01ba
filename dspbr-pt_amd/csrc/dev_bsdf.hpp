// Device implementation of the Enterprise PBR BSDF mixture used by the
// reference (slang-pbr 0.1.0 generated GLSL, packages/lib/shader/generated/
// slang_materials/webgl-full/material_kernel.glsl, "KF"; the webgl-lean
// profile is the same mixture with iridescence and dispersion forced to 0,
// KL:395,409,412).
//
// The reference adapter always enters with currentMediumIor = 1
// (SH/pbr_material_adapter.glsl:18), so every dielectric interface is
// air -> material: iorI = 1, iorO = ior, insideMedium = false. This file is
// specialised for that transport state; the closure is built once per hit
// and shared by the BSDF sample and the NEE eval/pdf.
#pragma once
#include "dev_math.hpp"

namespace bsdf {

constexpr float E5 = 0.00000999999974738f;   // 1e-5 as written by the generator
constexpr float E4 = 0.00009999999747379f;
constexpr float PIf = 3.14159274101257324f;
constexpr float INV_PI = 0.31830987334251404f;
constexpr float TWO_PIf = 6.28318548202514648f;
constexpr float E6 = 9.99999997475242708e-07f;
constexpr float E8 = 9.99999993922529029e-09f;
constexpr float E12 = 9.999999960041972e-13f;
constexpr float AVG_W = 0.0476190485060215f;     // 1/21, Schlick hemispherical average weight
constexpr float IOR_ONE = 1.00010001659393311f;

// glTF-side inputs after texture evaluation (PbrGltfMaterial facade).
struct GltfInputs {
  f3 baseColor;
  float metallic, roughness, anisotropy, transmission, ior;
  f3 specularColor;
  float specular;
  f3 sheenColor;
  float sheenRoughness, clearcoat, clearcoatRoughness, diffuseTransmission;
  f3 diffuseTransmissionColor;
  float iridescence, iridescenceIor, iridescenceThickness, dispersion;
  bool thin;
};

// SurfaceClosure fields the mixture reads (KF:330-422).
struct Closure {
  f3 diffRefl, albedo, dielF0, tanA, sheenColor, throughput, diffTrans;
  float diffuseWeight, sheenWeight, aniso, dielF90, dielectricWeight, specularWeight, clearcoatWeight;
  float rough, sheenRough, transparency, clearcoat, ccRough, metallic, specular, ior;
  float irid, iridIor, iridThick, dispersion, transmissionWeight, diffTransWeight;
  bool thin;
};

DI Closure build_closure(const GltfInputs& m, f3 anisoTangent, bool full_profile) {
  const f3 Z = sp3(0.0f), O = sp3(1.0f);
  Closure c;
  f3 albedo = vmin(vmax(m.baseColor, Z), O);
  float metallic = gclamp(m.metallic, 0.0f, 1.0f);
  float transparency = gclamp(m.transmission, 0.0f, 1.0f);
  float ior = gmax(m.ior, 1.0f);
  float specular = gclamp(m.specular, 0.0f, 1.0f);
  float translucency = gclamp(m.diffuseTransmission, 0.0f, 1.0f);
  float iridIor = gclamp(m.iridescenceIor, 1.0f, 2.5f);
  float irid = full_profile ? gclamp(m.iridescence, 0.0f, 1.0f) : 0.0f;
  float disp = full_profile ? gmax(m.dispersion, 0.0f) : 0.0f;
  f3 sheen = vclamp01(m.sheenColor);
  float ccRoughIn = gclamp(m.clearcoatRoughness, 0.0f, 1.0f);

  float dw = 1.0f - metallic;
  float opaqueFrac = 1.0f - transparency;
  float opaqueDiel = dw * opaqueFrac;
  float clearDiel = dw * transparency;
  c.diffRefl = albedo * opaqueDiel * (1.0f - translucency);
  c.diffTrans = opaqueDiel * vclamp01(m.diffuseTransmissionColor) * translucency;
  float r = (1.0f - ior) / (1.0f + ior);
  c.dielF0 = vclamp01(sp3(r * r * specular) * vclamp01(m.specularColor));
  f3 metalE = albedo + (O - albedo) * AVG_W;
  c.albedo = albedo;
  c.diffuseWeight = max3c(c.diffRefl);
  c.sheenWeight = max3c(sheen) * opaqueFrac;
  c.aniso = gclamp(m.anisotropy, 0.0f, 1.0f);
  c.tanA = anisoTangent;
  c.dielF90 = specular;
  c.dielectricWeight = dw;
  float s = gmax(iridIor, 1.0f);
  float film = gclamp(irid, 0.0f, 1.0f) * pm::ppow((s - 1.0f) / gmax(s + 1.0f, E5), 2.0f);
  c.specularWeight = gmax(max3c(metalE), film) * metallic;
  c.ccRough = gclamp(gmax(ccRoughIn, 0.0020000000949949f), 0.0f, 1.0f);
  c.clearcoat = gclamp(m.clearcoat, 0.0f, 1.0f);
  c.clearcoatWeight = c.clearcoat * gmix(0.20000000298023224f, 1.0f, 1.0f - c.ccRough);
  c.rough = gclamp(gmax(gclamp(m.roughness, 0.0f, 1.0f), 0.0020000000949949f), 0.0f, 1.0f);
  c.sheenColor = sheen;
  c.sheenRough = gclamp(m.sheenRoughness, 0.0f, 1.0f);
  c.transparency = transparency;
  c.throughput = vclamp01(c.diffRefl + c.diffTrans + albedo * clearDiel + vmax(metalE, sp3(film)) * metallic +
                          c.dielF0 * dw + sheen);
  c.metallic = metallic;
  c.specular = specular;
  c.ior = ior;
  c.irid = irid;
  c.iridIor = iridIor;
  c.iridThick = gmax(m.iridescenceThickness, 0.0f);
  c.dispersion = disp;
  c.thin = m.thin;
  c.transmissionWeight = max3c(albedo) * clearDiel;
  c.diffTransWeight = max3c(c.diffTrans);
  return c;
}

// ---------------------------------------------------------------- energy
DI float avg_albedo(float a) {
  float r = sqrtf(gmax(a, 0.0f));
  return gclamp(1.0f - 0.57999998331069946f * r + 0.07999999821186066f * r * r, 0.0f, 1.0f);
}
DI float avg_albedo_ms(float a, float e0) {
  return gclamp(gclamp(e0 + (1.0f - e0) * AVG_W, 0.0f, 1.0f) * avg_albedo(a), 0.0f, 1.0f);
}
DI float dir_albedo(float a, float ct) {
  return gclamp(1.0f - sqrtf(gmax(a, 0.0f)) * (0.5f + 0.5f * pm::ppow(1.0f - gclamp(ct, 0.0f, 1.0f), 5.0f)), 0.0f, 1.0f);
}
DI float dir_albedo_ms(float ct, float a, float e0) {
  float nv = gclamp(ct, 0.0f, 1.0f);
  return gclamp(gclamp(e0 + (1.0f - e0) * pm::ppow(1.0f - nv, 5.0f), 0.0f, 1.0f) * gmix(1.0f, dir_albedo(a, nv), 0.64999997615814209f),
                0.0f, 1.0f);
}

// ---------------------------------------------------------------- GGX
struct Frame3 { f3 t, b, n; };
DI f3 to_local(const Frame3& F, f3 d) { return f3{dot3(d, F.t), dot3(d, F.b), dot3(d, F.n)}; }
// projectToAnisotropicFrame basis (KF:483-490): tangent orthogonalised to n
DI Frame3 aniso_frame(f3 n, f3 t) {
  f3 st = nrm3(t - n * dot3(t, n));
  return Frame3{st, nrm3(cross3(n, st)), n};
}
DI float2 alpha_xy(float rough, float aniso) {
  float a = gclamp(aniso, 0.0f, 1.0f);
  float s = gmax(rough * rough, 0.0f);
  return make_float2(gmix(s, 1.0f, a * a), s);
}
DI float D_aniso(float2 a, f3 h) {
  float hx = h.x / a.x, hy = h.y / a.y, hz = h.z;
  float d = hx * hx + hy * hy + hz * hz;
  return 1.0f / gmax(PIf * a.x * a.y * d * d, E12);
}
DI float lambda_aniso(float2 a, f3 d) {
  float c2 = d.z * d.z;
  if (c2 <= E8) return 0.0f;
  if (gmax(1.0f - c2, 0.0f) <= E8) return 0.0f;
  return 0.5f * (-1.0f + sqrtf(1.0f + (d.x * d.x * a.x * a.x + d.y * d.y * a.y * a.y) / gmax(c2, E8)));
}
DI float G2_aniso(float2 a, f3 v, f3 l) { return 1.0f / (1.0f + lambda_aniso(a, v) + lambda_aniso(a, l)); }
DI float G1_aniso(float2 a, f3 d) { return 1.0f / (1.0f + lambda_aniso(a, d)); }
// evaluateMicrofacetResponseAnisotropic (KF:553-615); scalar D*G2/(4 nv nl)
DI float mf_aniso(const Closure& c, const Frame3& F, f3 v, f3 l) {
  f3 h = nrm3(v + l);
  f3 lv = to_local(F, v), ll = to_local(F, l), lh = to_local(F, h);
  float nv = fabsf(lv.z), nl = fabsf(ll.z), nh = fabsf(lh.z);
  float vh = gmax(dot3(v, h), 0.0f);
  if (nv <= 0.0f || nl <= 0.0f || nh <= 0.0f || vh <= 0.0f) return 0.0f;
  float2 a = alpha_xy(c.rough, c.aniso);
  return D_aniso(a, lh) * G2_aniso(a, lv, ll) / gmax(4.0f * nv * nl, E5);
}
DI float D_iso(float a, float nh) {
  float a2 = a * a;
  float d = nh * nh * (a2 - 1.0f) + 1.0f;
  return a2 / gmax(PIf * d * d, E12);
}
DI float lambda_iso(float a, float nx) {
  float ct = fabsf(nx);
  float c2 = ct * ct;
  float s2 = gmax(1.0f - c2, 0.0f);
  if (s2 <= E8 || ct <= E8) return 0.0f;
  return 0.5f * (-1.0f + sqrtf(1.0f + a * a * (s2 / gmax(c2, E8))));
}

// ---------------------------------------------------------------- Fresnel
DI f3 schlick(f3 f0, f3 f90, float ct) {
  float m = gclamp(1.0f - ct, 0.0f, 1.0f);
  float m2 = m * m;
  return f0 + (f90 - f0) * (m2 * m2 * m);
}
DI f3 schlick_avg(f3 f0, f3 f90) { return f0 + (f90 - f0) * AVG_W; }
// fresnelSchlickDielectric with ni = 1 (KF:712-739)
DI f3 schlick_dielectric(float ct, f3 f0, f3 f90, float nt, bool thin) {
  const float ni = 1.0f;
  if (fabsf(ni - nt) <= E4) return sp3(0.0f);
  float ci = gclamp(ct, 0.0f, 1.0f);
  float eta = ni / gmax(nt, E5);
  float s2 = eta * eta * gmax(1.0f - ci * ci, 0.0f);
  if (s2 >= 1.0f && !thin) return sp3(1.0f);
  float sc = (nt >= ni || thin) ? ci : sqrtf(gmax(1.0f - s2, 0.0f));
  return schlick(f0, f90, sc);
}

// ---------------------------------------------------------------- thin film (KF:741-936)
DI float film_f0(float t, float i) {
  float r = (t - i) / gmax(t + i, E5);
  return r * r;
}
DI f3 film_sensitivity(float opd, f3 shift) {
  float ph = 6.28318548202514648f * opd * 9.99999971718068537e-10f;
  float ph2 = ph * ph;
  const f3 var = f3{4.327799808e+09f, 9.304599552e+09f, 6.612100096e+09f};
  f3 xyzv = f3{5.48560003844900113e-13f, 4.42010008238832852e-13f, 5.24810013155518895e-13f} * vsqrt(6.28318548202514648f * var) *
            vcos(f3{1.681e+06f, 1.7953e+06f, 2.2084e+06f} * ph + shift) * vexp(-ph2 * var);
  xyzv.x = xyzv.x + 9.74699989231449238e-14f * sqrtf(2.845152256e+10f) * pm::pcos(2.2399e+06f * ph + shift.x) * pm::pexp(-4.528200192e+09f * ph2);
  f3 q = xyzv / 1.06850002623559703e-07f;
  return f3{3.24045419692993164f * q.x - 1.53713846206665039f * q.y - 0.49853140115737915f * q.z,
            -0.96926599740982056f * q.x + 1.87601077556610107f * q.y + 0.04155600070953369f * q.z,
            0.05564339831471443f * q.x - 0.20402589440345764f * q.y + 1.05722522735595703f * q.z};
}
#ifndef DSPT_FILM_INLINE
#define DSPT_FILM_INLINE 1
#endif
#if DSPT_FILM_INLINE
#define FILM_FN DI
#else
#define FILM_FN DN
#endif
FILM_FN f3 film_reflectance(float outside, float filmIor, f3 baseF0, float thick, float cos1) {
  float fi = gclamp(filmIor, 1.0f, 2.5f);
  float c1 = gclamp(cos1, 0.0f, 1.0f);
  float c2sq = 1.0f - outside / fi * (outside / fi) * (1.0f - c1 * c1);
  if (c2sq < 0.0f) return sp3(1.0f);
  float c2 = sqrtf(c2sq);
  const f3 O = sp3(1.0f);
  float r12 = schlick(sp3(film_f0(fi, outside)), O, c1).x;
  float t121 = 1.0f - r12;
  f3 sf0 = vsqrt(vmin(vmax(baseF0, sp3(0.0f)), sp3(0.99989998340606689f)));
  f3 baseIor = (O + sf0) / vmax(O - sf0, sp3(E5));
  f3 fiv = sp3(fi);
  f3 rr = (baseIor - fiv) / vmax(baseIor + fiv, sp3(E5));
  f3 r23 = schlick(rr * rr, O, c2);
  float opd = 2.0f * fi * gmax(thick, 0.0f) * c2;
  float phi12 = fi < outside ? PIf : 0.0f;
  float phi21 = PIf - phi12;
  f3 phi = sp3(phi21) + f3{baseIor.x < fi ? PIf : 0.0f, baseIor.y < fi ? PIf : 0.0f, baseIor.z < fi ? PIf : 0.0f};
  f3 r12v = sp3(r12);
  f3 r123 = vmin(vmax(r12v * r23, sp3(E5)), sp3(0.99989998340606689f));
  f3 sr = vsqrt(r123);
  f3 rs = t121 * t121 * r23 / vmax(O - r123, sp3(E5));
  f3 refl = r12v + rs;
  f3 cm = rs - sp3(t121);
  for (int m = 1; m <= 2; m++) {
    cm = cm * sr;
    float fm = float(m);
    refl = refl + cm * 2.0f * film_sensitivity(fm * opd, fm * phi);
  }
  return vmaxs(refl, 0.0f);
}
DI f3 film_mix(f3 regular, f3 film, float irid) { return vmix(regular, film, gclamp(irid, 0.0f, 1.0f)); }
// thin-film dielectric fresnel on the air side (incident ior 1)
DI f3 film_dielectric(const Closure& c, float ct, f3 regular) {
  if (c.irid <= E5) return regular;
  return film_mix(regular, film_reflectance(1.0f, c.iridIor, sp3(film_f0(gmax(c.ior, 1.0f), 1.0f)), c.iridThick, ct), c.irid);
}
DI f3 film_dielectric_if(const Closure& c, float ct, f3 regular) {
  return (c.irid > E5) ? film_dielectric(c, ct, regular) : regular;
}

// ---------------------------------------------------------------- shared terms
DI float clear_dielectric(const Closure& c) { return c.dielectricWeight * c.transparency; }
DI f3 trans_color(const Closure& c) { return c.albedo * clear_dielectric(c); }
DI f3 flip_across(f3 v, f3 n) { return nrm3(v - 2.0f * n * dot3(v, n)); }

struct TransCoeffs { f3 rc, tc; float rp, tp; };
// computeDielectricTransmissionCoefficients (KF:1074-1130), currentMediumIor = 1
DI TransCoeffs trans_coeffs(const Closure& c, f3 n, f3 v) {
  float cv = gmax(dot3(n, v), E5);
  f3 fr = schlick_dielectric(cv, c.dielF0, sp3(c.dielF90), c.ior, false);
  f3 film = film_dielectric_if(c, cv, fr);
  TransCoeffs t;
  t.rc = film * clear_dielectric(c);
  t.tc = (sp3(1.0f) - film) * trans_color(c);
  f3 tp = vmaxs(c.throughput, 1.0f);  // max(throughput, vec3(max(currentMediumIor, 1e-5)))
  float rw = max3c(vmaxs(tp, 0.0f) * vmaxs(t.rc, 0.0f));
  float tw = max3c(vmaxs(tp, 0.0f) * vmaxs(t.tc, 0.0f));
  float tot = rw + tw;
  if (tot <= E5) { t.rp = 1.0f; t.tp = 0.0f; }
  else { t.rp = rw / tot; t.tp = tw / tot; }
  return t;
}

// evaluateDielectricSpecularGGX with iorI = 1, iorO = ior (KF:948-1032)
DI f3 dielectric_reflection(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  float nl = gmax(dot3(n, l), 0.0f);
  float tw = clear_dielectric(c);
  if (nl <= 0.0f || tw <= 0.0f) return sp3(0.0f);
  float mr = mf_aniso(c, F, v, l);
  float vh = gmax(dot3(v, nrm3(v + l)), 0.0f);
  float a = gmax(c.rough * c.rough, 0.0f);
  float eAvg = avg_albedo(a);
  float ome = 1.0f - eAvg;
  float ms = (1.0f - dir_albedo(a, nl)) * (1.0f - dir_albedo(a, fabsf(dot3(n, v)))) / gmax(PIf * ome, E6);
  f3 f = film_dielectric_if(c, vh, schlick_dielectric(vh, c.dielF0, sp3(c.dielF90), c.ior, false));
  f3 fAvg = (fabsf(1.0f - c.ior) <= E4) ? sp3(0.0f) : schlick_avg(c.dielF0, sp3(c.dielF90));
  return tw * (sp3(mr) * f + ms * (fAvg * fAvg * eAvg / vmax(sp3(1.0f) - fAvg * ome, sp3(E5)))) * nl;
}
DI f3 thin_transmission(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  float nl = dot3(n, l);
  if (dot3(n, v) * nl >= 0.0f) return sp3(0.0f);
  f3 ml = flip_across(l, n);
  float ch = gmax(dot3(v, nrm3(v + ml)), E5);
  f3 fr = film_dielectric_if(c, ch, schlick_dielectric(ch, c.dielF0, sp3(c.dielF90), c.ior, true));
  return trans_color(c) * (sp3(1.0f) - fr) * sp3(mf_aniso(c, F, v, ml)) * fabsf(nl);
}
DI f3 thick_transmission(const Closure& c, const Frame3& F, f3 tc, f3 n, f3 v, f3 l, float iorO) {
  f3 lv = to_local(F, v), ll = to_local(F, l);
  float nv = fabsf(lv.z), nl = fabsf(ll.z);
  if (nv <= E5 || nl <= E5 || dot3(n, v) * dot3(n, l) >= 0.0f) return sp3(0.0f);
  float eta = iorO / gmax(1.0f, E5);
  f3 h = -nrm3(v + l * eta);
  f3 lh = to_local(F, h);
  float vh = dot3(v, h), lhd = dot3(l, h);
  if (fabsf(lh.z) <= E5 || vh * lhd >= 0.0f) return sp3(0.0f);
  float2 a = alpha_xy(c.rough, c.aniso);
  float d = D_aniso(a, lh);
  float g = G2_aniso(a, lv, ll);
  float avh = fabsf(vh);
  float factor = eta * eta * avh * fabsf(lhd) / gmax(pm::ppow(vh + eta * lhd, 2.0f) * nv, E5);
  f3 etc = tc;
  if (c.irid > E5) {
    f3 rr = schlick_dielectric(avh, c.dielF0, sp3(c.dielF90), iorO, false);
    etc = (sp3(1.0f) - film_dielectric(c, avh, rr)) * trans_color(c);
  }
  return etc * d * g * factor;
}
DI f3 disp_iors(const Closure& c) {
  float s = gmax(c.ior, 1.0f);
  float sc = (s - 1.0f) * gmax(c.dispersion, 0.0f) * 0.05000000074505806f;
  return vmaxs(f3{s + sc * -0.30094969272613525f, s, s + sc * 0.69905030727386475f}, 1.0f);
}
// evalDielectricTransmission (KF:1425-1497)
DI f3 dielectric_transmission(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  if (dot3(n, v) * dot3(n, l) > 0.0f) return dielectric_reflection(c, F, n, v, l);
  TransCoeffs co = trans_coeffs(c, n, v);
  if (c.thin) return thin_transmission(c, F, n, v, l);
  if (c.dispersion <= E5) return thick_transmission(c, F, co.tc, n, v, l, c.ior);
  f3 io = disp_iors(c);
  float r = thick_transmission(c, F, f3{co.tc.x, 0.0f, 0.0f}, n, v, l, io.x).x;
  float g = thick_transmission(c, F, f3{0.0f, co.tc.y, 0.0f}, n, v, l, io.y).y;
  float b = thick_transmission(c, F, f3{0.0f, 0.0f, co.tc.z}, n, v, l, io.z).z;
  return f3{r, g, b};
}

DI float coupled_diffuse(const Closure& c, float nv, float nl) {
  float a = gmax(c.rough * c.rough, 0.0f);
  float e0 = max3c(c.dielF0 * c.dielectricWeight);
  return gmix(0.31830987334251404f,
              (1.0f - dir_albedo_ms(nl, a, e0)) * (1.0f - dir_albedo_ms(nv, a, e0)) / gmax(PIf * (1.0f - avg_albedo_ms(a, e0)), E6),
              c.specular);
}
DI f3 thin_spec_f0(const Closure& c) { return c.dielF0 * c.dielectricWeight + c.albedo * c.metallic; }
DI f3 thin_spec_f90(const Closure& c) { return sp3(c.dielF90) * c.dielectricWeight + sp3(c.metallic); }
DI float transmissive_diffuse_comp(const Closure& c, float nv, float nl) {
  float coupled = coupled_diffuse(c, nv, nl);
  f3 sc = c.thin ? thin_spec_f0(c) : c.dielF0;
  f3 sf = c.thin ? thin_spec_f90(c) : sp3(c.dielF90);
  float f0 = max3c(sc);
  float d = max3c(sf) - f0;
  const f3 O = sp3(1.0f);
  f3 energy = vmin(vmax(O - (schlick_avg(c.albedo, O) * c.metallic + schlick_avg(c.dielF0, sp3(c.dielF90)) * c.dielectricWeight), sp3(0.0f)), O);
  return coupled * gmix(1.0f, 1.0f / gmax(max3c(energy), E4),
                        sqrtf(0.5f * (gclamp(f0 + d * pm::ppow(1.0f - gclamp(nv, 0.0f, 1.0f), 5.0f), 0.0f, 1.0f) +
                                      gclamp(f0 + d * pm::ppow(1.0f - gclamp(nl, 0.0f, 1.0f), 5.0f), 0.0f, 1.0f))));
}
DI f3 diffuse_reflection(const Closure& c, f3 n, f3 v, f3 l) {
  float nl = gmax(dot3(n, l), 0.0f);
  if (c.specular <= E5 && c.transparency <= 0.0f) return c.diffRefl * INV_PI * nl;
  float nv = gmax(fabsf(dot3(v, n)), 0.0f);
  float f = c.transparency > 0.0f ? transmissive_diffuse_comp(c, nv, nl) : coupled_diffuse(c, nv, nl);
  return c.diffRefl * f * nl;
}
DI bool thin_transmissive(const Closure& c) {
  return c.thin && (c.transparency > 0.0f || c.diffTransWeight > 0.0f) && c.metallic < 1.0f;
}
DI f3 base_specular(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  float nl = gmax(dot3(n, l), 0.0f);
  const f3 O = sp3(1.0f), E = sp3(E5);
  if (thin_transmissive(c)) {
    if (nl <= 0.0f) return sp3(0.0f);
    float mr = mf_aniso(c, F, v, l);
    float vh = gmax(dot3(v, nrm3(v + l)), 0.0f);
    float a = gmax(c.rough * c.rough, 0.0f);
    float eAvg = avg_albedo(a);
    float ome = 1.0f - eAvg;
    float ms = (1.0f - dir_albedo(a, nl)) * (1.0f - dir_albedo(a, fabsf(dot3(n, v)))) / gmax(PIf * ome, E6);
    f3 f0 = thin_spec_f0(c), f90 = thin_spec_f90(c);
    f3 f = schlick(f0, f90, vh);
    if (c.irid > E5)
      f = film_mix(f, film_reflectance(1.0f, c.iridIor, sp3(film_f0(gmax(c.ior, 1.0f), 1.0f)), c.iridThick, vh), c.irid);
    f3 fa = schlick_avg(f0, f90);
    return (sp3(mr) * f + ms * (fa * fa * eAvg / vmax(O - fa * ome, E))) * nl;
  }
  float mr = mf_aniso(c, F, v, l);
  float vh = gmax(dot3(v, nrm3(v + l)), 0.0f);
  float a = gmax(c.rough * c.rough, 0.0f);
  float eAvg = avg_albedo(a);
  float ome = 1.0f - eAvg;
  float ms = (1.0f - dir_albedo(a, fabsf(dot3(n, l)))) * (1.0f - dir_albedo(a, fabsf(dot3(n, v)))) / gmax(PIf * ome, E6);
  f3 metalF = schlick(c.albedo, O, vh);
  f3 dielF = schlick(c.dielF0, sp3(c.dielF90), vh);
  if (c.irid > E5) {
    metalF = film_mix(metalF, film_reflectance(1.0f, c.iridIor, c.albedo, c.iridThick, vh), c.irid);
    dielF = film_mix(dielF, film_reflectance(1.0f, c.iridIor, sp3(film_f0(gmax(c.ior, 1.0f), 1.0f)), c.iridThick, vh), c.irid);
  }
  f3 mfa = schlick_avg(c.albedo, O);
  f3 dfa = schlick_avg(c.dielF0, sp3(c.dielF90));
  f3 spec = c.metallic * (sp3(mr) * metalF + ms * (mfa * mfa * eAvg / vmax(O - mfa * ome, E))) +
            c.dielectricWeight * (1.0f - c.transparency) * (sp3(mr) * dielF + ms * (dfa * dfa * eAvg / vmax(O - dfa * ome, E)));
  return spec * nl;
}
DI float sheen_log(float x, float alpha) {
  float s = 1.0f - alpha;
  float o = s * s;
  return gmix(21.54730033874511719f, 25.32449913024902344f, o) /
             (1.0f + gmix(3.82986998558044434f, 3.32435011863708496f, o) * pm::ppow(fabsf(x), gmix(0.19822999835014343f, 0.16800999641418457f, o))) +
         gmix(-1.97759997844696045f, -1.27392995357513428f, o) * x + gmix(-4.32053995132446289f, -4.85967016220092773f, o);
}
DI float sheen_lambda(float ct, float alpha) {
  if (fabsf(ct) >= 0.5f) return pm::pexp(2.0f * sheen_log(0.5f, alpha) - sheen_log(1.0f - ct, alpha));
  return pm::pexp(sheen_log(ct, alpha));
}
DI f3 sheen(const Closure& c, f3 n, f3 v, f3 l) {
  if (max3c(c.sheenColor) <= 0.0f) return sp3(0.0f);
  float nv = gmax(dot3(n, v), 0.0f), nl = gmax(dot3(n, l), 0.0f), nh = gmax(dot3(n, nrm3(v + l)), 0.0f);
  if (nv <= 0.0f || nl <= 0.0f || nh <= 0.0f) return sp3(0.0f);
  float alpha = gmax(c.sheenRough * c.sheenRough, 0.0f);
  float inv = 1.0f / gmax(alpha, E5);
  return c.sheenColor *
         (1.0f / (1.0f + sheen_lambda(nv, alpha) + sheen_lambda(nl, alpha)) *
          ((2.0f + inv) * pm::ppow(fabsf(gmax(1.0f - nh * nh, 0.00100000004749745f)), 0.5f * inv) / 6.28318548202514648f) / gmax(4.0f * nv * nl, E5)) *
         nl;
}
DI f3 reflected_base(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  return diffuse_reflection(c, n, v, l) + base_specular(c, F, n, v, l) + sheen(c, n, v, l);
}
DI f3 clearcoat(const Closure& c, f3 n, f3 v, f3 l) {
  if (c.clearcoat <= 0.0f) return sp3(0.0f) * gmax(dot3(n, l), 0.0f);
  f3 h = nrm3(v + l);
  float nv = gmax(dot3(n, v), 0.0f), nl = gmax(dot3(n, l), 0.0f), nh = gmax(dot3(n, h), 0.0f), vh = gmax(dot3(v, h), 0.0f);
  f3 r = sp3(0.0f);
  if (!(nv <= 0.0f || nl <= 0.0f || nh <= 0.0f || vh <= 0.0f)) {
    float a = gmax(c.ccRough * c.ccRough, 0.0f);
    r = schlick(sp3(0.03999999910593033f), sp3(1.0f), vh) * D_iso(a, nh) * (1.0f / (1.0f + lambda_iso(a, nv) + lambda_iso(a, nl))) /
        gmax(4.0f * nv * nl, E5);
  }
  return c.clearcoat * r * gmax(dot3(n, l), 0.0f);
}
DI float clearcoat_base_weight(const Closure& c, f3 n, f3 v, f3 l) {
  if (c.clearcoat <= 0.0f) return 1.0f;
  const f3 f0 = sp3(0.03999999910593033f), O = sp3(1.0f);
  return gclamp(1.0f - c.clearcoat * gmax(max3c(schlick(f0, O, fabsf(dot3(n, v)))), max3c(schlick(f0, O, fabsf(dot3(n, l))))), 0.0f, 1.0f);
}
DI f3 diffuse_transmission(const Closure& c, f3 n, f3 l) { return c.diffTrans * INV_PI * gmax(fabsf(dot3(n, l)), 0.0f); }

// evalMaterialBsdfMixture (KF:2045-2089): returns bsdf * |cos|
DI f3 eval(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  if (!(dot3(n, v) * dot3(n, l) > 0.0f)) return diffuse_transmission(c, n, l) + dielectric_transmission(c, F, n, v, l);
  f3 base = reflected_base(c, F, n, v, l);
  if (!c.thin && c.transparency > 0.0f) base = base + dielectric_transmission(c, F, n, v, l);
  return clearcoat(c, n, v, l) + base * clearcoat_base_weight(c, n, v, l);
}

// ---------------------------------------------------------------- pdf
DI float spec_lobe_pdf(const Closure& c, const Frame3& F, f3 v, f3 d) {
  f3 h = nrm3(v + d);
  f3 lh = to_local(F, h);
  float nh = fabsf(lh.z);
  float vh = gmax(dot3(v, h), 0.0f);
  if (nh <= 0.0f || vh <= 0.0f) return 0.0f;
  return D_aniso(alpha_xy(c.rough, c.aniso), lh) * nh / gmax(4.0f * vh, E5);
}
DI float thick_trans_pdf(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  f3 lv = to_local(F, v);
  float nv = fabsf(lv.z);
  float nl = fabsf(to_local(F, l).z);
  if (nv <= E5 || nl <= E5 || dot3(n, v) * dot3(n, l) >= 0.0f) return 0.0f;
  float eta = c.ior / gmax(1.0f, E5);
  f3 h = -nrm3(v + l * eta);
  f3 lh = to_local(F, h);
  float vh = dot3(v, h), lhd = dot3(l, h);
  if (fabsf(lh.z) <= E5 || vh * lhd >= 0.0f) return 0.0f;
  float2 a = alpha_xy(c.rough, c.aniso);
  return D_aniso(a, lh) * G1_aniso(a, lv) * (eta * eta * fabsf(vh) * fabsf(lhd) / gmax(pm::ppow(vh + eta * lhd, 2.0f) * nv, E5));
}
DI float dielectric_trans_pdf(const Closure& c, const Frame3& F, f3 n, f3 v, f3 l) {
  TransCoeffs co = trans_coeffs(c, n, v);
  if (dot3(n, v) * dot3(n, l) < 0.0f) {
    float tp = gmax(co.tp, E5);
    float lobe = c.thin ? spec_lobe_pdf(c, F, v, flip_across(l, n)) : thick_trans_pdf(c, F, n, v, l);
    return gmax(tp * lobe, E5);
  }
  return gmax(gclamp(co.rp, 0.0f, 1.0f) * spec_lobe_pdf(c, F, v, l), E5);
}
DI float cos_pdf(f3 n, f3 l) { return gmax(dot3(n, l), 0.0f) * INV_PI; }
DI float clearcoat_pdf(const Closure& c, f3 n, f3 v, f3 l) {
  if (!(gmax(dot3(n, l), 0.0f) > 0.0f)) return 0.0f;
  f3 h = nrm3(v + l);
  float nh = gmax(dot3(n, h), 0.0f), vh = gmax(dot3(v, h), 0.0f);
  if (nh <= 0.0f || vh <= 0.0f) return 0.0f;
  return D_iso(gmax(c.ccRough * c.ccRough, 0.0f), nh) * nh / gmax(4.0f * vh, E5);
}
DI float thin_base_prob(const Closure& c, f3 n, f3 v) {
  return gclamp(1.0f - max3c(schlick(thin_spec_f0(c), thin_spec_f90(c), gmax(fabsf(dot3(n, v)), E5))), 0.0f, 1.0f);
}
DI float base_layer_weight(const Closure& c) {
  return gmax(c.diffuseWeight + c.sheenWeight + c.diffTransWeight + c.transmissionWeight, E5);
}
// pdfMaterialBsdfMixture (KF:2413-2615)
DI float pdf(const Closure& c, const Frame3& F, f3 n, f3 v, f3 d) {
  float nd = dot3(n, d);
  float nl = gmax(nd, 0.0f);
  float ccl = gmax(dot3(n, d), 0.0f);
  bool transSide = dot3(n, v) * nd < 0.0f;
  bool thinT = thin_transmissive(c);
  float tb = thinT ? thin_base_prob(c, n, v) : 0.0f;
  float ts = thinT ? 1.0f - tb : 0.0f;
  if (transSide) {
    float p;
    if (thinT) {
      float blw = base_layer_weight(c);
      p = c.diffTransWeight > 0.0f ? tb * (c.diffTransWeight / blw) * (gmax(dot3(-n, d), 0.0f) * INV_PI) : 0.0f;
      if (c.transmissionWeight > 0.0f) p = p + tb * (c.transmissionWeight / blw) * spec_lobe_pdf(c, F, v, flip_across(d, n));
      return p > 0.0f ? gmax(p, E5) : 0.0f;
    }
    float tw = gmax(c.diffTransWeight + c.transmissionWeight, E5);
    p = c.diffTransWeight > 0.0f ? c.diffTransWeight / tw * (gmax(dot3(-n, d), 0.0f) * INV_PI) : 0.0f;
    if (c.transmissionWeight > 0.0f) p = p + c.transmissionWeight / tw * dielectric_trans_pdf(c, F, n, v, d);
    return p > 0.0f ? gmax(p, E5) : 0.0f;
  }
  if (nl <= 0.0f) return 0.0f;
  if (thinT) {
    float blw = base_layer_weight(c);
    float tot = gmax(tb + ts + c.clearcoatWeight, E5);
    float dp = tb * c.diffuseWeight / blw / tot * cos_pdf(n, d);
    float sp = tb * c.sheenWeight / blw / tot * cos_pdf(n, d);
    float spp = ts / tot * spec_lobe_pdf(c, F, v, d);
    float ccw = c.clearcoatWeight / tot;
    float cc = ccl > 0.0f ? clearcoat_pdf(c, n, v, d) : 0.0f;
    return gmax(dp + sp + spp + ccw * cc, E5);
  }
  float dvf = max3c(schlick(c.dielF0, sp3(c.dielF90), gmax(dot3(n, v), 0.0f)));
  float omf = 1.0f - dvf;
  float dpw = c.diffuseWeight * omf, spw = c.sheenWeight * omf;
  float specw = c.specularWeight + c.dielectricWeight * (1.0f - c.transparency) * dvf;
  float tot = gmax(dpw + spw + c.diffTransWeight + specw + c.transmissionWeight + c.clearcoatWeight, E5);
  float dp = dpw / tot * cos_pdf(n, d);
  float sp = spw / tot * cos_pdf(n, d);
  float spp = specw / tot * spec_lobe_pdf(c, F, v, d);
  float tp = c.transmissionWeight > 0.0f ? c.transmissionWeight / tot * dielectric_trans_pdf(c, F, n, v, d) : 0.0f;
  float ccw = c.clearcoatWeight / tot;
  float cc = ccl > 0.0f ? clearcoat_pdf(c, n, v, d) : 0.0f;
  return gmax(dp + sp + spp + tp + ccw * cc, E5);
}

// ---------------------------------------------------------------- sampling
struct Sample { f3 dir; float pdf; f3 w; float specular; };

DI void branchless_onb(f3 n, f3& t, f3& b) {  // orthonormalBasis, KF:2631-2659
  float sg = n.z >= 0.0f ? 1.0f : -1.0f;
  float a = -1.0f / (sg + n.z);
  float bb = n.x * n.y * a;
  t = nrm3(f3{1.0f + sg * n.x * n.x * a, sg * bb, -sg * n.x});
  b = nrm3(f3{bb, sg + n.y * n.y * a, -n.y});
}
DI f3 cosine_dir(f3 n, float u0, float u1) {
  float r = sqrtf(u0);
  float phi = TWO_PIf * u1;
  float sphi, cphi;
  pm::psincos(phi, sphi, cphi);
  float x = r * cphi, y = r * sphi;
  float z = sqrtf(gmax(0.0f, 1.0f - u0));
  f3 t, b;
  branchless_onb(n, t, b);
  return nrm3(t * x + b * y + n * z);
}
DI f3 ggx_half(const Closure& c, const Frame3& F, float u0, float u1) {
  float2 a = alpha_xy(c.rough, c.aniso);
  float phi = TWO_PIf * u1;
  float sp, cp;
  pm::psincos(phi, sp, cp);
  float ct = 1.0f / sqrtf(1.0f + u0 / gmax(1.0f - u0, E6) / gmax(cp * cp / gmax(a.x * a.x, E12) + sp * sp / gmax(a.y * a.y, E12), E6));
  float st = sqrtf(gmax(0.0f, 1.0f - ct * ct));
  return nrm3(F.t * (st * cp) + F.b * (st * sp) + F.n * ct);
}
DI f3 ggx_half_iso(f3 n, float rough, float u0, float u1) {
  float a = gmax(rough * rough, 0.0f);
  float phi = TWO_PIf * u1;
  float ct = sqrtf((1.0f - u0) / gmax(1.0f + (a * a - 1.0f) * u0, E6));
  float st = sqrtf(gmax(0.0f, 1.0f - ct * ct));
  float sphi, cphi;
  pm::psincos(phi, sphi, cphi);
  float x = st * cphi, y = st * sphi;
  f3 t, b;
  branchless_onb(n, t, b);
  return nrm3(t * x + b * y + n * ct);
}
DI float half_pdf(const Closure& c, const Frame3& F, f3 h) {
  f3 lh = to_local(F, h);
  float nh = fabsf(lh.z);
  if (nh <= E5) return 0.0f;
  return D_aniso(alpha_xy(c.rough, c.aniso), lh) * nh;
}
DI f3 refract_safe(f3 i, f3 n, float eta) {
  float ni = dot3(n, i);
  float k = 1.0f - eta * eta * (1.0f - ni * ni);
  if (k <= 0.0f) return sp3(0.0f);
  return eta * i - (eta * ni + sqrtf(k)) * n;
}
// sampleDielectricTransmission (KF:2881-3145)
DI Sample sample_dielectric(const Closure& c, const Frame3& F, f3 n, f3 v, float u0, float u1, float eventU) {
  Sample s;
  if (!c.thin && fabsf(1.0f - c.ior) <= E4) {
    s.dir = -v; s.pdf = 1.0f; s.w = trans_color(c); s.specular = 2.0f;
    return s;
  }
  f3 h = ggx_half(c, F, u0, u1);
  TransCoeffs co = trans_coeffs(c, n, v);
  float hp = half_pdf(c, F, h);
  float pr = gclamp(co.rp, 0.0f, 1.0f);
  float tp = gmax(co.tp, E5);
  if (eventU <= pr) {
    f3 d = reflect3(-v, h);
    float lobe = hp / gmax(4.0f * fabsf(dot3(d, h)), E5);
    float p = gmax(pr * lobe, E5);
    s.dir = d; s.pdf = p; s.w = dielectric_reflection(c, F, n, v, d) / p; s.specular = 1.0f;
    return s;
  }
  if (c.thin) {
    f3 d = flip_across(reflect3(-v, h), n);
    float lobe = hp / gmax(4.0f * fabsf(dot3(d, h)), E5);
    float p = gmax(tp * lobe, E5);
    s.dir = d; s.pdf = p; s.w = thin_transmission(c, F, n, v, d) / p; s.specular = 1.0f;
    return s;
  }
  bool disp = c.dispersion > E5;
  float chU = gclamp((eventU - pr) / gmax(tp, E5), 0.0f, 0.99999898672103882f);
  f3 mask = sp3(1.0f);
  float tIor = c.ior;
  if (disp) {
    mask = chU < 0.3333333432674408f ? f3{1.0f, 0.0f, 0.0f} : (chU < 0.66666668653488159f ? f3{0.0f, 1.0f, 0.0f} : f3{0.0f, 0.0f, 1.0f});
    tIor = dot3(disp_iors(c), mask);
  }
  float chP = disp ? 0.3333333432674408f : 1.0f;
  f3 mv = -v;
  f3 rf = refract_safe(mv, h, 1.0f / gmax(tIor, E5));
  if (len3(rf) <= E5) {
    f3 d = reflect3(mv, h);
    float p = gmax(dielectric_trans_pdf(c, F, n, v, d), E5);
    s.dir = d; s.pdf = p; s.w = dielectric_reflection(c, F, n, v, d) / p; s.specular = 3.0f;
    return s;
  }
  float vh = dot3(v, h);
  float rh = dot3(rf, h);
  float lobe = hp * (tIor * tIor) * fabsf(rh) / gmax(pm::ppow(1.0f * vh + tIor * rh, 2.0f), E5);
  float g1 = G1_aniso(alpha_xy(c.rough, c.aniso), to_local(F, rf));
  f3 etc = co.tc;
  if (c.irid > E5) {
    float avh = fabsf(vh);
    f3 rr = schlick_dielectric(avh, c.dielF0, sp3(c.dielF90), tIor, false);
    etc = (sp3(1.0f) - film_dielectric(c, avh, rr)) * trans_color(c);
  }
  float ctp = tp * chP;
  s.dir = rf;
  s.pdf = gmax(ctp * lobe, E5);
  s.w = etc * mask / gmax(ctp, E5) * g1;
  s.specular = 1.0f;
  return s;
}

// sampleMaterialBsdfMixture (KF:3200-3567). u = {component, lobe.x, lobe.y, boundary}.
// `lobe_id` reports the branch taken (diagnostics / path signatures).
// sampleMaterialBsdfMixture (KF:3200-3567). The lobe choice, the
// dielectric sample and the mixture pdf each have one call site (the
// branches of the reference share them), which keeps the shading kernel's
// code small.
DI Sample sample(const Closure& c, const Frame3& F, f3 n, f3 v, float cu, float u0, float u1, float eu) {
  Sample s;
  f3 dir = sp3(0.0f);
  float spec = 0.0f;
  const bool thin = thin_transmissive(c);
  bool dielectric = false;  // the transmission lobe: sample_dielectric
  float deu = eu, transP = 0.0f;
  if (thin) {
    float tb = thin_base_prob(c, n, v);
    float ts = 1.0f - tb;
    float blw = base_layer_weight(c);
    float tot = gmax(tb + ts + c.clearcoatWeight, E5);
    float pd = tb * (c.diffuseWeight / blw) / tot;
    float psh = tb * (c.sheenWeight / blw) / tot;
    float pdt = tb * (c.diffTransWeight / blw) / tot;
    float ptt = tb * (c.transmissionWeight / blw) / tot;
    float psp = ts / tot;
    float c1 = pd + psh, c2 = c1 + pdt, c3 = c2 + ptt;
    if (cu < pd) { dir = cosine_dir(n, u0, u1); spec = 0.0f; }
    else if (cu < c1) { dir = cosine_dir(n, u0, u1); spec = 0.0f; }
    else if (cu < c2) { dir = cosine_dir(-n, u0, u1); spec = 0.0f; }
    else if (cu < c3) { dielectric = true; deu = 1.0f; }
    else if (cu < c3 + psp) { dir = reflect3(-v, ggx_half(c, F, u0, u1)); spec = 1.0f; }
    else { dir = reflect3(-v, ggx_half_iso(n, c.ccRough, u0, u1)); spec = 1.0f; }
  } else {
    float dvf = max3c(schlick(c.dielF0, sp3(c.dielF90), gmax(dot3(n, v), 0.0f)));
    float omf = 1.0f - dvf;
    float dpw = c.diffuseWeight * omf, spw = c.sheenWeight * omf;
    float specw = c.specularWeight + c.dielectricWeight * (1.0f - c.transparency) * dvf;
    float bpw = dpw + spw + c.diffTransWeight;
    float tot = gmax(bpw + specw + c.transmissionWeight + c.clearcoatWeight, E5);
    float baseP = bpw / tot, specP = specw / tot;
    transP = c.transmissionWeight / tot;
    if (cu < baseP) {
      float bu = cu / gmax(baseP, E5);
      float dp = bpw > 0.0f ? dpw / bpw : 0.0f;
      float sp = bpw > 0.0f ? spw / bpw : 0.0f;
      if (bu < dp) dir = cosine_dir(n, u0, u1);
      else if (bu < dp + sp) dir = cosine_dir(n, u0, u1);
      else dir = cosine_dir(-n, u0, u1);
      spec = 0.0f;
    } else if (cu < baseP + specP) {
      dir = reflect3(-v, ggx_half(c, F, u0, u1)); spec = 1.0f;
    } else if (cu < baseP + specP + transP) {
      dielectric = true;
    } else {
      dir = reflect3(-v, ggx_half_iso(n, c.ccRough, u0, u1)); spec = 1.0f;
    }
  }
  float nd = 0.0f, nl = 0.0f;
  bool need_pdf;
  f3 pdir;
  if (dielectric) {
    s = sample_dielectric(c, F, n, v, u0, u1, deu);
    need_pdf = thin;
    pdir = s.dir;
  } else {
    nd = dot3(n, dir);
    nl = gmax(nd, 0.0f);
    if (nl <= 0.0f && !(nd < 0.0f)) {
      f3 fb = cosine_dir(n, u0, u1);
      s.dir = fb; s.pdf = gmax(gmax(dot3(n, fb), 0.0f) * INV_PI, E5); s.w = sp3(0.0f); s.specular = 0.0f;
      return s;
    }
    need_pdf = true;
    pdir = dir;
  }
  const float mp = need_pdf ? pdf(c, F, n, v, pdir) : 0.0f;
  if (dielectric) {
    if (thin) {
      s.pdf = mp;
      s.w = dielectric_transmission(c, F, n, v, s.dir) / gmax(mp, E5);
    } else {
      s.w = s.w / gmax(transP, E5);
      if (s.pdf < 1.0f) s.pdf = s.pdf * transP;
    }
    return s;
  }
  if (nl <= 0.0f) {  // nd < 0: diffuse transmission
    float p = gmax(mp, E5);
    s.dir = dir; s.pdf = p; s.specular = 0.0f;
    s.w = thin ? diffuse_transmission(c, n, dir) / gmax(p, E5) : diffuse_transmission(c, n, dir) / p;
    return s;
  }
  f3 rb = reflected_base(c, F, n, v, dir) / gmax(nl, E5);
  f3 cce = nl > 0.0f ? clearcoat(c, n, v, dir) / gmax(nl, E5) : sp3(0.0f);
  float bw = clearcoat_base_weight(c, n, v, dir);
  s.dir = dir; s.pdf = mp; s.specular = spec;
  s.w = thin ? (rb * bw + cce) * nl / gmax(mp, E5) : (rb * bw + cce) * fabsf(nd) / mp;
  return s;
}

}  // namespace bsdf
