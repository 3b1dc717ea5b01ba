// CPU ORACLE — test infrastructure only (see pt_oracle.hpp).
//
// Restatement of the slang-pbr Enterprise PBR kernel as generated into GLSL
// (KF = .../webgl-full/material_kernel.glsl, KL = .../webgl-lean/...).
// Function names follow the generated names without the `_N` suffixes;
// each cites its KF line (KL line where the lean file differs in position).
// The adapter always passes currentMediumIor = 1 (SH/pbr_material_adapter.glsl:18),
// but the restatement keeps the general form of every ior branch.
#pragma once
#include "pt_oracle.hpp"

namespace orc {

template <class R> struct PbrMat {  // PbrGltfMaterial, KF facade
  V3<R> baseColor; R opacity;
  R metallic, roughness;
  V3<R> emissive; R emissiveStrength;
  R specular; V3<R> specularColor;
  R transmission, diffuseTransmission; V3<R> diffuseTransmissionColor;
  R ior; V3<R> attenuationColor; R attenuationDistance; R thickness;
  V3<R> multiscatterColor; R scatterAnisotropy;
  R clearcoat, clearcoatRoughness, clearcoatNormalScale;
  V3<R> sheenColor; R sheenRoughness;
  R anisotropyStrength, anisotropyRotation;
  R iridescence, iridescenceIor, iridescenceThickness, dispersion, normalScale;
};

template <class R> struct Closure {  // SurfaceClosure, KF:330-356 (KL:320-355)
  V3<R> diffuseColor, diffuseReflectionColor, transmissionColor;
  R diffuseWeight, sheenWeight, anisotropy;
  V3<R> anisotropyTangent, specularColor, specularF90, dielectricSpecularColor, dielectricSpecularF90;
  R dielectricWeight, specularWeight, clearcoatWeight, roughness;
  V3<R> sheenColor; R sheenRoughness, transparency, clearcoat, clearcoatRoughness;
  V3<R> throughput;
  R metallic, specular, ior, iridescence, iridescenceIor, iridescenceThickness, dispersion, thinWalled;
  V3<R> attenuationColor; R attenuationDistance;
  V3<R> multiscatterColor; R scatterAnisotropy, transmissionWeight;
  V3<R> diffuseTransmissionColor; R diffuseTransmissionWeight;
};

template <class R> struct BState { Closure<R> c; R currentMediumIor; };
template <class R> struct BSample { V3<R> dir; R pdf; V3<R> bsdfOverPdf; R specular, crossedBoundary, nextMediumIor; };

#define F1E5 R(0.00000999999974738)
#define F1E4 R(0.00009999999747379)
#define FPI R(3.14159274101257324)
#define F1PI R(0.31830987334251404)

// ---- analytic energy approximations, KF:5-91
template <class R> R approxAvgAlbedoGGX(R alpha) {
  R r = std::sqrt(gmax(alpha, R(0)));
  return gclamp(R(1) - R(0.57999998331069946) * r + R(0.07999999821186066) * r * r, R(0), R(1));
}
template <class R> R approxAvgAlbedoGGXMs(R alpha, R e0) {
  return gclamp(gclamp(e0 + (R(1) - e0) * R(0.0476190485060215), R(0), R(1)) * approxAvgAlbedoGGX(alpha), R(0), R(1));
}
template <class R> R approxDirAlbedoGGX(R alpha, R cosT) {
  return gclamp(R(1) - std::sqrt(gmax(alpha, R(0))) * (R(0.5) + R(0.5) * opow(R(1) - gclamp(cosT, R(0), R(1)), R(5))), R(0), R(1));
}
template <class R> R approxDirAlbedoGGXMs(R theta, R alpha, R e0) {
  R noV = gclamp(theta, R(0), R(1));
  return gclamp(gclamp(e0 + (R(1) - e0) * opow(R(1) - noV, R(5)), R(0), R(1)) * gmix(R(1), approxDirAlbedoGGX(alpha, noV), R(0.64999997615814209)), R(0), R(1));
}

// ---- closure build, KF:209-261 buildEnterpriseSurfaceFromGltf + KF:357-422 buildSurfaceClosure
template <class R> R iridescenceInterfaceWeight(R irid, R filmIor) {  // KF:307
  R s = gmax(filmIor, R(1));
  return gclamp(irid, R(0), R(1)) * opow((s - R(1)) / gmax(s + R(1), F1E5), R(2));
}
template <class R> Closure<R> buildClosure(const PbrMat<R>& m, V3<R> anisotropyTangent, int profile) {
  const V3<R> Z(R(0)), O(R(1));
  // buildEnterpriseSurfaceFromGltf
  V3<R> albedo = vclamp(m.baseColor, Z, O);
  R metallic = gclamp(m.metallic, R(0), R(1));
  R roughness = gclamp(m.roughness, R(0), R(1));
  R anisotropy = gclamp(m.anisotropyStrength, R(0), R(1));
  R transparency = gclamp(m.transmission, R(0), R(1));
  R ior = gmax(m.ior, R(1));
  V3<R> specularColor = vclamp(m.specularColor, Z, O);
  R specular = gclamp(m.specular, R(0), R(1));
  V3<R> attColor = vmax(m.attenuationColor, V3<R>(F1E4));
  R attDist = gmax(m.attenuationDistance, F1E4);
  V3<R> msColor = vclamp(m.multiscatterColor, Z, O);
  R scatterAniso = gclamp(m.scatterAnisotropy, R(-0.99900001287460327), R(1));
  R thinWalled = m.thickness > R(0) ? R(0) : R(1);
  R translucency = gclamp(m.diffuseTransmission, R(0), R(1));
  V3<R> translucencyColor = vclamp(m.diffuseTransmissionColor, Z, O);
  R iridescence = gclamp(m.iridescence, R(0), R(1));
  R iridescenceIor = gclamp(m.iridescenceIor, R(1), R(2.5));
  R iridescenceThickness = gmax(m.iridescenceThickness, R(0));
  R dispersion = gmax(m.dispersion, R(0));
  R clearcoat = gclamp(m.clearcoat, R(0), R(1));
  R clearcoatRoughness = gclamp(m.clearcoatRoughness, R(0), R(1));
  V3<R> sheenColor = vclamp(m.sheenColor, Z, O);
  R sheenRoughness = gclamp(m.sheenRoughness, R(0), R(1));
  if (profile == 0) { iridescence = R(0); dispersion = R(0); }  // KL:395,409,412

  // buildSurfaceClosure
  Closure<R> c;
  R dielectricWeight = R(1) - metallic;
  R s9 = R(1) - transparency;
  R opaqueDielectricWeight = dielectricWeight * s9;
  R transparentDielectricWeight = dielectricWeight * transparency;
  V3<R> diffRefl = albedo * opaqueDielectricWeight * (R(1) - translucency);
  V3<R> diffTrans = opaqueDielectricWeight * translucencyColor * translucency;
  R iorRatio = (R(1) - ior) / (R(1) + ior);
  V3<R> dielF0 = vclamp(V3<R>(iorRatio * iorRatio * specular) * specularColor, Z, O);
  V3<R> metalSpecEnergy = albedo + (O - albedo) * R(0.0476190485060215);
  V3<R> weightedTrans = albedo * transparentDielectricWeight;
  R matRough = gclamp(gmax(roughness, R(0.0020000000949949)), R(0), R(1));
  R ccRough = gclamp(gmax(clearcoatRoughness, R(0.0020000000949949)), R(0), R(1));
  c.diffuseColor = albedo;
  c.diffuseReflectionColor = diffRefl;
  c.transmissionColor = albedo;
  c.diffuseWeight = max3(diffRefl);
  c.sheenWeight = max3(sheenColor) * s9;
  c.anisotropy = anisotropy;
  c.anisotropyTangent = anisotropyTangent;
  c.specularColor = albedo;
  c.specularF90 = O;
  c.dielectricSpecularColor = dielF0;
  c.dielectricSpecularF90 = V3<R>(specular);
  c.dielectricWeight = dielectricWeight;
  R filmWeight = iridescenceInterfaceWeight(iridescence, iridescenceIor);
  c.specularWeight = gmax(max3(metalSpecEnergy), filmWeight) * metallic;
  c.clearcoatWeight = clearcoat * gmix(R(0.20000000298023224), R(1), R(1) - ccRough);
  c.roughness = matRough;
  c.sheenColor = sheenColor;
  c.sheenRoughness = sheenRoughness;
  c.transparency = transparency;
  c.clearcoat = clearcoat;
  c.clearcoatRoughness = ccRough;
  c.throughput = vclamp(diffRefl + diffTrans + weightedTrans + vmax(metalSpecEnergy, V3<R>(filmWeight)) * metallic +
                            dielF0 * dielectricWeight + sheenColor, Z, O);
  c.metallic = metallic;
  c.specular = specular;
  c.ior = ior;
  c.iridescence = iridescence;
  c.iridescenceIor = iridescenceIor;
  c.iridescenceThickness = iridescenceThickness;
  c.dispersion = dispersion;
  c.thinWalled = thinWalled;
  c.attenuationColor = attColor;
  c.attenuationDistance = attDist;
  c.multiscatterColor = msColor;
  c.scatterAnisotropy = scatterAniso;
  c.transmissionWeight = max3(albedo) * transparentDielectricWeight;
  c.diffuseTransmissionColor = diffTrans;
  c.diffuseTransmissionWeight = max3(diffTrans);
  return c;
}

// ---- microfacet helpers, KF:470-640
template <class R> V3<R> evalDiffuseTransmission(const Closure<R>& c, V3<R> n, V3<R> l) {
  return c.diffuseTransmissionColor * F1PI * gmax(std::fabs(dot(n, l)), R(0));
}
template <class R> R transparentDielectricWeight(const Closure<R>& c) { return c.dielectricWeight * c.transparency; }
template <class R> V3<R> projectAniso(V3<R> n, V3<R> t, V3<R> d) {
  V3<R> st = normalize(t - n * dot(t, n));
  return V3<R>(dot(d, st), dot(d, normalize(cross(n, st))), dot(d, n));
}
template <class R> V2<R> anisoAlpha(R rough, R aniso) {
  R ca = gclamp(aniso, R(0), R(1));
  R s = gmax(rough * rough, R(0));
  return V2<R>{gmix(s, R(1), ca * ca), s};
}
template <class R> R ggxDAniso(V2<R> a, V3<R> h) {
  R hx = h.x / a.x, hy = h.y / a.y, hz = h.z;
  R den = hx * hx + hy * hy + hz * hz;
  return R(1) / gmax(FPI * a.x * a.y * den * den, R(9.999999960041972e-13));
}
template <class R> R ggxLambdaAniso(V2<R> a, V3<R> d) {
  R c2 = d.z * d.z;
  if (c2 <= R(9.99999993922529029e-09)) return R(0);
  if (gmax(R(1) - c2, R(0)) <= R(9.99999993922529029e-09)) return R(0);
  return R(0.5) * (R(-1) + std::sqrt(R(1) + (d.x * d.x * a.x * a.x + d.y * d.y * a.y * a.y) / gmax(c2, R(9.99999993922529029e-09))));
}
template <class R> R ggxG2Aniso(V2<R> a, V3<R> v, V3<R> l) {
  return R(1) / (R(1) + ggxLambdaAniso(a, v) + ggxLambdaAniso(a, l));
}
template <class R> R ggxG1Aniso(V2<R> a, V3<R> d) { return R(1) / (R(1) + ggxLambdaAniso(a, d)); }
// KF:553-615
template <class R> V3<R> microfacetAniso(R rough, R aniso, V3<R> n, V3<R> t, V3<R> v, V3<R> l) {
  V3<R> h = normalize(v + l);
  V3<R> lv = projectAniso(n, t, v), ll = projectAniso(n, t, l), lh = projectAniso(n, t, h);
  R noV = std::fabs(lv.z), noL = std::fabs(ll.z), noH = std::fabs(lh.z);
  R voh = gmax(dot(v, h), R(0));
  if (noV <= R(0) || noL <= R(0) || noH <= R(0) || voh <= R(0)) return V3<R>(R(0));
  V2<R> a = anisoAlpha(rough, aniso);
  return V3<R>(ggxDAniso(a, lh) * ggxG2Aniso(a, lv, ll) / gmax(R(4) * noV * noL, F1E5));
}

// ---- Fresnel, KF:638-745
struct FresnelAngles { double cosT; bool tir; };
template <class R> void dielectricFresnelAngles(R cosTheta, R ni, R nt, bool thin, R& schlickCos, bool& tir) {
  R cosI = gclamp(cosTheta, R(0), R(1));
  R eta = ni / gmax(nt, F1E5);
  R sinT2 = eta * eta * gmax(R(1) - cosI * cosI, R(0));
  if (sinT2 >= R(1) && !thin) { schlickCos = R(0); tir = true; return; }
  R cosT = std::sqrt(gmax(R(1) - sinT2, R(0)));
  schlickCos = (nt >= ni || thin) ? cosI : cosT;
  tir = false;
}
template <class R> V3<R> fresnelSchlick(V3<R> f0, V3<R> f90, R cosTheta) {
  R m = gclamp(R(1) - cosTheta, R(0), R(1));
  R m2 = m * m;
  return f0 + (f90 - f0) * (m2 * m2 * m);
}
template <class R> V3<R> fresnelSchlickDielectric(R cosTheta, V3<R> f0, V3<R> f90, R ni, R nt, bool thin) {
  if (std::fabs(ni - nt) <= F1E4) return V3<R>(R(0));
  R sc; bool tir;
  dielectricFresnelAngles(cosTheta, ni, nt, thin, sc, tir);
  V3<R> s = fresnelSchlick(f0, f90, sc);
  return tir ? V3<R>(R(1)) : s;
}
template <class R> V3<R> fresnelSchlickAverage(V3<R> f0, V3<R> f90) { return f0 + (f90 - f0) * R(0.0476190485060215); }

// ---- thin film (KF:741-936, full profile only)
template <class R> R thinFilmIorToF0(R tIor, R iIor) {
  R r = (tIor - iIor) / gmax(tIor + iIor, F1E5);
  return r * r;
}
template <class R> V3<R> thinFilmDielectricSubstrateF0(R ior, R outside) {
  return V3<R>(thinFilmIorToF0(gmax(ior, R(1)), gmax(outside, R(1))));
}
template <class R> V3<R> thinFilmF0ToIor(V3<R> f0) {
  V3<R> s = vsqrt(vclamp(f0, V3<R>(R(0)), V3<R>(R(0.99989998340606689))));
  return (V3<R>(R(1)) + s) / vmax(V3<R>(R(1)) - s, V3<R>(F1E5));
}
template <class R> V3<R> thinFilmSensitivity(R opd, V3<R> shift) {
  R phase = R(6.28318548202514648) * opd * R(9.99999971718068537e-10);
  R phase2 = phase * phase;
  const V3<R> var(R(4.327799808e+09), R(9.304599552e+09), R(6.612100096e+09));
  V3<R> xyz = V3<R>(R(5.48560003844900113e-13), R(4.42010008238832852e-13), R(5.24810013155518895e-13)) *
              vsqrt(R(6.28318548202514648) * var) *
              vcos(V3<R>(R(1.681e+06), R(1.7953e+06), R(2.2084e+06)) * phase + shift) * vexp(-phase2 * var);
  xyz.x = xyz.x + R(9.74699989231449238e-14) * std::sqrt(R(2.845152256e+10)) * ocos(R(2.2399e+06) * phase + shift.x) *
                      oexp(R(-4.528200192e+09) * phase2);
  V3<R> s = xyz / R(1.06850002623559703e-07);
  R X = s.x, Y = s.y, Z = s.z;
  return V3<R>(R(3.24045419692993164) * X - R(1.53713846206665039) * Y - R(0.49853140115737915) * Z,
               R(-0.96926599740982056) * X + R(1.87601077556610107) * Y + R(0.04155600070953369) * Z,
               R(0.05564339831471443) * X - R(0.20402589440345764) * Y + R(1.05722522735595703) * Z);
}
template <class R> V3<R> thinFilmReflectance(R outside, R filmIor, V3<R> baseF0, R thickness, R cos1) {
  R fi = gclamp(filmIor, R(1), R(2.5));
  R c1 = gclamp(cos1, R(0), R(1));
  R c2sq = R(1) - outside / fi * (outside / fi) * (R(1) - c1 * c1);
  if (c2sq < R(0)) return V3<R>(R(1));
  R c2 = std::sqrt(c2sq);
  const V3<R> O(R(1));
  R r12 = fresnelSchlick(V3<R>(thinFilmIorToF0(fi, outside)), O, c1).x;
  R t121 = R(1) - r12;
  V3<R> baseIor = thinFilmF0ToIor(baseF0);
  V3<R> fiv(fi);
  V3<R> r1r = (baseIor - fiv) / vmax(baseIor + fiv, V3<R>(F1E5));
  V3<R> r23 = fresnelSchlick(r1r * r1r, O, c2);
  R opd = R(2) * fi * gmax(thickness, R(0)) * c2;
  R phi12 = fi < outside ? FPI : R(0);
  R phi21 = FPI - phi12;
  V3<R> phi23(baseIor.x < fi ? FPI : R(0), baseIor.y < fi ? FPI : R(0), baseIor.z < fi ? FPI : R(0));
  V3<R> phi = V3<R>(phi21) + phi23;
  V3<R> r12v(r12);
  V3<R> r123 = vclamp(r12v * r23, V3<R>(F1E5), V3<R>(R(0.99989998340606689)));
  V3<R> sr = vsqrt(r123);
  V3<R> rs = t121 * t121 * r23 / vmax(O - r123, V3<R>(F1E5));
  V3<R> refl = r12v + rs;
  V3<R> cm = rs - V3<R>(t121);
  for (int m = 1; m <= 2; m++) {
    cm = cm * sr;
    R fm = R(m);
    refl = refl + cm * R(2) * thinFilmSensitivity(fm * opd, fm * phi);
  }
  return vmax(refl, V3<R>(R(0)));
}
template <class R> V3<R> thinFilmMix(V3<R> regular, V3<R> film, R irid) {
  R a = gclamp(irid, R(0), R(1));
  return vmix(regular, film, V3<R>(a));
}
template <class R> V3<R> thinFilmDielectricFresnel(const Closure<R>& c, R cosT, V3<R> regular, R incidentIor) {
  if (c.iridescence <= F1E5) return regular;
  return thinFilmMix(regular,
                     thinFilmReflectance(gmax(incidentIor, R(1)), c.iridescenceIor,
                                         thinFilmDielectricSubstrateF0(c.ior, incidentIor), c.iridescenceThickness, cosT),
                     c.iridescence);
}

// ---- dielectric specular/transmission eval, KF:948-1440
template <class R> V3<R> evalDielectricSpecularGGX(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l, R iorI, R iorO) {
  R noL = gmax(dot(n, l), R(0));
  R tw = transparentDielectricWeight(c);
  if (noL <= R(0) || tw <= R(0)) return V3<R>(R(0));
  V3<R> mr = microfacetAniso(c.roughness, c.anisotropy, n, c.anisotropyTangent, v, l);
  R voh = gmax(dot(v, normalize(v + l)), R(0));
  R a = gmax(c.roughness * c.roughness, R(0));
  R eAvg = approxAvgAlbedoGGX(a);
  R oneMinusE = R(1) - eAvg;
  R ms = (R(1) - approxDirAlbedoGGX(a, noL)) * (R(1) - approxDirAlbedoGGX(a, std::fabs(dot(n, v)))) /
         gmax(FPI * oneMinusE, R(9.99999997475242708e-07));
  V3<R> regF = fresnelSchlickDielectric(voh, c.dielectricSpecularColor, c.dielectricSpecularF90, iorI, iorO, false);
  V3<R> f = (c.iridescence > F1E5 && iorI <= R(1.00010001659393311)) ? thinFilmDielectricFresnel(c, voh, regF, R(1)) : regF;
  V3<R> fAvg = (std::fabs(iorI - iorO) <= F1E4) ? V3<R>(R(0)) : fresnelSchlickAverage(c.dielectricSpecularColor, c.dielectricSpecularF90);
  return tw * (mr * f + ms * (fAvg * fAvg * eAvg / vmax(V3<R>(R(1)) - fAvg * oneMinusE, V3<R>(F1E5)))) * noL;
}
template <class R> V3<R> weightedTransmissionColor(const Closure<R>& c) { return c.transmissionColor * transparentDielectricWeight(c); }
template <class R> R maxWeightedComponent(V3<R> w, V3<R> v) { return max3(vmax(w, V3<R>(R(0))) * vmax(v, V3<R>(R(0)))); }
template <class R> V2<R> reflTransProbs(V3<R> tp, V3<R> rc, V3<R> tc) {
  R rw = maxWeightedComponent(tp, rc), tw = maxWeightedComponent(tp, tc);
  R tot = rw + tw;
  if (tot <= F1E5) return V2<R>{R(1), R(0)};
  return V2<R>{rw / tot, tw / tot};
}
template <class R> struct TransCoeffs { V3<R> rc, tc; R rp, tp; };
template <class R> TransCoeffs<R> transCoeffs(const Closure<R>& c, R curIor, V3<R> n, V3<R> v, R iorI, R iorO) {
  R cosV = gmax(dot(n, v), F1E5);
  V3<R> fr = fresnelSchlickDielectric(cosV, c.dielectricSpecularColor, c.dielectricSpecularF90, iorI, iorO, false);
  V3<R> film = (c.iridescence > F1E5 && iorI <= R(1.00010001659393311)) ? thinFilmDielectricFresnel(c, cosV, fr, R(1)) : fr;
  TransCoeffs<R> t;
  t.rc = film * transparentDielectricWeight(c);
  t.tc = (V3<R>(R(1)) - film) * weightedTransmissionColor(c);
  V2<R> p = reflTransProbs(vmax(c.throughput, V3<R>(gmax(curIor, F1E5))), t.rc, t.tc);
  t.rp = p.x; t.tp = p.y;
  return t;
}
template <class R> V3<R> flipAcross(V3<R> v, V3<R> n) { return normalize(v - R(2) * n * dot(v, n)); }
template <class R> V3<R> evalThinTransmission(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l, R iorI, R iorO) {
  R nl = dot(n, l);
  if (dot(n, v) * nl >= R(0)) return V3<R>(R(0));
  V3<R> ml = flipAcross(l, n);
  R cosH = gmax(dot(v, normalize(v + ml)), F1E5);
  V3<R> rr = fresnelSchlickDielectric(cosH, c.dielectricSpecularColor, c.dielectricSpecularF90, iorI, iorO, true);
  V3<R> fr = (c.iridescence > F1E5 && iorI <= R(1.00010001659393311)) ? thinFilmDielectricFresnel(c, cosH, rr, R(1)) : rr;
  return weightedTransmissionColor(c) * (V3<R>(R(1)) - fr) *
         microfacetAniso(c.roughness, c.anisotropy, n, c.anisotropyTangent, v, ml) * std::fabs(nl);
}
template <class R>
V3<R> evalThickTransmission(const Closure<R>& c, V3<R> tc, V3<R> n, V3<R> v, V3<R> l, R iorI, R iorO) {
  V3<R> lv = projectAniso(n, c.anisotropyTangent, v), ll = projectAniso(n, c.anisotropyTangent, l);
  R noV = std::fabs(lv.z), noL = std::fabs(ll.z);
  if (noV <= F1E5 || noL <= F1E5 || dot(n, v) * dot(n, l) >= R(0)) return V3<R>(R(0));
  R eta = iorO / gmax(iorI, F1E5);
  V3<R> h = -normalize(v + l * eta);
  V3<R> lh = projectAniso(n, c.anisotropyTangent, h);
  R voh = dot(v, h), loh = dot(l, h);
  if (std::fabs(lh.z) <= F1E5 || voh * loh >= R(0)) return V3<R>(R(0));
  V2<R> a = anisoAlpha(c.roughness, c.anisotropy);
  R d = ggxDAniso(a, lh);
  R g = ggxG2Aniso(a, lv, ll);
  R avoh = std::fabs(voh);
  R factor = eta * eta * avoh * std::fabs(loh) / gmax(opow(voh + eta * loh, R(2)) * noV, F1E5);
  V3<R> rr = fresnelSchlickDielectric(avoh, c.dielectricSpecularColor, c.dielectricSpecularF90, iorI, iorO, false);
  V3<R> etc = (c.iridescence > F1E5 && iorI <= R(1.00010001659393311))
                  ? (V3<R>(R(1)) - thinFilmDielectricFresnel(c, avoh, rr, R(1))) * weightedTransmissionColor(c)
                  : tc;
  return etc * d * g * factor;
}
template <class R> V3<R> dispersionIors(const Closure<R>& c) {  // KF:1289-1298
  R s = gmax(c.ior, R(1));
  R scale = (s - R(1)) * gmax(c.dispersion, R(0)) * R(0.05000000074505806);
  return vmax(V3<R>(s + scale * R(-0.30094969272613525), s, s + scale * R(0.69905030727386475)), V3<R>(R(1)));
}
template <class R>
V3<R> evalThickTransmissionDispersed(const Closure<R>& c, V3<R> tc, V3<R> n, V3<R> v, V3<R> l, R iorI, R iorO, bool inside) {
  if (c.dispersion <= F1E5 || c.thinWalled > R(0.5)) return evalThickTransmission(c, tc, n, v, l, iorI, iorO);
  V3<R> iors = dispersionIors(c);
  V3<R> red = evalThickTransmission(c, V3<R>(tc.x, R(0), R(0)), n, v, l, inside ? iors.x : R(1), inside ? R(1) : iors.x);
  V3<R> green = evalThickTransmission(c, V3<R>(R(0), tc.y, R(0)), n, v, l, inside ? iors.y : R(1), inside ? R(1) : iors.y);
  V3<R> blue = evalThickTransmission(c, V3<R>(R(0), R(0), tc.z), n, v, l, inside ? iors.z : R(1), inside ? R(1) : iors.z);
  return V3<R>(red.x, green.y, blue.z);
}
template <class R> void mediumIors(const Closure<R>& c, R curIor, bool& inside, R& iorI, R& iorO) {
  R s = gmax(curIor, R(1));
  inside = (s > R(1.00010001659393311)) ? (c.thinWalled < R(0.5)) : false;
  iorI = inside ? s : R(1);
  iorO = inside ? R(1) : c.ior;
}
template <class R> V3<R> evalDielectricTransmission(const Closure<R>& c, R curIor, V3<R> n, V3<R> v, V3<R> l) {
  bool inside; R iorI, iorO;
  mediumIors(c, curIor, inside, iorI, iorO);
  if (dot(n, v) * dot(n, l) > R(0)) return evalDielectricSpecularGGX(c, n, v, l, iorI, iorO);
  TransCoeffs<R> co = transCoeffs(c, curIor, n, v, iorI, iorO);
  if (c.thinWalled > R(0.5)) return evalThinTransmission(c, n, v, l, iorI, iorO);
  return evalThickTransmissionDispersed(c, co.tc, n, v, l, iorI, iorO, inside);
}

// ---- diffuse / specular / sheen / clearcoat, KF:1505-2089
template <class R> R coupledDiffuseFactor(const Closure<R>& c, R noV, R noL) {
  R a = gmax(c.roughness * c.roughness, R(0));
  R e0 = max3(c.dielectricSpecularColor * c.dielectricWeight);
  return gmix(R(0.31830987334251404),
              (R(1) - approxDirAlbedoGGXMs(noL, a, e0)) * (R(1) - approxDirAlbedoGGXMs(noV, a, e0)) /
                  gmax(FPI * (R(1) - approxAvgAlbedoGGXMs(a, e0)), R(9.99999997475242708e-07)),
              c.specular);
}
template <class R> V3<R> thinSharedSpecularColor(const Closure<R>& c) {
  return c.dielectricSpecularColor * c.dielectricWeight + c.specularColor * c.metallic;
}
template <class R> V3<R> thinSharedSpecularF90(const Closure<R>& c) {
  return c.dielectricSpecularF90 * c.dielectricWeight + V3<R>(c.metallic);
}
template <class R> V3<R> closureDiffuseEnergy(const Closure<R>& c) {
  const V3<R> O(R(1));
  return vclamp(O - (fresnelSchlickAverage(c.specularColor, c.specularF90) * c.metallic +
                     fresnelSchlickAverage(c.dielectricSpecularColor, c.dielectricSpecularF90) * c.dielectricWeight),
                V3<R>(R(0)), O);
}
template <class R> R transmissiveDiffuseCompFactor(const Closure<R>& c, R noV, R noL) {
  R coupled = coupledDiffuseFactor(c, noV, noL);
  bool thin = c.thinWalled > R(0.5);
  V3<R> sc = thin ? thinSharedSpecularColor(c) : c.dielectricSpecularColor;
  V3<R> sf = thin ? thinSharedSpecularF90(c) : c.dielectricSpecularF90;
  R f0 = max3(sc);
  R d = max3(sf) - f0;
  return coupled * gmix(R(1), R(1) / gmax(max3(closureDiffuseEnergy(c)), F1E4),
                        std::sqrt(R(0.5) * (gclamp(f0 + d * opow(R(1) - gclamp(noV, R(0), R(1)), R(5)), R(0), R(1)) +
                                            gclamp(f0 + d * opow(R(1) - gclamp(noL, R(0), R(1)), R(5)), R(0), R(1)))));
}
template <class R> V3<R> evalDiffuseReflection(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  R noL = gmax(dot(n, l), R(0));
  if (c.specular <= F1E5 && c.transparency <= R(0)) return c.diffuseReflectionColor * F1PI * noL;
  R noV = gmax(std::fabs(dot(v, n)), R(0));
  R factor = c.transparency > R(0) ? transmissiveDiffuseCompFactor(c, noV, noL) : coupledDiffuseFactor(c, noV, noL);
  return c.diffuseReflectionColor * factor * noL;
}
template <class R> bool isThinTransmissiveDielectric(const Closure<R>& c) {
  bool b = (c.thinWalled > R(0.5)) ? (c.transparency > R(0) ? true : c.diffuseTransmissionWeight > R(0)) : false;
  return b ? (c.metallic < R(1)) : false;
}
template <class R> V3<R> evalThinSharedSpecularGGX(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  R noL = gmax(dot(n, l), R(0));
  if (noL <= R(0)) return V3<R>(R(0));
  V3<R> mr = microfacetAniso(c.roughness, c.anisotropy, n, c.anisotropyTangent, v, l);
  R voh = gmax(dot(v, normalize(v + l)), R(0));
  R a = gmax(c.roughness * c.roughness, R(0));
  R eAvg = approxAvgAlbedoGGX(a);
  R ome = R(1) - eAvg;
  R ms = (R(1) - approxDirAlbedoGGX(a, noL)) * (R(1) - approxDirAlbedoGGX(a, std::fabs(dot(n, v)))) /
         gmax(FPI * ome, R(9.99999997475242708e-07));
  V3<R> f0 = thinSharedSpecularColor(c), f90 = thinSharedSpecularF90(c);
  V3<R> f = fresnelSchlick(f0, f90, voh);
  if (c.iridescence > F1E5)
    f = thinFilmMix(f, thinFilmReflectance(R(1), c.iridescenceIor, thinFilmDielectricSubstrateF0(c.ior, R(1)), c.iridescenceThickness, voh),
                    c.iridescence);
  V3<R> fAvg = fresnelSchlickAverage(f0, f90);
  return (mr * f + ms * (fAvg * fAvg * eAvg / vmax(V3<R>(R(1)) - fAvg * ome, V3<R>(F1E5)))) * noL;
}
template <class R> V3<R> evalBaseSpecularGGX(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  V3<R> mr = microfacetAniso(c.roughness, c.anisotropy, n, c.anisotropyTangent, v, l);
  R voh = gmax(dot(v, normalize(v + l)), R(0));
  R a = gmax(c.roughness * c.roughness, R(0));
  R eAvg = approxAvgAlbedoGGX(a);
  R ome = R(1) - eAvg;
  R ms = (R(1) - approxDirAlbedoGGX(a, std::fabs(dot(n, l)))) * (R(1) - approxDirAlbedoGGX(a, std::fabs(dot(n, v)))) /
         gmax(FPI * ome, R(9.99999997475242708e-07));
  V3<R> metalF = fresnelSchlick(c.specularColor, c.specularF90, voh);
  bool irid = c.iridescence > F1E5;
  if (irid) metalF = thinFilmMix(metalF, thinFilmReflectance(R(1), c.iridescenceIor, c.specularColor, c.iridescenceThickness, voh), c.iridescence);
  V3<R> metalSS = mr * metalF;
  V3<R> metalFAvg = fresnelSchlickAverage(c.specularColor, c.specularF90);
  const V3<R> O(R(1)), E(F1E5);
  V3<R> metalMS = ms * (metalFAvg * metalFAvg * eAvg / vmax(O - metalFAvg * ome, E));
  V3<R> dielF = fresnelSchlick(c.dielectricSpecularColor, c.dielectricSpecularF90, voh);
  if (irid)
    dielF = thinFilmMix(dielF, thinFilmReflectance(R(1), c.iridescenceIor, thinFilmDielectricSubstrateF0(c.ior, R(1)), c.iridescenceThickness, voh),
                        c.iridescence);
  V3<R> dielFAvg = fresnelSchlickAverage(c.dielectricSpecularColor, c.dielectricSpecularF90);
  return c.metallic * (metalSS + metalMS) +
         c.dielectricWeight * (R(1) - c.transparency) * (mr * dielF + ms * (dielFAvg * dielFAvg * eAvg / vmax(O - dielFAvg * ome, E)));
}
template <class R> V3<R> evalBaseSpecular(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  R noL = gmax(dot(n, l), R(0));
  if (isThinTransmissiveDielectric(c)) return evalThinSharedSpecularGGX(c, n, v, l);
  return evalBaseSpecularGGX(c, n, v, l) * noL;
}
template <class R> R sheenLogApprox(R x, R alpha) {
  R s = R(1) - alpha;
  R o = s * s;
  return gmix(R(21.54730033874511719), R(25.32449913024902344), o) /
             (R(1) + gmix(R(3.82986998558044434), R(3.32435011863708496), o) *
                         opow(std::fabs(x), gmix(R(0.19822999835014343), R(0.16800999641418457), o))) +
         gmix(R(-1.97759997844696045), R(-1.27392995357513428), o) * x +
         gmix(R(-4.32053995132446289), R(-4.85967016220092773), o);
}
template <class R> R lambdaSheen(R cosT, R alpha) {
  if (std::fabs(cosT) >= R(0.5))
    return oexp(R(2) * sheenLogApprox(R(0.5), alpha) - sheenLogApprox(R(1) - cosT, alpha));
  return oexp(sheenLogApprox(cosT, alpha));
}
template <class R> V3<R> evalSheen(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  if (max3(c.sheenColor) <= R(0)) return V3<R>(R(0));
  R noV = gmax(dot(n, v), R(0)), noL = gmax(dot(n, l), R(0)), noH = gmax(dot(n, normalize(v + l)), R(0));
  if (noV <= R(0) || noL <= R(0) || noH <= R(0)) return V3<R>(R(0));
  R alpha = gmax(c.sheenRoughness * c.sheenRoughness, R(0));
  R inv = R(1) / gmax(alpha, F1E5);
  return c.sheenColor *
         (R(1) / (R(1) + lambdaSheen(noV, alpha) + lambdaSheen(noL, alpha)) *
          ((R(2) + inv) * opow(std::fabs(gmax(R(1) - noH * noH, R(0.00100000004749745))), R(0.5) * inv) / R(6.28318548202514648)) /
          gmax(R(4) * noV * noL, F1E5)) *
         noL;
}
template <class R> V3<R> evalReflectedBase(const BState<R>& s, V3<R> n, V3<R> v, V3<R> l) {
  return evalDiffuseReflection(s.c, n, v, l) + evalBaseSpecular(s.c, n, v, l) + evalSheen(s.c, n, v, l);
}
template <class R> R ggxD(R alpha, R noH) {
  R a2 = alpha * alpha;
  R d = noH * noH * (a2 - R(1)) + R(1);
  return a2 / gmax(FPI * d * d, R(9.999999960041972e-13));
}
template <class R> R ggxLambda(R alpha, R noX) {
  R ct = std::fabs(noX);
  R c2 = ct * ct;
  R s2 = gmax(R(1) - c2, R(0));
  if (s2 <= R(9.99999993922529029e-09) || ct <= R(9.99999993922529029e-09)) return R(0);
  return R(0.5) * (R(-1) + std::sqrt(R(1) + alpha * alpha * (s2 / gmax(c2, R(9.99999993922529029e-09)))));
}
template <class R> R ggxG2(R alpha, R noV, R noL) { return R(1) / (R(1) + ggxLambda(alpha, noV) + ggxLambda(alpha, noL)); }
template <class R> V3<R> microfacetGGX(V3<R> f0, V3<R> f90, R rough, V3<R> n, V3<R> v, V3<R> l) {
  V3<R> h = normalize(v + l);
  R noV = gmax(dot(n, v), R(0)), noL = gmax(dot(n, l), R(0)), noH = gmax(dot(n, h), R(0)), voh = gmax(dot(v, h), R(0));
  if (noV <= R(0) || noL <= R(0) || noH <= R(0) || voh <= R(0)) return V3<R>(R(0));
  R a = gmax(rough * rough, R(0));
  return fresnelSchlick(f0, f90, voh) * ggxD(a, noH) * ggxG2(a, noV, noL) / gmax(R(4) * noV * noL, F1E5);
}
template <class R> V3<R> evalClearcoatGGX(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  if (c.clearcoat <= R(0)) return V3<R>(R(0));
  return c.clearcoat * microfacetGGX(V3<R>(R(0.03999999910593033)), V3<R>(R(1)), c.clearcoatRoughness, n, v, l);
}
template <class R> V3<R> evalClearcoat(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  return evalClearcoatGGX(c, n, v, l) * gmax(dot(n, l), R(0));
}
template <class R> R clearcoatBaseWeight(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  if (c.clearcoat <= R(0)) return R(1);
  const V3<R> f0(R(0.03999999910593033)), O(R(1));
  return gclamp(R(1) - c.clearcoat * gmax(max3(fresnelSchlick(f0, O, std::fabs(dot(n, v)))), max3(fresnelSchlick(f0, O, std::fabs(dot(n, l))))),
                R(0), R(1));
}
// KF:2045-2089 (KL:1540)
template <class R> V3<R> evalMixture(const BState<R>& s, V3<R> n, V3<R> ncc, V3<R> v, V3<R> l) {
  if (!(dot(n, v) * dot(n, l) > R(0)))
    return evalDiffuseTransmission(s.c, n, l) + evalDielectricTransmission(s.c, s.currentMediumIor, n, v, l);
  V3<R> base = evalReflectedBase(s, n, v, l);
  if (s.c.thinWalled < R(0.5) && s.c.transparency > R(0)) base = base + evalDielectricTransmission(s.c, s.currentMediumIor, n, v, l);
  return evalClearcoat(s.c, ncc, v, l) + base * clearcoatBaseWeight(s.c, ncc, v, l);
}

// ---- pdf, KF:2107-2615 (KL:1616-2120)
template <class R> R thinInterfaceBaseProb(const Closure<R>& c, V3<R> n, V3<R> v) {
  return gclamp(R(1) - max3(fresnelSchlick(thinSharedSpecularColor(c), thinSharedSpecularF90(c), gmax(std::fabs(dot(n, v)), F1E5))), R(0), R(1));
}
template <class R> R transmissiveBaseLayerWeight(const Closure<R>& c) {
  return gmax(c.diffuseWeight + c.sheenWeight + c.diffuseTransmissionWeight + c.transmissionWeight, F1E5);
}
template <class R> R pdfDiffuseTransmission(V3<R> n, V3<R> l) { return gmax(dot(-n, l), R(0)) * F1PI; }
template <class R> R specularLobePdf(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> d) {
  V3<R> h = normalize(v + d);
  V3<R> lh = projectAniso(n, c.anisotropyTangent, h);
  R noH = std::fabs(lh.z);
  R voh = gmax(dot(v, h), R(0));
  if (noH <= R(0) || voh <= R(0)) return R(0);
  return ggxDAniso(anisoAlpha(c.roughness, c.anisotropy), lh) * noH / gmax(R(4) * voh, F1E5);
}
template <class R> R thickTransmissionPdf(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l, R iorI, R iorO) {
  V3<R> lv = projectAniso(n, c.anisotropyTangent, v);
  R noV = std::fabs(lv.z);
  R noL = std::fabs(projectAniso(n, c.anisotropyTangent, l).z);
  if (noV <= F1E5 || noL <= F1E5 || dot(n, v) * dot(n, l) >= R(0)) return R(0);
  R eta = iorO / gmax(iorI, F1E5);
  V3<R> h = -normalize(v + l * eta);
  V3<R> lh = projectAniso(n, c.anisotropyTangent, h);
  R voh = dot(v, h), loh = dot(l, h);
  if (std::fabs(lh.z) <= F1E5 || voh * loh >= R(0)) return R(0);
  V2<R> a = anisoAlpha(c.roughness, c.anisotropy);
  return ggxDAniso(a, lh) * ggxG1Aniso(a, lv) *
         (eta * eta * std::fabs(voh) * std::fabs(loh) / gmax(opow(voh + eta * loh, R(2)) * noV, F1E5));
}
template <class R> R pdfDielectricTransmission(const Closure<R>& c, R curIor, V3<R> n, V3<R> v, V3<R> l) {
  bool inside; R iorI, iorO;
  mediumIors(c, curIor, inside, iorI, iorO);
  TransCoeffs<R> co = transCoeffs(c, curIor, n, v, iorI, iorO);
  if (dot(n, v) * dot(n, l) < R(0)) {
    R tp = gmax(co.tp, F1E5);
    R lobe = c.thinWalled > R(0.5) ? specularLobePdf(c, n, v, flipAcross(l, n)) : thickTransmissionPdf(c, n, v, l, iorI, iorO);
    return gmax(tp * lobe, F1E5);
  }
  return gmax(gclamp(co.rp, R(0), R(1)) * specularLobePdf(c, n, v, l), F1E5);
}
template <class R> R pdfDiffuseReflection(V3<R> n, V3<R> l) { return gmax(dot(n, l), R(0)) * F1PI; }
template <class R> R specularLobePdfIso(V3<R> n, V3<R> v, R rough, V3<R> d) {
  V3<R> h = normalize(v + d);
  R noH = gmax(dot(n, h), R(0)), voh = gmax(dot(v, h), R(0));
  if (noH <= R(0) || voh <= R(0)) return R(0);
  return ggxD(gmax(rough * rough, R(0)), noH) * noH / gmax(R(4) * voh, F1E5);
}
template <class R> R pdfClearcoat(const Closure<R>& c, V3<R> n, V3<R> v, V3<R> l) {
  return gmax(dot(n, l), R(0)) > R(0) ? specularLobePdfIso(n, v, c.clearcoatRoughness, l) : R(0);
}
template <class R> R dielectricViewFresnelMax(const Closure<R>& c, V3<R> n, V3<R> v) {
  return max3(fresnelSchlick(c.dielectricSpecularColor, c.dielectricSpecularF90, gmax(dot(n, v), R(0))));
}
template <class R> R pdfMixture(const BState<R>& s, V3<R> n, V3<R> ncc, V3<R> v, V3<R> d) {
  const Closure<R>& c = s.c;
  R nd = dot(n, d);
  R noL = gmax(nd, R(0));
  R ccL = gmax(dot(ncc, d), R(0));
  bool transSide = dot(n, v) * nd < R(0);
  bool thin = isThinTransmissiveDielectric(c);
  R thinBase = thin ? thinInterfaceBaseProb(c, n, v) : R(0);
  R thinSpec = thin ? R(1) - thinBase : R(0);
  if (transSide) {
    R pdf;
    if (thin) {
      R blw = transmissiveBaseLayerWeight(c);
      pdf = c.diffuseTransmissionWeight > R(0) ? thinBase * (c.diffuseTransmissionWeight / blw) * pdfDiffuseTransmission(n, d) : R(0);
      if (c.transmissionWeight > R(0))
        pdf = pdf + thinBase * (c.transmissionWeight / blw) * specularLobePdf(c, n, v, flipAcross(d, n));
      return pdf > R(0) ? gmax(pdf, F1E5) : R(0);
    }
    R tw = gmax(c.diffuseTransmissionWeight + c.transmissionWeight, F1E5);
    pdf = c.diffuseTransmissionWeight > R(0) ? c.diffuseTransmissionWeight / tw * pdfDiffuseTransmission(n, d) : R(0);
    if (c.transmissionWeight > R(0))
      pdf = pdf + c.transmissionWeight / tw * pdfDielectricTransmission(c, s.currentMediumIor, n, v, d);
    return pdf > R(0) ? gmax(pdf, F1E5) : R(0);
  }
  if (noL <= R(0)) return R(0);
  if (thin) {
    R blw = transmissiveBaseLayerWeight(c);
    R tot = gmax(thinBase + thinSpec + c.clearcoatWeight, F1E5);
    R diffusePdf = thinBase * c.diffuseWeight / blw / tot * pdfDiffuseReflection(n, d);
    R sheenPdf = thinBase * c.sheenWeight / blw / tot * pdfDiffuseReflection(n, d);
    R specPdf = thinSpec / tot * specularLobePdf(c, n, v, d);
    R ccw = c.clearcoatWeight / tot;
    R cc = ccL > R(0) ? pdfClearcoat(c, ncc, v, d) : R(0);
    return gmax(diffusePdf + sheenPdf + specPdf + ccw * cc, F1E5);
  }
  R dvf = dielectricViewFresnelMax(c, n, v);
  R omf = R(1) - dvf;
  R dpw = c.diffuseWeight * omf;
  R spw = c.sheenWeight * omf;
  R specw = c.specularWeight + c.dielectricWeight * (R(1) - c.transparency) * dvf;
  R tot = gmax(dpw + spw + c.diffuseTransmissionWeight + specw + c.transmissionWeight + c.clearcoatWeight, F1E5);
  R diffusePdf = dpw / tot * pdfDiffuseReflection(n, d);
  R sheenPdf = spw / tot * pdfDiffuseReflection(n, d);
  R specPdf = specw / tot * specularLobePdf(c, n, v, d);
  R transPdf = c.transmissionWeight > R(0) ? c.transmissionWeight / tot * pdfDielectricTransmission(c, s.currentMediumIor, n, v, d) : R(0);
  R ccw = c.clearcoatWeight / tot;
  R cc = ccL > R(0) ? pdfClearcoat(c, ncc, v, d) : R(0);
  return gmax(diffusePdf + sheenPdf + specPdf + transPdf + ccw * cc, F1E5);
}

// ---- sampling, KF:2631-3567 (KL:2134-2973)
template <class R> void orthonormalBasis(V3<R> n, V3<R>& t, V3<R>& b) {
  R sign = n.z >= R(0) ? R(1) : R(-1);
  R a = R(-1) / (sign + n.z);
  R bb = n.x * n.y * a;
  t = normalize(V3<R>(R(1) + sign * n.x * n.x * a, sign * bb, -sign * n.x));
  b = normalize(V3<R>(bb, sign + n.y * n.y * a, -n.y));
}
template <class R> V3<R> sampleCosineHemisphere(V3<R> n, R u0, R u1) {
  R r = std::sqrt(u0);
  R phi = R(6.28318548202514648) * u1;
  R x = r * ocos(phi), y = r * osin(phi);
  R z = std::sqrt(gmax(R(0), R(1) - u0));
  V3<R> t, b;
  orthonormalBasis(n, t, b);
  return normalize(t * x + b * y + n * z);
}
template <class R> BSample<R> sampleDiffuseReflection(const Closure<R>& c, R cur, V3<R> n, V3<R> v, R u0, R u1) {
  V3<R> d = sampleCosineHemisphere(n, u0, u1);
  R pdf = gmax(pdfDiffuseReflection(n, d), F1E5);
  return BSample<R>{d, pdf, evalDiffuseReflection(c, n, v, d) / pdf, R(0), R(0), cur};
}
template <class R> BSample<R> sampleSheen(const Closure<R>& c, R cur, V3<R> n, V3<R> v, R u0, R u1) {
  V3<R> d = sampleCosineHemisphere(n, u0, u1);
  R pdf = gmax(pdfDiffuseReflection(n, d), F1E5);
  return BSample<R>{d, pdf, evalSheen(c, n, v, d) / pdf, R(0), R(0), cur};
}
template <class R> BSample<R> sampleDiffuseTransmission(const Closure<R>& c, R cur, V3<R> n, V3<R> v, R u0, R u1) {
  V3<R> d = sampleCosineHemisphere(-n, u0, u1);
  R pdf = gmax(pdfDiffuseTransmission(n, d), F1E5);
  bool inside = (gmax(cur, R(1)) > R(1.00010001659393311)) ? (c.thinWalled < R(0.5)) : false;
  bool thick = c.thinWalled < R(0.5);
  BSample<R> s{d, pdf, evalDiffuseTransmission(c, n, d) / pdf, R(0), thick ? R(1) : R(0), R(0)};
  s.nextMediumIor = thick ? (inside ? R(1) : c.ior) : cur;
  return s;
}
template <class R> void anisotropicBasis(V3<R> n, V3<R> ti, V3<R>& t, V3<R>& b) {
  V3<R> s = normalize(ti - n * dot(ti, n));
  t = s;
  b = normalize(cross(n, s));
}
template <class R> V3<R> sampleGGXHalf(const Closure<R>& c, V3<R> n, R rough, R u0, R u1) {
  V2<R> a = anisoAlpha(rough, c.anisotropy);
  R phi = R(6.28318548202514648) * u1;
  R sp = osin(phi), cp = ocos(phi);
  R cosT = R(1) / std::sqrt(R(1) + u0 / gmax(R(1) - u0, R(9.99999997475242708e-07)) /
                                      gmax(cp * cp / gmax(a.x * a.x, R(9.999999960041972e-13)) + sp * sp / gmax(a.y * a.y, R(9.999999960041972e-13)),
                                           R(9.99999997475242708e-07)));
  R sinT = std::sqrt(gmax(R(0), R(1) - cosT * cosT));
  R x = sinT * cp, y = sinT * sp;
  V3<R> t, b;
  anisotropicBasis(n, c.anisotropyTangent, t, b);
  return normalize(t * x + b * y + n * cosT);
}
template <class R> R sampleHalfVectorPdf(const Closure<R>& c, V3<R> n, V3<R> h) {
  V3<R> lh = projectAniso(n, c.anisotropyTangent, h);
  R noH = std::fabs(lh.z);
  if (noH <= F1E5) return R(0);
  return ggxDAniso(anisoAlpha(c.roughness, c.anisotropy), lh) * noH;
}
template <class R> V3<R> safeRefract(V3<R> i, V3<R> n, R eta) {
  R noI = dot(n, i);
  R k = R(1) - eta * eta * (R(1) - noI * noI);
  if (k <= R(0)) return V3<R>(R(0));
  return eta * i - (eta * noI + std::sqrt(k)) * n;
}
template <class R> V3<R> dispersionChannelMask(R u) {
  if (u < R(0.3333333432674408)) return V3<R>(R(1), R(0), R(0));
  if (u < R(0.66666668653488159)) return V3<R>(R(0), R(1), R(0));
  return V3<R>(R(0), R(0), R(1));
}
// KF:2881-3145
template <class R>
BSample<R> sampleDielectricTransmission(const Closure<R>& c, R cur, V3<R> tn, V3<R> n, V3<R> v, R u0, R u1, R eventU) {
  bool inside; R iorI, iorO;
  mediumIors(c, cur, inside, iorI, iorO);
  if (c.thinWalled < R(0.5) && std::fabs(iorI - iorO) <= F1E4)
    return BSample<R>{-v, R(1), weightedTransmissionColor(c), R(2), R(1), iorO};
  V3<R> h = sampleGGXHalf(c, n, c.roughness, u0, u1);
  TransCoeffs<R> co = transCoeffs(c, cur, n, v, iorI, iorO);
  R halfPdf = sampleHalfVectorPdf(c, n, h);
  R probReflect = gclamp(co.rp, R(0), R(1));
  R tp = gmax(co.tp, F1E5);
  BSample<R> s;
  s.crossedBoundary = R(0);
  s.nextMediumIor = cur;
  if (eventU <= probReflect) {
    V3<R> d = reflect(-v, h);
    R lobe = halfPdf / gmax(R(4) * std::fabs(dot(d, h)), F1E5);
    s.dir = d;
    R p = gmax(probReflect * lobe, F1E5);
    s.pdf = p;
    s.bsdfOverPdf = evalDielectricSpecularGGX(c, n, v, d, iorI, iorO) / p;
    s.specular = R(1);
    return s;
  }
  if (c.thinWalled > R(0.5)) {
    V3<R> d = flipAcross(reflect(-v, h), n);
    R lobe = halfPdf / gmax(R(4) * std::fabs(dot(d, h)), F1E5);
    s.dir = d;
    R p = gmax(tp * lobe, F1E5);
    s.pdf = p;
    s.bsdfOverPdf = evalThinTransmission(c, n, v, d, iorI, iorO) / p;
    s.specular = R(1);
    return s;
  }
  bool disp = c.dispersion > F1E5;
  R channelU = gclamp((eventU - probReflect) / gmax(tp, F1E5), R(0), R(0.99999898672103882));
  V3<R> mask = disp ? dispersionChannelMask(channelU) : V3<R>(R(1));
  V3<R> ciors = dispersionIors(c);
  R tIor = disp ? dot(ciors, mask) : c.ior;
  R channelProb = disp ? R(0.3333333432674408) : R(1);
  R sI = inside ? tIor : R(1);
  R sO = inside ? R(1) : tIor;
  V3<R> mv = -v;
  V3<R> refr = safeRefract(mv, h, sI / gmax(sO, F1E5));
  if (length(refr) <= F1E5) {
    V3<R> d = reflect(mv, h);
    R p = gmax(pdfDielectricTransmission(c, cur, n, v, d), F1E5);
    s.dir = d;
    s.pdf = p;
    s.bsdfOverPdf = evalDielectricSpecularGGX(c, n, v, d, iorI, iorO) / p;
    s.specular = R(3);
    return s;
  }
  R voh = dot(v, h);
  R roh = dot(refr, h);
  R lobe = halfPdf * (sO * sO) * std::fabs(roh) / gmax(opow(sI * voh + sO * roh, R(2)), F1E5);
  R g1 = ggxG1Aniso(anisoAlpha(c.roughness, c.anisotropy), projectAniso(n, c.anisotropyTangent, refr));
  R avoh = std::fabs(voh);
  V3<R> rr = fresnelSchlickDielectric(avoh, c.dielectricSpecularColor, c.dielectricSpecularF90, sI, sO, false);
  V3<R> etc = (c.iridescence > F1E5 && iorI <= R(1.00010001659393311))
                  ? (V3<R>(R(1)) - thinFilmDielectricFresnel(c, avoh, rr, R(1))) * weightedTransmissionColor(c)
                  : co.tc;
  s.dir = refr;
  R ctp = tp * channelProb;
  s.pdf = gmax(ctp * lobe, F1E5);
  s.bsdfOverPdf = etc * mask / gmax(ctp, F1E5) * g1;
  s.specular = R(1);
  s.crossedBoundary = R(1);
  s.nextMediumIor = inside ? R(1) : c.ior;
  return s;
}
template <class R> BSample<R> sampleBaseSpecular(const Closure<R>& c, R cur, V3<R> n, V3<R> v, R u0, R u1) {
  V3<R> d = reflect(-v, sampleGGXHalf(c, n, c.roughness, u0, u1));
  R pdf = gmax(specularLobePdf(c, n, v, d), F1E5);
  return BSample<R>{d, pdf, evalBaseSpecular(c, n, v, d) / pdf, R(1), R(0), cur};
}
template <class R> V3<R> sampleGGXHalfIso(V3<R> n, R rough, R u0, R u1) {
  R a = gmax(rough * rough, R(0));
  R phi = R(6.28318548202514648) * u1;
  R cosT = std::sqrt((R(1) - u0) / gmax(R(1) + (a * a - R(1)) * u0, R(9.99999997475242708e-07)));
  R sinT = std::sqrt(gmax(R(0), R(1) - cosT * cosT));
  R x = sinT * ocos(phi), y = sinT * osin(phi);
  V3<R> t, b;
  orthonormalBasis(n, t, b);
  return normalize(t * x + b * y + n * cosT);
}
template <class R> BSample<R> sampleClearcoat(const Closure<R>& c, R cur, V3<R> n, V3<R> v, R u0, R u1) {
  V3<R> d = reflect(-v, sampleGGXHalfIso(n, c.clearcoatRoughness, u0, u1));
  R pdf = gmax(pdfClearcoat(c, n, v, d), F1E5);
  return BSample<R>{d, pdf, evalClearcoat(c, n, v, d) / pdf, R(1), R(0), cur};
}
// KF:3200-3567 (KL:2597-2973)
template <class R>
BSample<R> sampleMixture(const BState<R>& st, V3<R> rawNg, V3<R> tn, V3<R> n, V3<R> ncc, V3<R> v, R choiceU, R u0, R u1, R eventU) {
  const Closure<R>& c = st.c;
  const R cur = st.currentMediumIor;
  V3<R> dir;
  if (isThinTransmissiveDielectric(c)) {
    R tb = thinInterfaceBaseProb(c, n, v);
    R ts = R(1) - tb;
    R blw = transmissiveBaseLayerWeight(c);
    R tot = gmax(tb + ts + c.clearcoatWeight, F1E5);
    R pd = tb * (c.diffuseWeight / blw) / tot;
    R psh = tb * (c.sheenWeight / blw) / tot;
    R pdt = tb * (c.diffuseTransmissionWeight / blw) / tot;
    R ptt = tb * (c.transmissionWeight / blw) / tot;
    R psp = ts / tot;
    BSample<R> ts_;
    ts_.crossedBoundary = R(0);
    ts_.nextMediumIor = cur;
    if (choiceU < pd) {
      BSample<R> x = sampleDiffuseReflection(c, cur, n, v, u0, u1); ts_.specular = x.specular; dir = x.dir;
    } else if (choiceU < pd + psh) {
      BSample<R> x = sampleSheen(c, cur, n, v, u0, u1); ts_.specular = x.specular; dir = x.dir;
    } else if (choiceU < pd + psh + pdt) {
      BSample<R> x = sampleDiffuseTransmission(c, cur, n, v, u0, u1); ts_.specular = x.specular; dir = x.dir;
    } else if (choiceU < pd + psh + pdt + ptt) {
      ts_ = sampleDielectricTransmission(c, cur, tn, n, v, u0, u1, R(1));
      R mp = pdfMixture(st, n, ncc, v, ts_.dir);
      ts_.pdf = mp;
      ts_.bsdfOverPdf = evalDielectricTransmission(c, cur, n, v, ts_.dir) / gmax(mp, F1E5);
      return ts_;
    } else if (choiceU < pd + psh + pdt + ptt + psp) {
      BSample<R> x = sampleBaseSpecular(c, cur, n, v, u0, u1); ts_.specular = x.specular; dir = x.dir;
    } else {
      BSample<R> x = sampleClearcoat(c, cur, ncc, v, u0, u1); ts_.specular = x.specular; dir = x.dir;
    }
    R nd = dot(n, dir);
    R noL = gmax(nd, R(0));
    R ccL = gmax(dot(ncc, dir), R(0));
    if (noL <= R(0)) {
      if (nd < R(0)) {
        R mp = pdfMixture(st, n, ncc, v, dir);
        V3<R> dte = evalDiffuseTransmission(c, n, dir);
        ts_.dir = dir;
        R p = gmax(mp, F1E5);
        ts_.pdf = p;
        ts_.bsdfOverPdf = dte / gmax(p, F1E5);
        ts_.specular = R(0);
        return ts_;
      }
      V3<R> fb = sampleCosineHemisphere(n, u0, u1);
      ts_.dir = fb;
      ts_.pdf = gmax(gmax(dot(n, fb), R(0)) * F1PI, F1E5);
      ts_.bsdfOverPdf = V3<R>(R(0));
      ts_.specular = R(0);
      return ts_;
    }
    R mp = pdfMixture(st, n, ncc, v, dir);
    V3<R> rb = evalReflectedBase(st, n, v, dir) / gmax(noL, F1E5);
    V3<R> cce = ccL > R(0) ? evalClearcoat(c, ncc, v, dir) / gmax(ccL, F1E5) : V3<R>(R(0));
    R bw = clearcoatBaseWeight(c, ncc, v, dir);
    ts_.dir = dir;
    ts_.pdf = mp;
    ts_.bsdfOverPdf = (rb * bw + cce) * noL / gmax(mp, F1E5);
    return ts_;
  }

  R dvf = dielectricViewFresnelMax(c, n, v);
  R omf = R(1) - dvf;
  R dpw = c.diffuseWeight * omf;
  R spw = c.sheenWeight * omf;
  R specw = c.specularWeight + c.dielectricWeight * (R(1) - c.transparency) * dvf;
  R bpw = dpw + spw + c.diffuseTransmissionWeight;
  R tot = gmax(bpw + specw + c.transmissionWeight + c.clearcoatWeight, F1E5);
  R baseProb = bpw / tot;
  R specProb = specw / tot;
  R transProb = c.transmissionWeight / tot;
  BSample<R> s;
  s.crossedBoundary = R(0);
  s.nextMediumIor = cur;
  if (choiceU < baseProb) {
    R bu = choiceU / gmax(baseProb, F1E5);
    bool pos = bpw > R(0);
    R dp = pos ? dpw / bpw : R(0);
    R sp = pos ? spw / bpw : R(0);
    if (bu < dp) {
      BSample<R> x = sampleDiffuseReflection(c, cur, n, v, u0, u1); s.specular = x.specular; dir = x.dir;
    } else if (bu < dp + sp) {
      BSample<R> x = sampleSheen(c, cur, n, v, u0, u1); s.specular = x.specular; dir = x.dir;
    } else {
      BSample<R> x = sampleDiffuseTransmission(c, cur, n, v, u0, u1);
      s.specular = x.specular; s.crossedBoundary = x.crossedBoundary; s.nextMediumIor = x.nextMediumIor; dir = x.dir;
    }
  } else if (choiceU < baseProb + specProb) {
    BSample<R> x = sampleBaseSpecular(c, cur, n, v, u0, u1); s.specular = x.specular; dir = x.dir;
  } else if (choiceU < baseProb + specProb + transProb) {
    s = sampleDielectricTransmission(c, cur, tn, n, v, u0, u1, eventU);
    s.bsdfOverPdf = s.bsdfOverPdf / gmax(transProb, F1E5);
    if (s.pdf < R(1)) s.pdf = s.pdf * transProb;
    return s;
  } else {
    BSample<R> x = sampleClearcoat(c, cur, ncc, v, u0, u1); s.specular = x.specular; dir = x.dir;
  }
  R nd = dot(n, dir);
  R noL = gmax(nd, R(0));
  R ccL = gmax(dot(ncc, dir), R(0));
  R absNoL = std::fabs(nd);
  if (noL <= R(0)) {
    if (nd < R(0)) {
      R mp = pdfMixture(st, n, ncc, v, dir);
      V3<R> dte = evalDiffuseTransmission(c, n, dir);
      s.dir = dir;
      R p = gmax(mp, F1E5);
      s.pdf = p;
      s.bsdfOverPdf = dte / p;
      s.specular = R(0);
      return s;
    }
    V3<R> fb = sampleCosineHemisphere(n, u0, u1);
    s.dir = fb;
    s.pdf = gmax(gmax(dot(n, fb), R(0)) * F1PI, F1E5);
    s.bsdfOverPdf = V3<R>(R(0));
    s.specular = R(0);
    s.nextMediumIor = cur;
    return s;
  }
  R mp = pdfMixture(st, n, ncc, v, dir);
  V3<R> rb = evalReflectedBase(st, n, v, dir) / gmax(noL, F1E5);
  V3<R> cce = ccL > R(0) ? evalClearcoat(c, ncc, v, dir) / gmax(ccL, F1E5) : V3<R>(R(0));
  R bw = clearcoatBaseWeight(c, ncc, v, dir);
  s.dir = dir;
  s.pdf = mp;
  s.bsdfOverPdf = (rb * bw + cce) * absNoL / mp;
  s.nextMediumIor = cur;
  return s;
}

#undef F1E5
#undef F1E4
#undef FPI
#undef F1PI
}  // namespace orc
