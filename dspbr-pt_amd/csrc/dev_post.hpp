// Display chain of renderer.ts:665-700 as one fused pass per canvas pixel:
// the display resample (display.frag) picks a texel of the display buffer,
// FXAA (fxaa.frag) — when on — fetches up to nine texels of the tonemapped
// buffer around it, and every fetched texel is tonemapped on the fly from the
// accumulation (tonemap.frag), so no intermediate buffer is written.
//
// GL behaviour pinned where the reference leaves it to the implementation:
// every render buffer is RGBA32F with NEAREST + CLAMP_TO_EDGE
// (gl_utils.ts:22-33), so a fetch at normalised u reads texel
// clamp(floor(u * W), 0, W - 1); fragment centres are (x + 0.5, y + 0.5) and
// the full-screen varying is v_uv = (2(x+0.5))/W - 1; `mediump` is f32; the
// canvas stores floor(clamp(c, 0, 1) * 255 + 0.5) (NaN -> 0); constant
// expressions fold in f32.
#pragma once
#include "dev_math.hpp"

namespace post {

// renderer.ts:192: ["None", "Reinhard", "Cineon", "AcesFilm"] -> 0..3; the
// shader's mode 4 (Uncharted2, tonemap.frag:89-90) is not in that list, and
// an unknown name gives indexOf() == -1: no operator at all.
DI float sat(float a) { return gclamp(a, 0.0f, 1.0f); }

DI float u2_helper(float x) {  // tonemap.frag:30-31
  const float n = x * (0.15f * x + 0.10f * 0.50f) + 0.20f * 0.02f;
  const float d = x * (0.15f * x + 0.50f) + 0.20f * 0.30f;
  return gmax(n / d - 0.02f / 0.30f, 0.0f);
}

DI float cineon(float x) {  // tonemap.frag:39-44, per channel after exposure
  x = gmax(0.0f, x - 0.004f);
  return pm::ppow((x * (6.2f * x + 0.5f)) / (x * (6.2f * x + 1.7f) + 0.06f), 2.2f);
}

DI float rrt_odt(float v) {  // tonemap.frag:46-51
  const float a = v * (v + 0.0245786f) - 0.000090537f;
  const float b = v * (0.983729f * v + 0.4329510f) + 0.238081f;
  return a / b;
}

DI f3 aces(f3 c) {  // tonemap.frag:53-71; the mat3 constructors list columns
  const f3 a{rrt_odt(0.59719f * c.x + 0.35458f * c.y + 0.04823f * c.z),
             rrt_odt(0.07600f * c.x + 0.90834f * c.y + 0.01566f * c.z),
             rrt_odt(0.02840f * c.x + 0.13383f * c.y + 0.83777f * c.z)};
  return f3{sat(1.60475f * a.x + -0.53108f * a.y + -0.07367f * a.z), sat(-0.10208f * a.x + 1.10813f * a.y + -0.00605f * a.z),
            sat(-0.00327f * a.x + -0.07276f * a.y + 1.07602f * a.z)};
}

struct Params {
  int mode, gamma;
  float exposure;
};

// tonemap.frag:73-93
DI f3 tonemap(f3 c, const Params& P) {
  const float e = P.exposure;
  if (P.mode == 0) c = c * e;
  if (P.mode == 1) {
    c = c * e;
    c = f3{sat(c.x / (1.0f + c.x)), sat(c.y / (1.0f + c.y)), sat(c.z / (1.0f + c.z))};
  }
  if (P.mode == 2) {
    c = c * e;
    c = f3{cineon(c.x), cineon(c.y), cineon(c.z)};
  }
  if (P.mode == 3) c = aces(c * (e / 0.6f));
  if (P.mode == 4) {
    c = c * e;
    const float w = u2_helper(1.0f);
    c = f3{sat(u2_helper(c.x) / w), sat(u2_helper(c.y) / w), sat(u2_helper(c.z) / w)};
  }
  if (P.gamma) {
    const float g = 1.0f / 2.2f;
    c = f3{pm::ppow(c.x, g), pm::ppow(c.y, g), pm::ppow(c.z, g)};
  }
  return c;
}

DI int nearest(float u, int n) {
  const int i = (int)floorf(u * (float)n);
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}
DI float vuv(int p, int n) { return (2.0f * (float(p) + 0.5f)) / float(n) - 1.0f; }
DI float luma(f3 c) { return (c.x * 0.299f + c.y * 0.587f) + c.z * 0.114f; }

struct Source {
  const float4* __restrict__ accum;  // RGBA32F, row 0 = bottom
  int W, H;
  Params P;
  // texel (i, j) of the tonemapped buffer = fragment (i, j) of the tonemap
  // pass, which reads the accumulation at its own uv
  DD float4 texel(int i, int j) const {
    const int ti = nearest((vuv(i, W) + 1.0f) * 0.5f, W), tj = nearest((vuv(j, H) + 1.0f) * 0.5f, H);
    const float4 a = accum[(size_t)tj * W + ti];
    const f3 c = tonemap(f3{a.x, a.y, a.z}, P);
    return make_float4(c.x, c.y, c.z, a.w);
  }
  DD f3 fetch(float u, float v) const { return xyz(texel(nearest(u, W), nearest(v, H))); }
};

// fxaa.frag:22-59 at fragment (x, y); ix, iy = u_inv_res
DD float4 fxaa(const Source& S, int x, int y, float ix, float iy) {
  const float fx = float(x) + 0.5f, fy = float(y) + 0.5f;
  const f3 nw = S.fetch((fx + -1.0f) * ix, (fy + -1.0f) * iy);
  const f3 ne = S.fetch((fx + 1.0f) * ix, (fy + -1.0f) * iy);
  const f3 sw = S.fetch((fx + -1.0f) * ix, (fy + 1.0f) * iy);
  const f3 se = S.fetch((fx + 1.0f) * ix, (fy + 1.0f) * iy);
  const float bx = fx * ix, by = fy * iy;
  const float4 m = S.texel(nearest(bx, S.W), nearest(by, S.H));
  const float lNW = luma(nw), lNE = luma(ne), lSW = luma(sw), lSE = luma(se), lM = luma(xyz(m));
  const float lmin = gmin(lM, gmin(gmin(lNW, lNE), gmin(lSW, lSE)));
  const float lmax = gmax(lM, gmax(gmax(lNW, lNE), gmax(lSW, lSE)));
  float dx = -((lNW + lNE) - (lSW + lSE));
  float dy = (lNW + lSW) - (lNE + lSE);
  const float reduce = gmax((((lNW + lNE) + lSW) + lSE) * (0.25f * (1.0f / 8.0f)), 1.0f / 128.0f);
  const float rcp = 1.0f / (gmin(fabsf(dx), fabsf(dy)) + reduce);
  dx = gmin(8.0f, gmax(-8.0f, dx * rcp)) * ix;
  dy = gmin(8.0f, gmax(-8.0f, dy * rcp)) * iy;
  const float k1 = 1.0f / 3.0f - 0.5f, k2 = 2.0f / 3.0f - 0.5f;
  const f3 rgbA = (S.fetch(bx + dx * k1, by + dy * k1) + S.fetch(bx + dx * k2, by + dy * k2)) * 0.5f;
  const f3 rgbB = rgbA * 0.5f + (S.fetch(bx + dx * -0.5f, by + dy * -0.5f) + S.fetch(bx + dx * 0.5f, by + dy * 0.5f)) * 0.25f;
  const float lB = luma(rgbB);
  const f3 c = (lB < lmin || lB > lmax) ? rgbA : rgbB;
  return make_float4(c.x, c.y, c.z, m.w);
}

DI unsigned unorm8(float c) {
  if (!(c > 0.0f)) return 0u;
  if (c >= 1.0f) return 255u;
  return (unsigned)floorf(c * 255.0f + 0.5f);
}

}  // namespace post
