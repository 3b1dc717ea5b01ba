// CPU ORACLE — test infrastructure only, never part of the product path.
//
// Line-by-line C++ restatement of the reference's GLSL path tracer (the
// `misptdl` integrator of dspbr-pt 0.5.0) used as the correctness checker
// for the HIP engine. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.
//
// Templated on Real: Real=float with -ffp-contract=off is the parity target;
// Real=double is the converged-image reference.
//
// Citations are relative to the reference repository root:
//   SH/ = packages/lib/shader/
//   KF  = packages/lib/shader/generated/slang_materials/webgl-full/material_kernel.glsl
//   KL  = packages/lib/shader/generated/slang_materials/webgl-lean/material_kernel.glsl
// The lean profile (KL) is the full kernel (KF) with iridescence and
// dispersion forced to zero in the closure (KL:395,409,412); every KF-only
// code path is gated on those two closure fields, so one restatement with a
// profile flag covers both files.
//
// Pinned definitions (GLSL leaves them to the implementation; the HIP engine
// uses the same ones):
//   normalize(v) = v * (1/sqrt(dot(v,v)))       (correctly rounded sqrt, div)
//   min(x,y) = y<x ? y : x, max(x,y) = x<y ? y : x (GLSL spec, NaN-sensitive)
//   mix(x,y,a) = x*(1-a) + y*a                   (GLSL spec)
//   dot = (x*x'+y*y')+z*z', no fused multiply-add anywhere
//   v_uv = (2(px+0.5))/W - 1 (the interpolated full-screen-quad varying)
//   texture() bilinear on the IBL: GL LINEAR + REPEAT, weights in Real
//   sin, cos, exp, log, pow, atan2, acos in float: the pinned implementations
//     of dspbr-pt_amd/csrc/pin_math.hpp (double internally, rounded once),
//     the same code the kernels run; the double build uses the C library
//
// Parity status per function (how each is pinned) is listed in the header of
// tests/test_oracle.py.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../dspbr-pt_amd/csrc/pin_math.hpp"

namespace orc {

// Transcendentals: pinned float versions, C library double versions.
inline float osin(float x) { return pm::psin(x); }
inline float ocos(float x) { return pm::pcos(x); }
inline float oexp(float x) { return pm::pexp(x); }
inline float olog(float x) { return pm::plog(x); }
inline float opow(float x, float y) { return pm::ppow(x, y); }
inline float oatan2(float y, float x) { return pm::patan2(y, x); }
inline float oacos(float x) { return pm::pacos(x); }
inline double osin(double x) { return std::sin(x); }
inline double ocos(double x) { return std::cos(x); }
inline double oexp(double x) { return std::exp(x); }
inline double olog(double x) { return std::log(x); }
inline double opow(double x, double y) { return std::pow(x, y); }
inline double oatan2(double y, double x) { return std::atan2(y, x); }
inline double oacos(double x) { return std::acos(x); }

// ----------------------------------------------------------------- vectors
template <class R> struct V2 { R x, y; };
template <class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  explicit V3(R s) : x(s), y(s), z(s) {}
  R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> struct V4 { R x, y, z, w; };

template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return V3<R>(s * a.x, s * a.y, s * a.z); }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }

template <class R> inline R gmin(R x, R y) { return (y < x) ? y : x; }
template <class R> inline R gmax(R x, R y) { return (x < y) ? y : x; }
template <class R> inline R gclamp(R x, R lo, R hi) { return gmin(gmax(x, lo), hi); }
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return V3<R>(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return V3<R>(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
template <class R> inline V3<R> vclamp(V3<R> v, V3<R> lo, V3<R> hi) { return vmin(vmax(v, lo), hi); }
template <class R> inline R gmix(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <class R> inline V3<R> vmix(V3<R> x, V3<R> y, V3<R> a) {
  return V3<R>(gmix(x.x, y.x, a.x), gmix(x.y, y.y, a.y), gmix(x.z, y.z, a.z));
}
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R length(V3<R> v) { return std::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(V3<R> v) { return v * (R(1) / std::sqrt(dot(v, v))); }
// GLSL reflect(I, N) = I - 2*dot(N,I)*N
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - (R(2) * dot(n, i)) * n; }
template <class R> inline R max3(V3<R> v) { return gmax(v.x, gmax(v.y, v.z)); }
template <class R> inline V3<R> vabs(V3<R> v) { return V3<R>(std::fabs(v.x), std::fabs(v.y), std::fabs(v.z)); }
template <class R> inline V3<R> vsqrt(V3<R> v) { return V3<R>(std::sqrt(v.x), std::sqrt(v.y), std::sqrt(v.z)); }
template <class R> inline V3<R> vexp(V3<R> v) { return V3<R>(oexp(v.x), oexp(v.y), oexp(v.z)); }
template <class R> inline V3<R> vlog(V3<R> v) { return V3<R>(olog(v.x), olog(v.y), olog(v.z)); }
template <class R> inline V3<R> vcos(V3<R> v) { return V3<R>(ocos(v.x), ocos(v.y), ocos(v.z)); }
template <class R> inline V3<R> vpow(V3<R> v, R e) { return V3<R>(opow(v.x, e), opow(v.y, e), opow(v.z, e)); }
template <class R> inline R sat(R v) { return gclamp(v, R(0), R(1)); }

// SH/constants.glsl:1-21 (values as f32 literals)
template <class R> struct K {
  static constexpr R PI = R(3.14159265358979323f);
  static constexpr R TWO_PI = R(6.28318530717958648f);
  static constexpr R EPS = R(1e-8f);
  static constexpr R EPS_COS = R(0.001f);
  static constexpr R EPS_PDF = R(0.001f);
  static constexpr R MINIMUM_ROUGHNESS = R(0.0001f);
  static constexpr R TFAR_MAX = R(100000.0f);
  static constexpr R BVH_FAR = R(1e30f);
};
constexpr int RR_START_DEPTH = 2;           // SH/constants.glsl:20
constexpr float RR_TERMINATION_PROB = 0.1f; // SH/constants.glsl:21
constexpr int BVH_STACK_SIZE = 64;          // SH/bvh.glsl:8
constexpr int MAX_BVH_TRAVERSAL_STEPS = 4096;
constexpr int MAX_BVH_LEAF_TRIANGLES = 64;

// ----------------------------------------------------------------- RNG
// SH/rng.glsl:5-29
struct Rng { uint32_t x, y; };
inline uint32_t gm_rng(Rng& s) {
  s.x = 36969u * (s.x & 65535u) + (s.x >> 16u);
  s.y = 18000u * (s.y & 65535u) + (s.y >> 16u);
  return (s.x << 16u) + s.y;
}
inline float rng_float(Rng& s) { return float(gm_rng(s)) / float(0xFFFFFFFFu); }
// fragcoord = pixel centre (px+0.5, py+0.5), bottom-left origin.
inline Rng rng_init(int seed, int px, int py) {
  float off = float(seed * 17);
  float fx = float(px) + 0.5f, fy = float(py) + 0.5f;
  Rng s;
  s.x = (uint32_t)(397.6432f * (fx + off));
  s.y = (uint32_t)(397.6432f * (fy + 0.0f));
  s.x ^= (uint32_t)(32.9875f * (fy + off));
  s.y ^= (uint32_t)(32.9875f * (fx + 0.0f));
  return s;
}

// ----------------------------------------------------------------- scene
struct Scene {
  const float* verts = nullptr;    // f32[20*V]  packages/lib/scene_data.ts:21
  int V = 0;
  const uint32_t* tris = nullptr;  // u32[3*T]
  int T = 0;
  const float* nodes = nullptr;    // f32[16*N]  SH/bvh.glsl:1-6
  int N = 0;
  const int32_t* bvh_idx = nullptr;  // i32[M]
  int M = 0;
  const float* mats = nullptr;     // f32[68*K]  packages/lib/material.ts
  int K = 0;
  // IBL (RGBA f32) + importance-sampling tables from ibl_cdf()
  const float* env = nullptr;
  int ew = 0, eh = 0;
  const float* pdf = nullptr;
  const float* cdf = nullptr;
  const float* ypdf = nullptr;
  const float* ycdf = nullptr;
  int profile = 0;  // 0 = webgl-lean, 1 = webgl-full
  // material textures, scene_data_webgl2.ts:233-363: TexInfo f32[8*num_tex]
  // (material.ts:483-523), RGBA8 arrays of (w, h, layers) in texdims,
  // texels of all arrays concatenated, row 0 first
  const float* texinfo = nullptr;
  int num_tex = 0;
  const int32_t* texdims = nullptr;
  int num_texarr = 0;
  const uint8_t* texels = nullptr;
};

// 52-float uniform block, SH/renderer.frag:10-31 <-> renderer.ts:376-438
struct Frame {
  float view[16];
  float bg[4];
  float cam[4];     // xyz position, w = filmHeight
  float inv_res[2];
  float ibl_res[2];
  float frame_count, debug_mode, pad1, use_ibl;
  float ibl_rotation, bg_from_ibl, max_bounces, focal_length;
  float ibl_total_sum, ray_eps, clamp_threshold, pad0;
  float scene_counts[4];
  float light_pos[4];
  float light_emission[4];
};
static_assert(sizeof(Frame) == 52 * 4, "frame block is 52 floats");

struct Counters {
  int64_t node_visits = 0, tri_tests = 0, rays = 0;
};

// ----------------------------------------------------------------- BVH
// SH/bvh.glsl:12-27
template <class R> struct Ray { V3<R> dir, org; R tfar; V3<R> inv_dir; };
template <class R> struct Hit { int tri; R t; R u, v; };
template <class R> inline Ray<R> make_ray(V3<R> d, V3<R> o, R tfar) {
  Ray<R> r; r.dir = d; r.org = o; r.tfar = tfar;
  r.inv_dir = V3<R>(R(1) / d.x, R(1) / d.y, R(1) / d.z);
  return r;
}

// SH/bvh.glsl:31-41
template <class R> inline R intersect_bounds(V3<R> bmin, V3<R> bmax, V3<R> org, V3<R> inv, R tfar) {
  V3<R> t0 = (bmin - org) * inv;
  V3<R> t1 = (bmax - org) * inv;
  V3<R> lo = vmin(t0, t1), hi = vmax(t0, t1);
  R tmin = gmax(gmax(lo.x, lo.y), lo.z);
  R tmax = gmin(gmin(hi.x, hi.y), hi.z);
  if (std::isnan(tmin) || std::isnan(tmax) || tmin > tmax || tmax < R(0) || tmin >= tfar) return K<R>::BVH_FAR;
  return tmin;
}

// SH/bvh.glsl:44-62 (Moeller-Trumbore)
template <class R>
inline bool intersect_triangle(const Ray<R>& r, V3<R> p0, V3<R> p1, V3<R> p2, R tfar, R& t, R& u_out, R& v_out) {
  V3<R> e0 = p1 - p0, e1 = p2 - p0;
  V3<R> pvec = cross(r.dir, e1);
  R det = dot(e0, pvec);
  if (std::fabs(det) < K<R>::EPS) return false;
  R f = R(1) / det;
  V3<R> s = r.org - p0;
  R u = f * dot(s, pvec);
  if (u < R(0) || u > R(1)) return false;
  V3<R> q = cross(s, e0);
  R v = f * dot(r.dir, q);
  if (v < R(0) || u + v > R(1)) return false;
  R tt = f * dot(e1, q);
  if (tt < K<R>::EPS || tt >= tfar) return false;
  t = tt; u_out = u; v_out = v;
  return true;
}

template <class R> inline V3<R> vtx_attr(const Scene& s, uint32_t tri, int corner, int attr, R* w = nullptr) {
  uint32_t vid = s.tris[tri * 3 + corner];
  const float* p = s.verts + (size_t)vid * 20 + attr * 4;
  if (w) *w = R(p[3]);
  return V3<R>(R(p[0]), R(p[1]), R(p[2]));
}

// get_texture_value (scene_data_webgl2.ts:325-357): texture(array,
// vec3(uv * scale + offset, layer)) on RGBA8 UNORM with GL LINEAR + REPEAT:
// texel / 255, bilinear weights as in the environment lookup. The layer is
// round-to-nearest, clamped; an array index out of range yields vec4(1) (the
// generated if-chain falls through). uv set 1 is never written by the render
// state (renderer.frag:72-91): taken as (0, 0).
template <class R> inline V4<R> tex_value(const Scene& s, float id, R u0, R v0) {
  const int idx = (int)id;
  const float* ti = s.texinfo + (size_t)idx * 8;
  const int ai = (int)ti[2];
  if (ai < 0 || ai >= s.num_texarr) return V4<R>{R(1), R(1), R(1), R(1)};
  int64_t off = 0;
  for (int k = 0; k < ai; k++) off += (int64_t)s.texdims[3 * k] * s.texdims[3 * k + 1] * s.texdims[3 * k + 2];
  const int w = s.texdims[3 * ai], h = s.texdims[3 * ai + 1], layers = s.texdims[3 * ai + 2];
  const R u = ti[6] > 0.5f ? R(0) : u0, v = ti[6] > 0.5f ? R(0) : v0;
  const R tu = u * R(ti[4]) + R(ti[0]), tv = v * R(ti[5]) + R(ti[1]);
  R lf = std::floor(R(ti[3]) + R(0.5f));
  lf = gmin(gmax(lf, R(0)), R(layers - 1));
  const int layer = (int)lf;
  const R x = tu * R(w) - R(0.5f), y = tv * R(h) - R(0.5f);
  R fx = std::floor(x), fy = std::floor(y);
  const R a = x - fx, b = y - fy;
  fx = gmin(gmax(fx, R(-1.0e9f)), R(1.0e9f));
  fy = gmin(gmax(fy, R(-1.0e9f)), R(1.0e9f));
  const int ix = (int)fx, iy = (int)fy;
  int x0 = ix % w; if (x0 < 0) x0 += w;
  int x1 = (ix + 1) % w; if (x1 < 0) x1 += w;
  int y0 = iy % h; if (y0 < 0) y0 += h;
  int y1 = (iy + 1) % h; if (y1 < 0) y1 += h;
  const uint8_t* L = s.texels + (off + (int64_t)layer * w * h) * 4;
  auto texel = [&](int xx, int yy, int c) { return R(L[((int64_t)yy * w + xx) * 4 + c]) / R(255); };
  R o[4];
  for (int c = 0; c < 4; c++) {
    const R ia = R(1) - a, ib = R(1) - b;
    const R top = texel(x0, y0, c) * ia + texel(x1, y0, c) * a;
    const R bot = texel(x0, y1, c) * ia + texel(x1, y1, c) * a;
    o[c] = top * ib + bot * b;
  }
  return V4<R>{o[0], o[1], o[2], o[3]};
}

// SH/bvh.glsl:137-212. Occlusion uses the same closest-hit search
// (SH/bvh.glsl:225-229), as the reference does.
template <class R> inline bool intersect_nearest(const Scene& s, const Ray<R>& r, Hit<R>& hit, Counters* cnt) {
  hit.t = r.tfar; hit.tri = -1; hit.u = R(0); hit.v = R(0);
  int stack[BVH_STACK_SIZE];
  int top = 1;
  stack[0] = 0;
  int64_t nv = 0, nt = 0;
  auto valid = [&](int ni) { return ni >= 0 && ni < s.N; };
  auto push = [&](int child, int parent) {
    if (top < BVH_STACK_SIZE && valid(child) && child != parent) stack[top++] = child;
  };
  for (int step = 0; step < MAX_BVH_TRAVERSAL_STEPS && top > 0; step++) {
    int ni = stack[--top];
    if (!valid(ni)) continue;
    nv++;
    const float* nd = s.nodes + (size_t)ni * 16;
    int count = (int)nd[11];
    int first = (int)nd[15];
    if (count > 0) {
      int end = std::min(first + std::min(count, MAX_BVH_LEAF_TRIANGLES), s.M);
      for (int i = first; i < end; i++) {
        int idx = (i < 0 || i >= s.M) ? -1 : s.bvh_idx[i];
        if (idx < 0 || idx >= s.T) continue;
        nt++;
        V3<R> p0 = vtx_attr<R>(s, (uint32_t)idx, 0, 0);
        V3<R> p1 = vtx_attr<R>(s, (uint32_t)idx, 1, 0);
        V3<R> p2 = vtx_attr<R>(s, (uint32_t)idx, 2, 0);
        R t, u, v;
        if (intersect_triangle(r, p0, p1, p2, hit.t, t, u, v)) {
          hit.t = t; hit.tri = idx; hit.u = u; hit.v = v;
        }
      }
    } else {
      int left = (int)nd[3], right = (int)nd[7];
      V3<R> lmin{R(nd[0]), R(nd[1]), R(nd[2])}, lmax{R(nd[4]), R(nd[5]), R(nd[6])};
      V3<R> rmin{R(nd[8]), R(nd[9]), R(nd[10])}, rmax{R(nd[12]), R(nd[13]), R(nd[14])};
      R d0 = intersect_bounds(lmin, lmax, r.org, r.inv_dir, hit.t);
      R d1 = intersect_bounds(rmin, rmax, r.org, r.inv_dir, hit.t);
      const R F = K<R>::BVH_FAR;
      if (d0 < F && d1 < F) {
        if (d0 < d1) { push(right, ni); push(left, ni); }
        else { push(left, ni); push(right, ni); }
      } else if (d0 < F) {
        push(left, ni);
      } else if (d1 < F) {
        push(right, ni);
      }
    }
  }
  if (cnt) { cnt->node_visits += nv; cnt->tri_tests += nt; cnt->rays += 1; }
  return hit.tri >= 0;
}

// ----------------------------------------------------------------- utils
// SH/utils.glsl:104-109
template <class R> inline V3<R> onb_tangent(V3<R> n) {
  return normalize(cross(n, V3<R>(-n.z, n.x, -n.y)));
}
// SH/utils.glsl:88-98
template <class R> inline bool fix_normals(V3<R>& n, V3<R>& ng, V3<R> wi) {
  bool backside = false;
  if (dot(wi, ng) < R(0)) { ng = -ng; backside = true; }
  if (dot(ng, n) < R(0)) n = -n;
  return backside;
}
template <class R> inline V3<R> fix_normal(V3<R> n, V3<R> wi) { return dot(n, wi) < R(0) ? -n : n; }

// SH/utils.glsl:125-154
template <class R> inline R compute_theta(V3<R> d) { return oacos(gmax(R(-1), gmin(R(1), d.y))); }
template <class R> inline R compute_phi(V3<R> d) {
  R t = oatan2(d.z, d.x);
  return t < R(0) ? K<R>::TWO_PI + t : t;
}
template <class R> inline V2<R> dir_to_uv(V3<R> d, R& pdf) {
  R theta = compute_theta(d);
  R u = compute_phi(d) / K<R>::TWO_PI;
  R v = theta / K<R>::PI;
  pdf = R(1) / (R(2) * K<R>::PI * K<R>::PI * gmax(K<R>::EPS_COS, osin(theta)));
  return V2<R>{u, v};
}
template <class R> inline V3<R> uv_to_dir(V2<R> uv, R& pdf) {
  R theta = uv.y * K<R>::PI;
  R phi = uv.x * K<R>::TWO_PI;
  pdf = R(1) / (R(2) * K<R>::PI * K<R>::PI * gmax(K<R>::EPS_COS, osin(theta)));
  return V3<R>(osin(theta) * ocos(phi), ocos(theta), osin(theta) * osin(phi));
}

// ----------------------------------------------------------------- IBL
// renderer.ts:777-800, precompute1DPdfAndCdf. `sum` is a JS double; pdf and
// cdf are Float32Array stores; each cdf add is double math on two f32 values
// rounded to f32 (== an f32 add).
inline double cdf_1d(const float* data, float* pdf, float* cdf, int num) {
  double sum = 0;
  for (int i = 0; i < num; i++) sum += (double)data[i];
  if (sum == 0) sum = 1;
  for (int i = 0; i < num; i++) pdf[i] = (float)((double)data[i] / sum);
  cdf[0] = pdf[0];
  for (int i = 1; i < num; i++) cdf[i] = (float)((double)cdf[i - 1] + (double)pdf[i]);
  cdf[num - 1] = 1.0f;
  return sum;
}
// renderer.ts:802-838, precomputeIBLImportanceSamplingData
inline double ibl_cdf(const float* rgba, int w, int h, float* pdf, float* cdf, float* ypdf, float* ycdf) {
  std::vector<float> f((size_t)w * h), sumx(h);
  for (size_t i = 0; i < (size_t)w * h; i++)
    f[i] = (float)((double)rgba[i * 4] * 0.299 + (double)rgba[i * 4 + 1] * 0.587 + (double)rgba[i * 4 + 2] * 0.114);
  for (int y = 0; y < h; y++)
    sumx[y] = (float)cdf_1d(&f[(size_t)y * w], pdf + (size_t)y * w, cdf + (size_t)y * w, w);
  return cdf_1d(sumx.data(), ypdf, ycdf, h);
}

// SH/utils.glsl:196-217
inline int lower_bound(const float* row, int size, float value) {
  int count = size, first = 0;
  while (count > 0) {
    int step = count / 2;
    int idx = first + step;
    float v = row[idx];
    if (v < value) { first = ++idx; count -= step + 1; }
    else count = step;
  }
  return first;
}

// SH/lighting.glsl:66-74 ibl_eval texture(): GL LINEAR filtering, REPEAT wrap.
template <class R> inline V3<R> env_bilinear(const Scene& s, V2<R> uv) {
  int w = s.ew, h = s.eh;
  R x = uv.x * R(w) - R(0.5), y = uv.y * R(h) - R(0.5);
  R fx = std::floor(x), fy = std::floor(y);
  R a = x - fx, b = y - fy;
  long ix = (long)fx, iy = (long)fy;
  auto wrap = [](long i, int n) { long m = i % n; return (int)(m < 0 ? m + n : m); };
  int x0 = wrap(ix, w), x1 = wrap(ix + 1, w), y0 = wrap(iy, h), y1 = wrap(iy + 1, h);
  auto tex = [&](int xx, int yy) {
    const float* p = s.env + ((size_t)yy * w + xx) * 4;
    return V3<R>(R(p[0]), R(p[1]), R(p[2]));
  };
  V3<R> top = tex(x0, y0) * (R(1) - a) + tex(x1, y0) * a;
  V3<R> bot = tex(x0, y1) * (R(1) - a) + tex(x1, y1) * a;
  return top * (R(1) - b) + bot * b;
}

// texture() of a one-channel R32F importance table (the pdf and cdf textures
// of renderer.ts:858-859, created with LINEAR filtering and REPEAT wrap by
// gl_utils.ts:203-216): the red channel of the bilinear fetch.
template <class R> inline R table_bilinear(const float* t, int w, int h, V2<R> uv) {
  R x = uv.x * R(w) - R(0.5), y = uv.y * R(h) - R(0.5);
  R fx = std::floor(x), fy = std::floor(y);
  R a = x - fx, b = y - fy;
  long ix = (long)fx, iy = (long)fy;
  auto wrap = [](long i, int n) { long m = i % n; return (int)(m < 0 ? m + n : m); };
  int x0 = wrap(ix, w), x1 = wrap(ix + 1, w), y0 = wrap(iy, h), y1 = wrap(iy + 1, h);
  auto tex = [&](int xx, int yy) { return R(t[(size_t)yy * w + xx]); };
  R top = tex(x0, y0) * (R(1) - a) + tex(x1, y0) * a;
  R bot = tex(x0, y1) * (R(1) - a) + tex(x1, y1) * a;
  return top * (R(1) - b) + bot * b;
}

// SH/lighting.glsl:58-65, mat3 column-major times vector
template <class R> inline V3<R> rotate_dir_phi(V3<R> d, R rot, bool inverse) {
  R a = inverse ? -rot : rot;
  R c = ocos(a), sn = osin(a);
  // columns (c,0,s) (0,1,0) (-s,0,c)
  return V3<R>(c * d.x + R(0) * d.y + (-sn) * d.z,
               R(0) * d.x + R(1) * d.y + R(0) * d.z,
               sn * d.x + R(0) * d.y + c * d.z);
}
// Without an uploaded IBL the reference samples an empty texture: radiance
// and pdf are 0.
template <class R> inline V3<R> ibl_eval(const Scene& s, const Frame& f, V3<R> d) {
  if (!s.env || s.ew <= 0 || s.eh <= 0) return V3<R>(R(0));
  R pdf;
  V3<R> sd = rotate_dir_phi(d, R(f.ibl_rotation), false);
  V2<R> uv = dir_to_uv(sd, pdf);
  return env_bilinear(s, uv);
}
// SH/lighting.glsl:26-44, 76-86
template <class R> inline V3<R> ibl_sample_direction(const Scene& s, const Frame& f, float r0, float r1, R& o_pdf) {
  int w = (int)f.ibl_res[0], h = (int)f.ibl_res[1];
  if (!s.env || w <= 0 || h <= 0) { o_pdf = R(0); return V3<R>(R(0), R(1), R(0)); }
  // sampleRow1D reads the pdf at the unclamped index (0 past the end), then
  // clamps the index (SH/lighting.glsl:26-31).
  int y = lower_bound(s.ycdf, h, r1);
  float pdfY = y < h ? s.ypdf[y] : 0.0f;
  y = (y >= h) ? h - 1 : y;
  int x = lower_bound(s.cdf + (size_t)y * w, w, r0);
  float pdfX = x < w ? s.pdf[(size_t)y * w + x] : 0.0f;
  x = (x >= w) ? w - 1 : x;
  R sample_pdf = R(f.ibl_res[0]) * R(f.ibl_res[1]) * R(pdfY) * R(pdfX);
  V2<R> uv{R(x) / R(f.ibl_res[0]), R(y) / R(f.ibl_res[1])};
  R pdf_w;
  V3<R> d = rotate_dir_phi(uv_to_dir(uv, pdf_w), R(f.ibl_rotation), true);
  o_pdf = sample_pdf * pdf_w;
  return d;
}
// SH/lighting.glsl:89-98
template <class R> inline R ibl_pdf(const Scene& s, const Frame& f, V3<R> d) {
  if (!s.env || s.ew <= 0 || s.eh <= 0 || f.ibl_res[0] <= 0.0f || f.ibl_res[1] <= 0.0f) return R(0);
  R pdf_w;
  V3<R> sd = rotate_dir_phi(d, R(f.ibl_rotation), false);
  V2<R> uv = dir_to_uv(sd, pdf_w);
  R w = R(f.ibl_res[0]), h = R(f.ibl_res[1]);
  R x = gmin(uv.x * w, w - R(1));
  R y = gmin(uv.y * h, h - R(1));
  int ix = (int)x, iy = (int)y;
  return w * h * R(s.ypdf[iy]) * R(s.pdf[(size_t)iy * s.ew + ix]) * pdf_w;
}

}  // namespace orc
