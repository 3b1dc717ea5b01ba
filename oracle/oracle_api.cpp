// CPU ORACLE — test infrastructure only (see pt_oracle.hpp). C entry points
// for the Python tests (ctypes) and for bench.py's cpu_baseline leg.
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "integrator_oracle.hpp"
#include "post_oracle.hpp"

using namespace orc;

extern "C" {

// renderer.ts:777-838. Returns the marginal sum (totalSum, a JS double).
double oracle_ibl_cdf(const float* rgba, int w, int h, float* pdf, float* cdf, float* ypdf, float* ycdf) {
  return ibl_cdf(rgba, w, h, pdf, cdf, ypdf, ycdf);
}

// Generator probe: initial state of pixel (px, py) for `seed`, then n steps:
// raw states (x, y) after each step, the u32 outputs and rng_float values.
uint32_t oracle_rng_seq(int seed, int px, int py, int n, uint32_t* states, uint32_t* out_u32, float* out_f,
                        uint32_t* init_xy) {
  Rng s = rng_init(seed, px, py);
  if (init_xy) { init_xy[0] = s.x; init_xy[1] = s.y; }
  for (int i = 0; i < n; i++) {
    Rng t = s;
    uint32_t u = gm_rng(t);
    float f = rng_float(s);
    if (states) { states[2 * i] = s.x; states[2 * i + 1] = s.y; }
    if (out_u32) out_u32[i] = u;
    if (out_f) out_f[i] = f;
  }
  return s.x;
}

// Steps an explicit state (x, y) n times (for generator properties).
void oracle_rng_step(uint32_t x, uint32_t y, int n, uint32_t* states) {
  Rng s{x, y};
  for (int i = 0; i < n; i++) {
    gm_rng(s);
    states[2 * i] = s.x; states[2 * i + 1] = s.y;
  }
}

// Closest-hit traversal of rays given as {org.xyz, tfar, dir.xyz, pad}.
// counters: [node_visits, tri_tests] summed.
int oracle_trace(const Scene* s, const float* rays, int n, int32_t* tri, float* t, float* u, float* v, int64_t* counters,
                 int threads) {
  std::atomic<int64_t> nv{0}, nt{0};
  auto work = [&](int lo, int hi) {
    Counters c;
    for (int i = lo; i < hi; i++) {
      const float* r = rays + (size_t)i * 8;
      Ray<float> ray = make_ray(V3<float>(r[4], r[5], r[6]), V3<float>(r[0], r[1], r[2]), r[3]);
      Hit<float> h;
      intersect_nearest(*s, ray, h, &c);
      tri[i] = h.tri; t[i] = h.t; u[i] = h.u; v[i] = h.v;
    }
    nv += c.node_visits; nt += c.tri_tests;
  };
  threads = std::max(1, threads);
  std::vector<std::thread> th;
  int chunk = (n + threads - 1) / threads;
  for (int k = 0; k < threads; k++) {
    int lo = k * chunk, hi = std::min(n, lo + chunk);
    if (lo < hi) th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
  if (counters) { counters[0] = nv.load(); counters[1] = nt.load(); }
  return 0;
}

// Renders frames [first_frame, first_frame+nframes) of the pixel rectangle
// [x0,x1)x[y0,y1) of a W x H image. accum is RGBA f32 for the whole image
// (row 0 = bottom row, the GL convention) and is updated with the running
// mean of SH/renderer.frag:153-156. samples (optional): per frame, per rect
// pixel RGBA radiance; sigs / prim (optional): per-sample path signature and
// primary triangle id. counters: [node_visits, tri_tests, rays].
// use_double selects the FP64 instantiation.
int oracle_render(const Scene* s, const Frame* f0, int W, int H, int x0, int y0, int x1, int y1, int first_frame,
                  int nframes, float* accum, float* samples, uint32_t* sigs, int32_t* prim, int64_t* counters,
                  int threads, int use_double) {
  int rw = x1 - x0, rh = y1 - y0;
  if (rw <= 0 || rh <= 0) return 0;
  std::atomic<int64_t> nv{0}, nt{0}, nr{0};
  auto work = [&](int rlo, int rhi) {
    Counters c;
    for (int fi = 0; fi < nframes; fi++) {
      Frame f = *f0;
      f.frame_count = float(first_frame + fi);
      for (int yy = rlo; yy < rhi; yy++)
        for (int xx = 0; xx < rw; xx++) {
          int px = x0 + xx, py = y0 + yy;
          float L[4];
          uint32_t sg;
          int pt;
          if (use_double) {
            PathOut<double> o = trace_path<double>(*s, f, W, H, px, py, &c);
            L[0] = (float)o.L.x; L[1] = (float)o.L.y; L[2] = (float)o.L.z; L[3] = (float)o.alpha;
            sg = o.sig; pt = o.primary_tri;
          } else {
            PathOut<float> o = trace_path<float>(*s, f, W, H, px, py, &c);
            L[0] = o.L.x; L[1] = o.L.y; L[2] = o.L.z; L[3] = o.alpha;
            sg = o.sig; pt = o.primary_tri;
          }
          size_t ri = ((size_t)fi * rh + yy) * rw + xx;
          if (samples) for (int k = 0; k < 4; k++) samples[ri * 4 + k] = L[k];
          if (sigs) sigs[ri] = sg;
          if (prim) prim[ri] = pt;
          if (accum) {
            float* a = accum + ((size_t)py * W + px) * 4;
            float fc = f.frame_count;
            for (int k = 0; k < 4; k++) a[k] = (a[k] * (fc - 1.0f) + L[k]) / fc;
          }
        }
    }
    nv += c.node_visits; nt += c.tri_tests; nr += c.rays;
  };
  threads = std::max(1, threads);
  std::vector<std::thread> th;
  int chunk = (rh + threads - 1) / threads;
  for (int k = 0; k < threads; k++) {
    int lo = k * chunk, hi = std::min(rh, lo + chunk);
    if (lo < hi) th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
  if (counters) { counters[0] = nv.load(); counters[1] = nt.load(); counters[2] = nr.load(); }
  return 0;
}

// BSDF probes on one shading configuration (no textures, white vertex colour):
// mat = 68 floats, geo = {n.xyz, ng.xyz, t.xyz, t.w, wi.xyz}.
static MClosure<double> probe_closure_d(const float* mat, const float* geo) {
  Scene s; s.mats = mat; s.K = 1;
  RState<double> rs;
  rs.n = V3<double>(geo[0], geo[1], geo[2]);
  rs.ng = V3<double>(geo[3], geo[4], geo[5]);
  rs.tangent = V3<double>(geo[6], geo[7], geo[8]); rs.tangent_w = geo[9];
  rs.wi = V3<double>(geo[10], geo[11], geo[12]);
  return configure_material<double>(s, 0, rs, V3<double>(1.0), 1.0);
}
static MClosure<float> probe_closure_f(const float* mat, const float* geo) {
  Scene s; s.mats = mat; s.K = 1;
  RState<float> rs;
  rs.n = V3<float>(geo[0], geo[1], geo[2]);
  rs.ng = V3<float>(geo[3], geo[4], geo[5]);
  rs.tangent = V3<float>(geo[6], geo[7], geo[8]); rs.tangent_w = geo[9];
  rs.wi = V3<float>(geo[10], geo[11], geo[12]);
  return configure_material<float>(s, 0, rs, V3<float>(1.0f), 1.0f);
}

// out: eval.xyz (bsdf*|cos|, as pbrEvalGltfState), pdf; plus the flipped
// shading normal used (n.xyz) and backside flag.
int oracle_bsdf_eval(const float* mat, int profile, const float* geo, const float* wo, int use_double, double* out) {
  if (use_double) {
    MClosure<double> c = probe_closure_d(mat, geo);
    BState<double> st = make_state(c, profile);
    V3<double> l(wo[0], wo[1], wo[2]);
    V3<double> e = evalMixture(st, c.n, c.n, V3<double>(geo[10], geo[11], geo[12]), l);
    double p = pdfMixture(st, c.n, c.n, V3<double>(geo[10], geo[11], geo[12]), l);
    out[0] = e.x; out[1] = e.y; out[2] = e.z; out[3] = p;
    out[4] = c.n.x; out[5] = c.n.y; out[6] = c.n.z; out[7] = c.backside ? 1 : 0;
  } else {
    MClosure<float> c = probe_closure_f(mat, geo);
    BState<float> st = make_state(c, profile);
    V3<float> l(wo[0], wo[1], wo[2]);
    V3<float> e = evalMixture(st, c.n, c.n, V3<float>(geo[10], geo[11], geo[12]), l);
    float p = pdfMixture(st, c.n, c.n, V3<float>(geo[10], geo[11], geo[12]), l);
    out[0] = e.x; out[1] = e.y; out[2] = e.z; out[3] = p;
    out[4] = c.n.x; out[5] = c.n.y; out[6] = c.n.z; out[7] = c.backside ? 1 : 0;
  }
  return 0;
}

// rnd = {component, lobe.x, lobe.y, boundary}; out = {dir.xyz, pdf, w.xyz, specular}
int oracle_bsdf_sample(const float* mat, int profile, const float* geo, const float* rnd, int n, int use_double,
                       double* out) {
  for (int i = 0; i < n; i++) {
    const float* r = rnd + (size_t)i * 4;
    double* o = out + (size_t)i * 8;
    if (use_double) {
      MClosure<double> c = probe_closure_d(mat, geo);
      BState<double> st = make_state(c, profile);
      BSample<double> b = sampleMixture(st, c.ng, c.n, c.n, c.n, V3<double>(geo[10], geo[11], geo[12]), (double)r[0],
                                        (double)r[1], (double)r[2], (double)r[3]);
      o[0] = b.dir.x; o[1] = b.dir.y; o[2] = b.dir.z; o[3] = b.pdf;
      o[4] = b.bsdfOverPdf.x; o[5] = b.bsdfOverPdf.y; o[6] = b.bsdfOverPdf.z; o[7] = b.specular;
    } else {
      MClosure<float> c = probe_closure_f(mat, geo);
      BState<float> st = make_state(c, profile);
      BSample<float> b = sampleMixture(st, c.ng, c.n, c.n, c.n, V3<float>(geo[10], geo[11], geo[12]), r[0], r[1], r[2], r[3]);
      o[0] = b.dir.x; o[1] = b.dir.y; o[2] = b.dir.z; o[3] = b.pdf;
      o[4] = b.bsdfOverPdf.x; o[5] = b.bsdfOverPdf.y; o[6] = b.bsdfOverPdf.z; o[7] = b.specular;
    }
  }
  return 0;
}

// IBL sampling probes: dir + pdf for (r0, r1), and ibl_pdf / ibl_eval of a dir.
int oracle_ibl_sample(const Scene* s, const Frame* f, const float* r01, int n, float* out_dir_pdf) {
  for (int i = 0; i < n; i++) {
    float pdf;
    V3<float> d = ibl_sample_direction<float>(*s, *f, r01[2 * i], r01[2 * i + 1], pdf);
    out_dir_pdf[4 * i] = d.x; out_dir_pdf[4 * i + 1] = d.y; out_dir_pdf[4 * i + 2] = d.z; out_dir_pdf[4 * i + 3] = pdf;
  }
  return 0;
}
int oracle_ibl_lookup(const Scene* s, const Frame* f, const float* dirs, int n, float* out_rgb_pdf) {
  for (int i = 0; i < n; i++) {
    V3<float> d(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
    V3<float> e = ibl_eval<float>(*s, *f, d);
    float p = ibl_pdf<float>(*s, *f, d);
    out_rgb_pdf[4 * i] = e.x; out_rgb_pdf[4 * i + 1] = e.y; out_rgb_pdf[4 * i + 2] = e.z; out_rgb_pdf[4 * i + 3] = p;
  }
  return 0;
}

int oracle_scene_struct_size() { return (int)sizeof(Scene); }

// get_texture_value(id, uv) probes (f32): out f32[4n]
int oracle_tex_value(const Scene* s, float id, const float* uv, int n, float* out) {
  for (int i = 0; i < n; i++) {
    V4<float> t{1.0f, 1.0f, 1.0f, 1.0f};
    if (id >= 0.0f) t = tex_value<float>(*s, id, uv[2 * i], uv[2 * i + 1]);
    out[4 * i] = t.x; out[4 * i + 1] = t.y; out[4 * i + 2] = t.z; out[4 * i + 3] = t.w;
  }
  return 0;
}
int oracle_frame_struct_size() { return (int)sizeof(Frame); }

// Thin-film reflectance probes (KF:789-915): in f32[5n] = {outsideIor,
// filmIor, baseF0 (one value for all channels, or < 0 to use in[5n + 3i..]),
// thicknessNm, cosTheta1}; out f64[3n]. base3 (f32[3n]) holds per-channel
// base F0 for the rows whose baseF0 is negative (may be NULL otherwise).
int oracle_thin_film(const float* in, const float* base3, int n, int use_double, double* out) {
  for (int i = 0; i < n; i++) {
    const float* r = in + (size_t)5 * i;
    float b[3] = {r[2], r[2], r[2]};
    if (r[2] < 0.0f) {
      if (!base3) return -1;
      for (int k = 0; k < 3; k++) b[k] = base3[(size_t)3 * i + k];
    }
    if (use_double) {
      V3<double> f = thinFilmReflectance<double>(r[0], r[1], V3<double>(b[0], b[1], b[2]), r[3], r[4]);
      out[3 * i] = f.x; out[3 * i + 1] = f.y; out[3 * i + 2] = f.z;
    } else {
      V3<float> f = thinFilmReflectance<float>(r[0], r[1], V3<float>(b[0], b[1], b[2]), r[3], r[4]);
      out[3 * i] = f.x; out[3 * i + 1] = f.y; out[3 * i + 2] = f.z;
    }
  }
  return 0;
}

// Dispersion channel IORs (KF:1289-1298) of (ior, dispersion) pairs: out f64[3n].
int oracle_dispersion_iors(const float* ior_disp, int n, int use_double, double* out) {
  for (int i = 0; i < n; i++) {
    if (use_double) {
      Closure<double> c{};
      c.ior = ior_disp[2 * i]; c.dispersion = ior_disp[2 * i + 1];
      V3<double> f = dispersionIors(c);
      out[3 * i] = f.x; out[3 * i + 1] = f.y; out[3 * i + 2] = f.z;
    } else {
      Closure<float> c{};
      c.ior = ior_disp[2 * i]; c.dispersion = ior_disp[2 * i + 1];
      V3<float> f = dispersionIors(c);
      out[3 * i] = f.x; out[3 * i + 1] = f.y; out[3 * i + 2] = f.z;
    }
  }
  return 0;
}

// renderer.ts:665-700 over an ow x oh canvas: tonemap (+ gamma), optional
// FXAA, display resample, 8-bit conversion. rgba32 / rgba8 may be NULL.
void oracle_postprocess(const float* accum, int W, int H, int mode, float exposure, int gamma, int fxaa, int ow, int oh,
                        float* rgba32, uint8_t* rgba8) {
  orc_post::Post P{accum, W, H, mode, gamma, exposure};
  for (int y = 0; y < oh; y++)
    for (int x = 0; x < ow; x++) {
      float o[4];
      orc_post::display(P, fxaa, x, y, ow, oh, o);
      const size_t k = 4 * ((size_t)y * ow + x);
      for (int c = 0; c < 4; c++) {
        if (rgba32) rgba32[k + c] = o[c];
        if (rgba8) rgba8[k + c] = orc_post::unorm8(o[c]);
      }
    }
}

// Camera ray of pixel (px, py) with jitter (r0, r1), SH/renderer.frag:129-142:
// out = {org.xyz, dir.xyz}.
int oracle_camera_ray(const Frame* f, int W, int H, int px, int py, float r0, float r1, float* out) {
  Ray<float> r = camera_ray<float>(*f, W, H, px, py, r0, r1);
  out[0] = r.org.x; out[1] = r.org.y; out[2] = r.org.z;
  out[3] = r.dir.x; out[4] = r.dir.y; out[5] = r.dir.z;
  return 0;
}

// The analytic GGX albedo fits of the energy compensation (KF:5-91), in
// double: out = {approxAvgAlbedoGGX(alpha), approxDirAlbedoGGX(alpha, cosT)}.
int oracle_ggx_albedo_fits(double alpha, double cosT, double* out) {
  out[0] = approxAvgAlbedoGGX<double>(alpha);
  out[1] = approxDirAlbedoGGX<double>(alpha, cosT);
  return 0;
}

// The pinned float transcendentals (pin_math.hpp) over n arguments:
// fn 0 sin, 1 cos, 2 exp, 3 log, 4 pow(x, y), 5 atan2(x = y-argument, y = x-argument), 6 acos,
// 7 sin and 8 cos through psincos.
int oracle_pin_math(int fn, const float* x, const float* y, float* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    float s, c;
    switch (fn) {
      case 0: out[i] = pm::psin(x[i]); break;
      case 1: out[i] = pm::pcos(x[i]); break;
      case 2: out[i] = pm::pexp(x[i]); break;
      case 3: out[i] = pm::plog(x[i]); break;
      case 4: out[i] = pm::ppow(x[i], y[i]); break;
      case 5: out[i] = pm::patan2(x[i], y[i]); break;
      case 6: out[i] = pm::pacos(x[i]); break;
      case 7: pm::psincos(x[i], s, c); out[i] = s; break;
      case 8: pm::psincos(x[i], s, c); out[i] = c; break;
      default: return -1;
    }
  }
  return 0;
}

}  // extern "C"
