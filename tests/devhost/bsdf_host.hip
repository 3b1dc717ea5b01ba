// CPU build of the engine's shading code (test infrastructure).
//
// Compiles dev_shade.hpp / dev_bsdf.hpp for the host so tests can compare the
// engine's material set-up and BSDF mixture with the oracle on the same CPU
// and libm: any difference left is a formula difference, not a
// transcendental-ulp difference between the host libm and the GPU math library.
#include "../../dspbr-pt_amd/csrc/dev_shade.hpp"

static void probe(const float* mat, int profile, const float* geo, Shading& sh, bsdf::Closure& cl, bsdf::Frame3& fr, f3& wi) {
  SceneView S{};
  S.mats = reinterpret_cast<const float4*>(mat);
  S.K = 1;
  Hitpoint hp{};
  hp.n = f3{geo[0], geo[1], geo[2]};
  hp.ng_raw = f3{geo[3], geo[4], geo[5]};
  hp.t = f3{geo[6], geo[7], geo[8]};
  hp.tw = geo[9];
  hp.vc = f3{1.0f, 1.0f, 1.0f};
  hp.vcw = 1.0f;
  hp.mat = 0;
  wi = f3{geo[10], geo[11], geo[12]};
  sh = configure(S, hp, wi);
  cl = bsdf::build_closure(sh.in, sh.tanA, profile == 1);
  fr = bsdf::aniso_frame(sh.n, cl.tanA);
}

extern "C" int devhost_bsdf_eval(const float* mat, int profile, const float* geo, const float* wo, double* out) {
  Shading sh;
  bsdf::Closure cl;
  bsdf::Frame3 fr;
  f3 wi;
  probe(mat, profile, geo, sh, cl, fr, wi);
  f3 l{wo[0], wo[1], wo[2]};
  f3 e = bsdf::eval(cl, fr, sh.n, wi, l);
  float p = bsdf::pdf(cl, fr, sh.n, wi, l);
  out[0] = e.x; out[1] = e.y; out[2] = e.z; out[3] = p;
  out[4] = sh.n.x; out[5] = sh.n.y; out[6] = sh.n.z; out[7] = sh.backside ? 1 : 0;
  return 0;
}

extern "C" int devhost_bsdf_sample(const float* mat, int profile, const float* geo, const float* rnd, int n, double* out) {
  Shading sh;
  bsdf::Closure cl;
  bsdf::Frame3 fr;
  f3 wi;
  probe(mat, profile, geo, sh, cl, fr, wi);
  for (int i = 0; i < n; i++) {
    const float* r = rnd + 4 * i;
    bsdf::Sample s = bsdf::sample(cl, fr, sh.n, wi, r[0], r[1], r[2], r[3]);
    double* o = out + 8 * i;
    o[0] = s.dir.x; o[1] = s.dir.y; o[2] = s.dir.z; o[3] = s.pdf;
    o[4] = s.w.x; o[5] = s.w.y; o[6] = s.w.z; o[7] = s.specular;
  }
  return 0;
}

// The refill's item order (dev_order.hpp) over a whole task: item i of a
// task of nframes frames over a rw x rh region -> frame slot and pixel
// offset (dx, dy).
#include "../../dspbr-pt_amd/csrc/dev_order.hpp"
extern "C" int devhost_task_items(int rw, int rh, int nframes, int* fslot, int* dx, int* dy) {
  const int npix = rw * rh;
  for (int64_t i = 0; i < (int64_t)npix * nframes; i++) {
    int f, r, x, y;
    item_frame_pixel(i, npix, nframes, f, r);
    region_tile_pixel(r, rw, rh, x, y);
    fslot[i] = f; dx[i] = x; dy[i] = y;
  }
  return 0;
}
