// Host runtime and C ABI (include/dspt.h) of the gfx950 path-tracing engine.
// Kernels are in dspt_kernels.hip (same translation unit).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dspt.h"
#include "dspt_kernels.hip"

int dspt_bvh_build_impl(const float* pos, int64_t nv, const uint32_t* ind, int64_t nt, int max_leaf,
                        std::vector<float>& out_nodes, std::vector<int32_t>& out_idx, std::string& err);

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(DSPT_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kNumCounters = 32;

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
};

}  // namespace

struct dspt_ctx {
  int device = 0;
  int profile = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> scene_allocs, ibl_allocs, state_allocs, frame_allocs;
  size_t bytes_scene = 0, bytes_ibl = 0, bytes_state = 0, bytes_frame = 0;
  // scene
  float4* nodes = nullptr;
  float4* tri = nullptr;
  float4* verts = nullptr;
  uint32_t* tris = nullptr;
  float4* mats = nullptr;
  int32_t N = 0, M = 0, T = 0, K = 0;
  int64_t V = 0;
  bool have_scene = false;
  // ibl
  float4* env = nullptr;
  float *pdf = nullptr, *cdf = nullptr, *ypdf = nullptr, *ycdf = nullptr;
  int ew = 0, eh = 0;
  bool have_ibl = false;
  // frame
  FrameParams F{};
  bool have_frame = false;
  int W = 0, H = 0;
  float4* accum = nullptr;
  // path state
  int64_t cap = 0;
  PathState ps{};
  int32_t *qa = nullptr, *qb = nullptr, *shq = nullptr, *counts = nullptr;
  int* spill = nullptr;
  int trace_blocks = 0;
  unsigned long long* cnt = nullptr;
  int64_t max_paths = 8 << 20;
  // sample recording
  bool record = false;
  float4* rec_L = nullptr;
  uint32_t* rec_sig = nullptr;
  int32_t* rec_prim = nullptr;
  int64_t rec_n = 0;
  int last_frames = 0;
  // kernel timing
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  double kstat[DSPT_K_COUNT][2] = {};
};

namespace {

template <class T>
int dalloc(dspt_ctx* c, std::vector<void*>& list, size_t& bytes, T** out, size_t n) {
  void* p = nullptr;
  size_t sz = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) return fail(DSPT_ERR_MEMORY, std::string("hipMalloc(") + std::to_string(sz) + "): " + hipGetErrorString(e));
  list.push_back(p);
  bytes += sz;
  *out = (T*)p;
  return 0;
}
void free_list(std::vector<void*>& list, size_t& bytes) {
  for (void* p : list) (void)hipFree(p);
  list.clear();
  bytes = 0;
}

SceneView view_of(dspt_ctx* c) {
  SceneView S;
  S.nodes = c->nodes; S.tri = c->tri; S.verts = c->verts; S.tris = c->tris; S.mats = c->mats;
  S.N = c->N; S.M = c->M; S.T = c->T; S.K = c->K;
  S.env = c->env; S.pdf = c->pdf; S.cdf = c->cdf; S.ypdf = c->ypdf; S.ycdf = c->ycdf;
  S.ew = c->ew; S.eh = c->eh;
  S.has_ibl = c->have_ibl ? 1 : 0;
  S.full_profile = c->profile == DSPT_PROFILE_FULL;
  return S;
}

int grid_for(int64_t n, int block) { return (int)std::max<int64_t>(1, (n + block - 1) / block); }

hipEvent_t take_event(dspt_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Launch wrapper: with timing on, brackets the launch with events on the
// context stream (the stream the kernel runs on).
template <class F>
void timed(dspt_ctx* c, int kclass, F&& launch) {
  if (!c->timing) { launch(); return; }
  hipEvent_t a = take_event(c), b = take_event(c);
  (void)hipEventRecord(a, c->stream);
  launch();
  (void)hipEventRecord(b, c->stream);
  c->ev_pending.push_back({kclass, {a, b}});
}

void drain_events(dspt_ctx* c) {
  if (c->ev_pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& pe : c->ev_pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      c->kstat[pe.first][0] += 1.0;
      c->kstat[pe.first][1] += ms;
    }
    c->ev_pool.push_back(pe.second.first);
    c->ev_pool.push_back(pe.second.second);
  }
  c->ev_pending.clear();
}

// Path state for `cap` paths in flight.
int ensure_state(dspt_ctx* c, int64_t cap) {
  if (c->cap >= cap && ((c->ps.sig != nullptr) == c->record)) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  free_list(c->state_allocs, c->bytes_state);
  c->cap = 0;
  int r = 0;
  PathState& ps = c->ps;
  ps = PathState{};
  auto& L = c->state_allocs;
  auto& B = c->bytes_state;
  if ((r = dalloc(c, L, B, &ps.org, cap)) || (r = dalloc(c, L, B, &ps.dir, cap)) || (r = dalloc(c, L, B, &ps.hit, cap)) ||
      (r = dalloc(c, L, B, &ps.pw, cap)) || (r = dalloc(c, L, B, &ps.L, cap)) || (r = dalloc(c, L, B, &ps.rng, cap)) ||
      (r = dalloc(c, L, B, &ps.sh_org, cap)) || (r = dalloc(c, L, B, &ps.sh_dir, cap)) || (r = dalloc(c, L, B, &ps.sh_c, cap)) ||
      (r = dalloc(c, L, B, &ps.sh_pw, cap)) || (r = dalloc(c, L, B, &ps.sh_pdir, cap)) || (r = dalloc(c, L, B, &ps.sh_pc, cap)) ||
      (r = dalloc(c, L, B, &c->qa, cap)) || (r = dalloc(c, L, B, &c->qb, cap)) || (r = dalloc(c, L, B, &c->shq, cap)))
    return r;
  if (c->record) {
    if ((r = dalloc(c, L, B, &ps.sig, cap)) || (r = dalloc(c, L, B, &ps.prim, cap))) return r;
  }
  // traversal grid: enough resident waves to fill 256 CUs, capped by paths
  int max_blocks = 2048;
  c->trace_blocks = (int)std::min<int64_t>(max_blocks, grid_for(cap, kTraceBlock));
  if ((r = dalloc(c, L, B, &c->spill, (size_t)c->trace_blocks * kTraceBlock * (kStackSize - kLdsStack)))) return r;
  c->cap = cap;
  return 0;
}

int validate_nodes(const float* nodes, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    const float* d = nodes + i * 16;
    for (int k = 0; k < 4; k++) {
      float v = d[k * 4 + 3];
      if (!std::isfinite(v) || std::fabs(v) >= 1073741824.0f)
        return fail(DSPT_ERR_ARG, "bvh node " + std::to_string(i) + ": index/count field is not a finite value below 2^30");
    }
  }
  return 0;
}

}  // namespace

extern "C" {

const char* dspt_last_error(void) { return g_err.c_str(); }
const char* dspt_version(void) { return "dspt 0.1.0 gfx950"; }

int dspt_create(int profile, int device, dspt_ctx** out) {
  if (!out) return fail(DSPT_ERR_ARG, "out is NULL");
  if (profile != DSPT_PROFILE_LEAN && profile != DSPT_PROFILE_FULL) return fail(DSPT_ERR_ARG, "profile must be 0 (lean) or 1 (full)");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return fail(DSPT_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(DSPT_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  dspt_ctx* c = new dspt_ctx();
  c->device = device;
  c->profile = profile;
  if (const char* mp = std::getenv("DSPT_MAX_PATHS")) c->max_paths = std::max<int64_t>(65536, std::atoll(mp));
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(DSPT_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  void* p = nullptr;
  e = hipMalloc(&p, kNumCounters * sizeof(unsigned long long));
  if (e != hipSuccess) { (void)hipStreamDestroy(c->stream); delete c; return fail(DSPT_ERR_MEMORY, "counter allocation"); }
  c->cnt = (unsigned long long*)p;
  (void)hipMemset(c->cnt, 0, kNumCounters * sizeof(unsigned long long));
  e = hipMalloc(&p, 8 * sizeof(int32_t));
  if (e != hipSuccess) { (void)hipFree(c->cnt); (void)hipStreamDestroy(c->stream); delete c; return fail(DSPT_ERR_MEMORY, "count allocation"); }
  c->counts = (int32_t*)p;
  *out = c;
  return 0;
}

int dspt_destroy(dspt_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_list(c->scene_allocs, c->bytes_scene);
  free_list(c->ibl_allocs, c->bytes_ibl);
  free_list(c->state_allocs, c->bytes_state);
  free_list(c->frame_allocs, c->bytes_frame);
  if (c->rec_L) (void)hipFree(c->rec_L);
  if (c->rec_sig) (void)hipFree(c->rec_sig);
  if (c->rec_prim) (void)hipFree(c->rec_prim);
  drain_events(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  (void)hipFree(c->cnt);
  (void)hipFree(c->counts);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int dspt_upload_scene(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* nodes,
                      int64_t N, const int32_t* bvh_idx, int64_t M, const float* mats, int32_t K) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (!vertices || !tri_idx || !nodes || !bvh_idx || !mats) return fail(DSPT_ERR_ARG, "NULL scene buffer");
  if (V <= 0 || T <= 0 || N <= 0 || M <= 0 || K <= 0) return fail(DSPT_ERR_ARG, "empty scene");
  if (T >= (1LL << 24) || N >= (1LL << 30) || M >= (1LL << 30)) return fail(DSPT_ERR_UNSUPPORTED, "scene exceeds float-encoded index range");
  for (int64_t i = 0; i < 3 * T; i++)
    if ((int64_t)tri_idx[i] >= V) return fail(DSPT_ERR_ARG, "triangle index " + std::to_string(i) + " out of range");
  for (int64_t v = 0; v < V; v++) {
    float mi = vertices[v * 20 + 3];
    if (!(mi >= 0.0f) || mi >= (float)K) return fail(DSPT_ERR_ARG, "vertex " + std::to_string(v) + ": material index out of range");
  }
  for (int32_t k = 0; k < K; k++)
    for (int i = 44; i < 60; i++)
      if (mats[(size_t)k * 68 + i] >= 0.0f) return fail(DSPT_ERR_UNSUPPORTED, "material " + std::to_string(k) + " uses a texture");
  if (int r = validate_nodes(nodes, N)) return r;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_list(c->scene_allocs, c->bytes_scene);
  c->have_scene = false;
  int r = 0;
  auto& L = c->scene_allocs;
  auto& B = c->bytes_scene;
  if ((r = dalloc(c, L, B, &c->nodes, (size_t)N * 4)) || (r = dalloc(c, L, B, &c->tri, (size_t)M * 3)) ||
      (r = dalloc(c, L, B, &c->verts, (size_t)V * 5)) || (r = dalloc(c, L, B, &c->tris, (size_t)T * 3)) ||
      (r = dalloc(c, L, B, &c->mats, (size_t)K * 17)))
    return r;
  int32_t* d_bvh = nullptr;
  HIPCHK(hipMalloc(&d_bvh, (size_t)M * sizeof(int32_t)));
  HIPCHK(hipMemcpyAsync(c->nodes, nodes, (size_t)N * 64, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->verts, vertices, (size_t)V * 80, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->tris, tri_idx, (size_t)T * 12, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->mats, mats, (size_t)K * 272, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(d_bvh, bvh_idx, (size_t)M * 4, hipMemcpyHostToDevice, c->stream));
  k_prep_tris<<<grid_for(M, 256), 256, 0, c->stream>>>(d_bvh, c->tris, c->verts, (int32_t)M, (int32_t)T, c->tri);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipFree(d_bvh));
  c->N = (int32_t)N; c->M = (int32_t)M; c->T = (int32_t)T; c->K = K; c->V = V;
  c->have_scene = true;
  return 0;
}

int dspt_upload_ibl(dspt_ctx* c, const float* rgba, int32_t w, int32_t h, double* total_sum) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (!rgba || w <= 0 || h <= 0) return fail(DSPT_ERR_ARG, "invalid IBL image");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_list(c->ibl_allocs, c->bytes_ibl);
  c->have_ibl = false;
  size_t n = (size_t)w * h;
  int r = 0;
  auto& L = c->ibl_allocs;
  auto& B = c->bytes_ibl;
  float* lum = nullptr;
  float* rowsum = nullptr;
  double* total = nullptr;
  if ((r = dalloc(c, L, B, &c->env, n)) || (r = dalloc(c, L, B, &c->pdf, n)) || (r = dalloc(c, L, B, &c->cdf, n)) ||
      (r = dalloc(c, L, B, &c->ypdf, (size_t)h)) || (r = dalloc(c, L, B, &c->ycdf, (size_t)h)))
    return r;
  HIPCHK(hipMalloc(&lum, n * 4));
  HIPCHK(hipMalloc(&rowsum, (size_t)h * 4));
  HIPCHK(hipMalloc(&total, 8));
  HIPCHK(hipMemcpyAsync(c->env, rgba, n * 16, hipMemcpyHostToDevice, c->stream));
  k_ibl_luminance<<<grid_for((int64_t)n, 256), 256, 0, c->stream>>>(c->env, (int64_t)n, lum);
  k_ibl_rows<<<grid_for(h, 64), 64, 0, c->stream>>>(lum, w, h, c->pdf, c->cdf, rowsum);
  k_ibl_marginal<<<1, 64, 0, c->stream>>>(rowsum, h, c->ypdf, c->ycdf, total);
  HIPCHK(hipGetLastError());
  double tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, total, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(lum);
  (void)hipFree(rowsum);
  (void)hipFree(total);
  if (total_sum) *total_sum = tot;
  c->ew = w; c->eh = h;
  c->have_ibl = true;
  return 0;
}

int dspt_read_ibl_tables(dspt_ctx* c, float* pdf, float* cdf, float* ypdf, float* ycdf) {
  if (!c || !c->have_ibl) return fail(DSPT_ERR_STATE, "no IBL uploaded");
  size_t n = (size_t)c->ew * c->eh;
  if (pdf) HIPCHK(hipMemcpy(pdf, c->pdf, n * 4, hipMemcpyDeviceToHost));
  if (cdf) HIPCHK(hipMemcpy(cdf, c->cdf, n * 4, hipMemcpyDeviceToHost));
  if (ypdf) HIPCHK(hipMemcpy(ypdf, c->ypdf, (size_t)c->eh * 4, hipMemcpyDeviceToHost));
  if (ycdf) HIPCHK(hipMemcpy(ycdf, c->ycdf, (size_t)c->eh * 4, hipMemcpyDeviceToHost));
  return 0;
}

int dspt_resize(dspt_ctx* c, int32_t w, int32_t h) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (w <= 0 || h <= 0 || (int64_t)w * h > (1LL << 28)) return fail(DSPT_ERR_ARG, "invalid resolution");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  free_list(c->frame_allocs, c->bytes_frame);
  c->accum = nullptr;
  int r = dalloc(c, c->frame_allocs, c->bytes_frame, &c->accum, (size_t)w * h);
  if (r) return r;
  c->W = w; c->H = h;
  k_fill_f4<<<grid_for((int64_t)w * h, 256), 256, 0, c->stream>>>(c->accum, (int64_t)w * h, make_float4(0, 0, 0, 0));
  HIPCHK(hipGetLastError());
  return 0;
}

int dspt_set_frame(dspt_ctx* c, const float* block) {
  if (!c || !block) return fail(DSPT_ERR_ARG, "NULL argument");
  static_assert(sizeof(FrameParams) == DSPT_FRAME_FLOATS * 4, "frame block layout");
  std::memcpy(&c->F, block, sizeof(FrameParams));
  if (!(c->F.max_bounces >= 0.0f) || c->F.max_bounces > 1024.0f) return fail(DSPT_ERR_ARG, "max_bounces out of range");
  c->have_frame = true;
  return 0;
}

int dspt_set_record_samples(dspt_ctx* c, int enable) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  c->record = enable != 0;
  return 0;
}

int dspt_render(dspt_ctx* c, int32_t first_frame, int32_t num_frames) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (!c->have_scene) return fail(DSPT_ERR_STATE, "render before dspt_upload_scene");
  if (!c->have_frame) return fail(DSPT_ERR_STATE, "render before dspt_set_frame");
  if (!c->accum) return fail(DSPT_ERR_STATE, "render before dspt_resize");
  if (first_frame < 1 || num_frames < 0) return fail(DSPT_ERR_ARG, "frames are 1-based and num_frames >= 0");
  bool needs_ibl = c->F.use_ibl > 0.0f || c->F.bg_from_ibl > 0.0f;
  if (needs_ibl && !c->have_ibl) return fail(DSPT_ERR_STATE, "IBL enabled but no IBL uploaded");
  if (needs_ibl && ((int)c->F.ibl_res[0] != c->ew || (int)c->F.ibl_res[1] != c->eh))
    return fail(DSPT_ERR_ARG, "u_ibl_resolution does not match the uploaded IBL");
  HIPCHK(hipSetDevice(c->device));
  const int W = c->W, H = c->H;
  const int64_t npix = (int64_t)W * H;
  if (first_frame == 1)
    k_fill_f4<<<grid_for(npix, 256), 256, 0, c->stream>>>(c->accum, npix, make_float4(0, 0, 0, 0));
  if (num_frames == 0) return 0;
  int64_t slots = std::max<int64_t>(1, std::min<int64_t>(num_frames, c->max_paths / npix));
  int64_t cap = slots * npix;
  if (int r = ensure_state(c, cap)) return r;
  if (c->record) {
    int64_t need = (int64_t)num_frames * npix;
    if (c->rec_n < need) {
      if (c->rec_L) (void)hipFree(c->rec_L);
      if (c->rec_sig) (void)hipFree(c->rec_sig);
      if (c->rec_prim) (void)hipFree(c->rec_prim);
      c->rec_L = nullptr; c->rec_sig = nullptr; c->rec_prim = nullptr; c->rec_n = 0;
      HIPCHK(hipMalloc(&c->rec_L, need * sizeof(float4)));
      HIPCHK(hipMalloc(&c->rec_sig, need * sizeof(uint32_t)));
      HIPCHK(hipMalloc(&c->rec_prim, need * sizeof(int32_t)));
      c->rec_n = need;
    }
  }
  SceneView S = view_of(c);
  FrameParams F = c->F;
  const int maxB = (int)F.max_bounces;
  const int shade_blocks = (int)std::min<int64_t>(grid_for(cap, 256), 8192);
  for (int64_t f0 = 0; f0 < num_frames; f0 += slots) {
    int ns = (int)std::min<int64_t>(slots, num_frames - f0);
    int64_t P = (int64_t)ns * npix;
    int ff = first_frame + (int)f0;
    F.frame_count = (float)ff;
    timed(c, DSPT_K_RAYGEN, [&] {
      k_raygen<<<grid_for(P, 256), 256, 0, c->stream>>>(c->ps, F, W, H, (int)npix, ff, P, c->counts, c->cnt);
    });
    int32_t *qin = nullptr, *qout = c->qa;
    int32_t *cin = c->counts + 0, *cout = c->counts + 1, *csh = c->counts + 2;
    for (int it = 0; it <= maxB; it++) {
      timed(c, DSPT_K_TRACE_EXT, [&] {
        k_trace_ext<<<c->trace_blocks, kTraceBlock, 0, c->stream>>>(S, c->ps, qin, cin, c->spill, c->cnt);
      });
      k_fill_i32<<<1, 64, 0, c->stream>>>(cout, 1, 0);
      k_fill_i32<<<1, 64, 0, c->stream>>>(csh, 1, 0);
      timed(c, DSPT_K_SHADE, [&] {
        k_shade<<<shade_blocks, 256, 0, c->stream>>>(S, F, c->ps, it, (int)npix, qin, cin, qout, cout, c->shq, csh, c->cnt);
      });
      if (it < maxB)
        timed(c, DSPT_K_TRACE_SHADOW, [&] {
          k_trace_shadow<<<c->trace_blocks, kTraceBlock, 0, c->stream>>>(S, c->ps, c->shq, csh, c->spill, c->cnt);
        });
      // next iteration consumes qout
      qin = qout;
      qout = (qout == c->qa) ? c->qb : c->qa;
      std::swap(cin, cout);
    }
    HIPCHK(hipGetLastError());
    timed(c, DSPT_K_ACCUM, [&] {
      k_accumulate<<<grid_for(npix, 256), 256, 0, c->stream>>>(c->ps.L, c->accum, (int)npix, ns, ff);
    });
    if (c->record) {
      HIPCHK(hipMemcpyAsync(c->rec_L + f0 * npix, c->ps.L, P * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->rec_sig + f0 * npix, c->ps.sig, P * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(c->rec_prim + f0 * npix, c->ps.prim, P * sizeof(int32_t), hipMemcpyDeviceToDevice, c->stream));
    }
  }
  HIPCHK(hipGetLastError());
  c->last_frames = num_frames;
  return 0;
}

int dspt_synchronize(dspt_ctx* c) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_read_accum(dspt_ctx* c, float* rgba) {
  if (!c || !rgba) return fail(DSPT_ERR_ARG, "NULL argument");
  if (!c->accum) return fail(DSPT_ERR_STATE, "no render target (dspt_resize)");
  HIPCHK(hipMemcpyAsync(rgba, c->accum, (size_t)c->W * c->H * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_read_samples(dspt_ctx* c, float* rgba, uint32_t* sigs, int32_t* prim) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (!c->record || !c->rec_L) return fail(DSPT_ERR_STATE, "sample recording was not enabled for the last render");
  size_t n = (size_t)c->last_frames * c->W * c->H;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rgba) HIPCHK(hipMemcpy(rgba, c->rec_L, n * 16, hipMemcpyDeviceToHost));
  if (sigs) HIPCHK(hipMemcpy(sigs, c->rec_sig, n * 4, hipMemcpyDeviceToHost));
  if (prim) HIPCHK(hipMemcpy(prim, c->rec_prim, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int dspt_get_counters(dspt_ctx* c, int64_t* out) {
  if (!c || !out) return fail(DSPT_ERR_ARG, "NULL argument");
  unsigned long long h[kNumCounters];
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  for (int i = 0; i < DSPT_CNT_COUNT; i++) out[i] = (int64_t)h[i];
  out[DSPT_CNT_NODE_VISITS] = (int64_t)(h[4] + h[10]);
  out[DSPT_CNT_TRI_TESTS] = (int64_t)(h[5] + h[11]);
  // every traced extension-kernel ray that is not an extension ray is primary
  out[DSPT_CNT_RAYS_PRIMARY] = (int64_t)h[9] - (int64_t)h[2];
  return 0;
}

int dspt_reset_counters(dspt_ctx* c) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  drain_events(c);
  std::memset(c->kstat, 0, sizeof(c->kstat));
  HIPCHK(hipMemsetAsync(c->cnt, 0, kNumCounters * sizeof(unsigned long long), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_set_kernel_timing(dspt_ctx* c, int enable) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  drain_events(c);
  c->timing = enable != 0;
  return 0;
}

int dspt_get_kernel_stats(dspt_ctx* c, double* out) {
  if (!c || !out) return fail(DSPT_ERR_ARG, "NULL argument");
  drain_events(c);
  unsigned long long h[kNumCounters];
  HIPCHK(hipMemcpy(h, c->cnt, sizeof(h), hipMemcpyDeviceToHost));
  for (int k = 0; k < DSPT_K_COUNT; k++) {
    out[k * 4 + 0] = c->kstat[k][0];
    out[k * 4 + 1] = c->kstat[k][1];
    out[k * 4 + 2] = 0.0;
    out[k * 4 + 3] = 0.0;
  }
  out[DSPT_K_TRACE_EXT * 4 + 2] = (double)h[4];
  out[DSPT_K_TRACE_EXT * 4 + 3] = (double)h[5];
  out[DSPT_K_TRACE_SHADOW * 4 + 2] = (double)h[10];
  out[DSPT_K_TRACE_SHADOW * 4 + 3] = (double)h[11];
  return 0;
}

static int trace_common(dspt_ctx* c, const float* rays, int64_t n, int any, int32_t* tri, float* t, float* u, float* v) {
  if (!c || !rays || !tri) return fail(DSPT_ERR_ARG, "NULL argument");
  if (!c->have_scene) return fail(DSPT_ERR_STATE, "trace before dspt_upload_scene");
  if (n <= 0) return 0;
  HIPCHK(hipSetDevice(c->device));
  float4* d_rays = nullptr;
  int32_t* d_tri = nullptr;
  float *d_t = nullptr, *d_u = nullptr, *d_v = nullptr;
  int* d_spill = nullptr;
  int blocks = (int)std::min<int64_t>(2048, grid_for(n, kTraceBlock));
  HIPCHK(hipMalloc(&d_rays, n * 32));
  HIPCHK(hipMalloc(&d_tri, n * 4));
  if (t) {
    HIPCHK(hipMalloc(&d_t, n * 4));
    HIPCHK(hipMalloc(&d_u, n * 4));
    HIPCHK(hipMalloc(&d_v, n * 4));
  }
  HIPCHK(hipMalloc(&d_spill, (size_t)blocks * kTraceBlock * (kStackSize - kLdsStack) * 4));
  HIPCHK(hipMemcpyAsync(d_rays, rays, n * 32, hipMemcpyHostToDevice, c->stream));
  k_trace_rays<<<blocks, kTraceBlock, 0, c->stream>>>(view_of(c), d_rays, n, any, d_tri, d_t, d_u, d_v, d_spill, c->cnt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tri, d_tri, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (t) {
    HIPCHK(hipMemcpyAsync(t, d_t, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(u, d_u, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(v, d_v, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  (void)hipFree(d_rays); (void)hipFree(d_tri); (void)hipFree(d_spill);
  if (d_t) { (void)hipFree(d_t); (void)hipFree(d_u); (void)hipFree(d_v); }
  return 0;
}

int dspt_trace_closest(dspt_ctx* c, const float* rays, int64_t n, int32_t* tri, float* t, float* u, float* v) {
  if (!t || !u || !v) return fail(DSPT_ERR_ARG, "NULL output");
  return trace_common(c, rays, n, 0, tri, t, u, v);
}

int dspt_trace_occluded(dspt_ctx* c, const float* rays, int64_t n, int32_t* occluded) {
  if (!occluded) return fail(DSPT_ERR_ARG, "NULL output");
  std::vector<int32_t> tri(std::max<int64_t>(n, 1));
  int r = trace_common(c, rays, n, 1, tri.data(), nullptr, nullptr, nullptr);
  if (r) return r;
  for (int64_t i = 0; i < n; i++) occluded[i] = tri[i] >= 0 ? 1 : 0;
  return 0;
}

int dspt_build_bvh(const float* positions, int64_t nv, const uint32_t* indices, int64_t nt, int32_t max_leaf, float* nodes,
                   int64_t* num_nodes, int32_t* bvh_indices, int64_t* num_bvh_indices) {
  if (!positions || !indices || !num_nodes || !num_bvh_indices) return fail(DSPT_ERR_ARG, "NULL argument");
  static thread_local std::vector<float> cache_nodes;
  static thread_local std::vector<int32_t> cache_idx;
  static thread_local const float* cache_pos = nullptr;
  static thread_local int64_t cache_nt = -1;
  if (!nodes) {
    std::string err;
    if (dspt_bvh_build_impl(positions, nv, indices, nt, max_leaf, cache_nodes, cache_idx, err)) return fail(DSPT_ERR_ARG, err);
    cache_pos = positions; cache_nt = nt;
    *num_nodes = (int64_t)cache_nodes.size() / 16;
    *num_bvh_indices = (int64_t)cache_idx.size();
    return 0;
  }
  if (cache_pos != positions || cache_nt != nt) return fail(DSPT_ERR_STATE, "dspt_build_bvh: call with nodes == NULL first");
  if (*num_nodes < (int64_t)cache_nodes.size() / 16 || *num_bvh_indices < (int64_t)cache_idx.size())
    return fail(DSPT_ERR_ARG, "dspt_build_bvh: output buffers too small");
  std::memcpy(nodes, cache_nodes.data(), cache_nodes.size() * 4);
  std::memcpy(bvh_indices, cache_idx.data(), cache_idx.size() * 4);
  *num_nodes = (int64_t)cache_nodes.size() / 16;
  *num_bvh_indices = (int64_t)cache_idx.size();
  cache_nodes.clear(); cache_nodes.shrink_to_fit();
  cache_idx.clear(); cache_idx.shrink_to_fit();
  cache_pos = nullptr; cache_nt = -1;
  return 0;
}

int dspt_memory_usage(dspt_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return fail(DSPT_ERR_ARG, "NULL argument");
  *bytes = (int64_t)(c->bytes_scene + c->bytes_ibl + c->bytes_state + c->bytes_frame);
  return 0;
}

}  // extern "C"
