// Kernel declarations and the device-side job / pass / path-state layouts
// shared by the wavefront kernels and the host runtime.
#pragma once
#include "dev_scene.hpp"

// One render job as the kernels see it: its scene + IBL, its 52-float frame
// block, resolution, accumulation target and optional per-sample recording.
struct DevJob {
  SceneView S;
  FrameParams F;
  int32_t W, H, npix, rec_first;  // rec_first: first frame held by the record buffers
  float4* accum;
  float4* rec_L;                  // [frame - rec_first][pixel] finished samples (or null)
  uint32_t* rec_sig;
  int32_t* rec_prim;
};

// A contiguous run of paths in one pass: frames first..first+nframes-1 of
// one job, path ids [offset, offset + nframes*npix).
struct PassTask {
  int32_t job, first_frame, nframes, npix;
  int64_t offset;
};

// Structure-of-arrays path state, capacity = paths in flight. Arrays are
// indexed by path id.
struct PathState {
  float4* org;   // extension ray origin
  float4* dir;   // extension ray direction
  float4* hit;   // (tri id bits, t, u, v) of the last extension ray
  float4* pw;    // path weight xyz, last BSDF pdf in w
  float4* L;     // radiance xyz (+alpha); the finished sample at exit
  uint2* rng;    // generator state
  int32_t* job;  // job index of the path
  uint32_t* sig; // path signature (only when some job records samples)
  int32_t* prim; // primary hit triangle (only when recording)
  float4* sh_org;   // shadow origin xyz, flags bits in w (1 = IBL ray, 2 = point light ray)
  float4* sh_dir;   // IBL shadow direction, tmax
  float4* sh_c;     // IBL NEE radiance before path weight
  float4* sh_pw;    // path weight at the time of NEE
  float4* sh_pdir;  // point-light shadow direction, tmax
  float4* sh_pc;    // point-light radiance before path weight
};

__global__ void k_prep_tris(const int32_t* bvh_idx, const uint32_t* tris, const float4* verts, int32_t M, int32_t T,
                            float4* out);
__global__ void k_ibl_luminance(const float4* env, int64_t n, float* f);
__global__ void k_ibl_rows(const float* f, int w, int h, float* pdf, float* cdf, float* rowsum);
__global__ void k_ibl_marginal(const float* rowsum, int h, float* ypdf, float* ycdf, double* total);
__global__ void k_fill_i32(int32_t* p, int n, int32_t v);
__global__ void k_fill_f4(float4* p, int64_t n, float4 v);
__global__ void k_raygen(PathState ps, const DevJob* jobs, const PassTask* tasks, int ntasks, int64_t P, int32_t* counts);
__global__ void k_trace_ext(const DevJob* jobs, PathState ps, const int32_t* queue, const int32_t* count_ptr, int* spill,
                            unsigned long long* cnt);
__global__ void k_trace_shadow(const DevJob* jobs, PathState ps, const int32_t* queue, const int32_t* count_ptr, int* spill,
                               unsigned long long* cnt);
__global__ void k_shade(const DevJob* jobs, PathState ps, int iter, const int32_t* queue_in, const int32_t* count_in,
                        int32_t* queue_out, int32_t* count_out, int32_t* shq, int32_t* count_sh, unsigned long long* cnt);
__global__ void k_accumulate(PathState ps, const DevJob* jobs, const PassTask* tasks);
__global__ void k_trace_rays(SceneView S, const float4* rays, int64_t n, int any, int32_t* tri, float* t, float* u, float* v,
                             int* spill, unsigned long long* cnt);
