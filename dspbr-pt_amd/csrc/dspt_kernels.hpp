// Kernel declarations and the device-side job / task / path-state layouts
// shared by the wavefront kernels and the host runtime.
#pragma once
#include "dev_scene.hpp"

// One render job as the kernels see it: its scene + IBL, its 52-float frame
// block, resolution, accumulation target, the ring of finished frames that
// are not yet folded into the running mean, and optional per-sample records.
struct DevJob {
  SceneView S;
  FrameParams F;
  int32_t W, H, npix, rec_first;  // rec_first: first frame held by the record buffers
  // render region: the pixels [rx0, rx0 + rw) x [ry0, ry0 + rn / rw) of the
  // image (rows from the bottom); the whole image unless a caller renders a
  // tile (the reference's scissored tile passes, renderer.ts:632-643). The
  // region's pixels are the work items of a frame, its count completes the
  // frame and the fold covers them only.
  int32_t rx0, ry0, rw, rn;
  int32_t scene, pad;             // scene id (jobs of one scene share a tree and materials)
  float4* accum;
  float4* ring;                   // [frame % ring_frames][pixel] finished samples
  int32_t* done;                  // [frame % ring_frames] finished paths of that frame
  int32_t ring_frames;
  int32_t frame_end;              // one past the last frame of the current range
  float4* rec_L;                  // [frame - rec_first][pixel] finished samples (or null)
  uint32_t* rec_sig;
  int32_t* rec_prim;
};

// A contiguous run of work items: frames first..first+nframes-1 of one job,
// item ids [offset, offset + nframes*npix), frame-major then pixel.
struct PassTask {
  int32_t job, first_frame, nframes, npix;
  int64_t offset;
};

// Device-resident control block of the path pool. Items (job, frame, pixel)
// are issued in id order. Path state is dense: the live paths of a bounce
// occupy positions [0, qcount) of the current state buffer; the paths that
// continue are compacted into the other buffer, and the next bounce's new
// paths are appended behind them.
// Queue counts of one state buffer, side by side so the shading pass claims
// a path's next-state position and its shadow-queue slot with one 64-bit
// atomic (paths in the low word, shadow rays in the high word).
struct alignas(8) QueuePair {
  int32_t paths;   // live paths in the state buffer
  int32_t shadow;  // shadow rays queued by the paths written into it
};

struct StreamCtl {
  int64_t next_item;   // first item not yet issued
  int64_t item0;       // first item of the current refill
  int32_t regen_n;     // items issued by the current refill
  int32_t q_base;      // position where the refill appends
  QueuePair q[2];      // state buffer 0 / 1
  int32_t ext_next;    // next position a traversal wave takes
  int32_t sh_next;     // next shadow-queue entry a traversal wave takes
  int32_t pad;
};

// Written by the host before each chunk of bounces.
struct StreamSched {
  int64_t limit;   // items [next_item, limit) may be issued
  int32_t ntasks;  // entries in the task table
  int32_t solo;    // the one job with paths in flight, or -1
  int32_t jbits;   // top bits of the ray-order key given to the scene id (0: one scene)
  int32_t pad;
};

// Structure-of-arrays path state, capacity = pool size, indexed by
// position. Every per-path field the shading kernel reads is a 16-byte
// record (the job and bounce ride in the ray records' w), so a path read
// through the shading order touches six cache lines. Two instances alternate
// per bounce (the per-path fields are double-buffered); hit and the
// shadow-ray fields are shared: hit is written and read at the current
// position, the shadow fields at the path's slot in the shadow-ray queue.
struct PathState {
  float4* org;    // extension ray origin xyz, job index of the path (int bits) in w
  float4* dir;    // extension ray direction xyz, bounce index (int bits, 0 = camera ray) in w
  float4* hit;    // (tri id bits, t, u, v) of the last extension ray
  int32_t* hinst; // instance of that hit (two-level scenes; -1 otherwise)
  float4* pw;     // path weight xyz, last BSDF pdf in w
  float4* L;      // radiance xyz (+alpha)
  uint4* meta;    // generator state (x, y), frame, pixel: one 16-byte record
  uint32_t* sig;  // path signature (recording jobs only)
  int32_t* prim;  // primary hit triangle (recording jobs only)
  // shadow records, indexed by shadow queue slot (the queue holds the path)
  float4* sh_org;   // shadow origin xyz; w (int bits) = flags (1 = IBL ray, 2 = point light ray) | job << 2
  float4* sh_dir;   // IBL shadow direction, tmax
  float4* sh_c;     // IBL NEE radiance before path weight
  float4* sh_pw;    // path weight at the time of NEE
  float4* sh_pdir;  // point-light shadow direction, tmax
  float4* sh_pc;    // point-light radiance before path weight
};

// Work counters: kCntSlots copies of kNumCounters u64 (a 256-byte row each);
// a block adds into row blockIdx.x % kCntSlots, so the per-wave atomics of a
// launch spread over 256 addresses instead of queueing on one; readers sum
// the rows.
constexpr int kNumCounters = 32;
constexpr int kCntSlots = 256;
__device__ inline void stat_add(unsigned long long* cnt, int i, unsigned long long v) {
  wave_add_u64(cnt + (size_t)(blockIdx.x & (kCntSlots - 1)) * kNumCounters + i, v);
}

__global__ void k_prep_tris(const int32_t* bvh_idx, const uint32_t* tris, const float4* verts, int32_t M, int32_t T,
                            float4* out);
__global__ void k_boxes_finite(const float4* nodes, int32_t limit, int32_t* ok);
__global__ void k_tri_mat(const uint32_t* tris, const float4* verts, int32_t T, uint32_t* out);
__global__ void k_tri_verts(const uint32_t* tris, const float4* verts, int32_t T, float4* out);
__global__ void k_ibl_luminance(const float4* env, int64_t n, float* f);
__global__ void k_ibl_rows(const float* f, int w, int h, float* pdf, float* cdf, float* rowsum);
__global__ void k_ibl_marginal(const float* rowsum, int h, float* ypdf, float* ycdf, double* total);
__global__ void k_ibl_guides(const float* cdf, const float* ycdf, int w, int h, int32_t* guide, int32_t* yguide);
__global__ void k_fill_i32(int32_t* p, int n, int32_t v);
__global__ void k_fill_f4(float4* p, int64_t n, float4 v);
__global__ void k_stream_init(StreamCtl* ctl);
__global__ void k_plan(StreamCtl* ctl, const StreamSched* sch, int a, int32_t P, uint32_t* order_bins,
                       uint32_t* shade_bins);
__global__ void k_regen(PathState ps, const DevJob* jobs, const PassTask* tasks, const StreamSched* sch, const StreamCtl* ctl,
                        unsigned long long* cnt);
// Ray order of the extension queue (engine extension; the reference traces
// one pixel per fragment): the continuing paths [0, q_base) are bucketed by
// (origin cell of a 16x16x16 grid over the scene's root box, up or down)
// into perm; camera rays appended behind them keep pixel order.
constexpr int kOrderBins = 8192;
// launch shape of the key and scatter passes: blocks of 1024 threads (more
// loads in flight per CU), at most 512 of them (each block adds its bucket
// counts with one global atomic per non-empty bucket)
#ifndef DSPT_ORDER_THREADS
#define DSPT_ORDER_THREADS 1024
#endif
#ifndef DSPT_ORDER_BLOCKS
#define DSPT_ORDER_BLOCKS 512
#endif
static_assert(DSPT_ORDER_THREADS % 64 == 0 && DSPT_ORDER_THREADS <= 1024, "order passes: whole waves, at most 1024 threads");
__global__ void k_ray_keys(const DevJob* jobs, PathState ps, const StreamCtl* ctl, uint16_t* keys, uint32_t* bins,
                           const StreamSched* sch);
__global__ void k_bin_scan(uint32_t* bins);
__global__ void k_bin_scatter(const uint16_t* keys, const int32_t* n_ptr, uint32_t* bins, int32_t* perm, const int32_t* src,
                              const int32_t* ns_ptr);
__global__ void k_hit_keys(const DevJob* jobs, PathState ps, const int32_t* n_ptr, uint16_t* keys, uint32_t* bins,
                           const StreamSched* sch);
template <bool INST>
__global__ void k_trace_ext(const DevJob* jobs, PathState ps, const int32_t* count_ptr, int32_t* next_ptr, int* spill,
                            unsigned long long* cnt, const int32_t* perm, const int32_t* nsorted_ptr);
template <bool INST>
__global__ void k_trace_shadow(const DevJob* jobs, PathState ps, const int32_t* queue, const int32_t* count_ptr,
                               int32_t* next_ptr, int* spill, unsigned long long* cnt);
// DEBUG: the instance that also runs the debug program for jobs whose frame
// has a debug mode; the host launches it only while such a job is live.
template <bool DEBUG>
__global__ void k_shade(const DevJob* jobs, PathState cur, PathState nxt, const int32_t* count_in, QueuePair* qout,
                        int32_t* shq, unsigned long long* cnt, const int32_t* perm, const int32_t* nsorted_ptr);
__global__ void k_accum_ready(const DevJob* jobs, const int32_t* job_list, const int32_t* prog);
__global__ void k_advance(const DevJob* jobs, const int32_t* job_list, int n, int32_t* prog);
__global__ void k_trace_rays(SceneView S, const float4* rays, int64_t n, int any, int32_t* tri, float* t, float* u, float* v,
                             int32_t* inst, int* spill, unsigned long long* cnt);
