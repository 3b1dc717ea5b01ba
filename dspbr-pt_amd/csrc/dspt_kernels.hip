// Wavefront path-tracing kernels for gfx950 (wave64, VALU only).
//
// A fixed pool of path slots serves a stream of work items (job, frame,
// pixel); a job = one scene + IBL + camera/frame block + resolution. Every
// path carries its job index, so paths of several scenes, resolutions and
// frame counts share one set of launches. Per bounce:
//   plan + regen refill the slots freed on the previous bounce with camera
//                rays of the next items (renderer.frag:129-147)
//   trace_ext    closest hit of the extension rays in the active queue
//   shade        render state, material, RR, emission, BSDF sample, NEE sample
//                + BSDF eval/pdf; appends extension rays and shadow rays;
//                a finished path writes its sample into its job's frame ring
//                and frees its slot
//   trace_shadow any-hit test of the NEE shadow rays, adds the deferred
//                NEE radiance in the reference's summation order
// accum_ready folds every complete frame into its job's running mean in
// frame order (renderer.frag:153-156), so results do not depend on how
// items of different frames and jobs were interleaved in the pool.
#include "dspt_kernels.hpp"
#include "dev_shade.hpp"
#include "dev_post.hpp"
#include "dev_order.hpp"

// Occupancy target of the shading kernel (waves per SIMD; 0 = compiler's
// choice). Left alone the compiler gives the full-profile shading kernel
// (BSDF mixture + material textures) more than 256 registers, i.e. one wave
// per SIMD. At 3 (168 registers) it spills some of them to scratch and is
// still faster than at 2 (no spills): the kernel waits on memory, and the
// extra resident waves hide more of it.
#ifndef DSPT_SHADE_WAVES
#define DSPT_SHADE_WAVES 3
#endif
#if DSPT_SHADE_WAVES > 0
#define SHADE_BOUNDS __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DSPT_SHADE_WAVES, DSPT_SHADE_WAVES)))
#else
#define SHADE_BOUNDS __launch_bounds__(256)
#endif
// Traversal kernels: full occupancy (8 waves per SIMD, <= 64 VGPRs) keeps
// enough node fetches in flight. The two-level walk carries the world ray
// and the instance pointers on top (per lane: lanes may serve different
// scenes); at 7 waves (72 VGPRs, a few values spilled) it measured faster
// than at 6 without spills or at 8 with many.
#ifndef DSPT_TRACE_WAVES
#define DSPT_TRACE_WAVES 8
#endif
#ifndef DSPT_TRACE_WAVES_INST
#define DSPT_TRACE_WAVES_INST 7
#endif
#define TRACE_BOUNDS __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(INST ? DSPT_TRACE_WAVES_INST : DSPT_TRACE_WAVES)))

// --------------------------------------------------------------------------
// Scene preparation: per BVH index slot, the triangle's three positions in
// leaf order (bvh.glsl:72-99 resolved once at upload).
__global__ void k_prep_tris(const int32_t* __restrict__ bvh_idx, const uint32_t* __restrict__ tris,
                            const float4* __restrict__ verts, int32_t M, int32_t T, float4* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  int id = bvh_idx[i];
  float4 p[3] = {make_float4(0, 0, 0, 0), make_float4(0, 0, 0, 0), make_float4(0, 0, 0, 0)};
  if (id >= 0 && id < T) {
    for (int k = 0; k < 3; k++) p[k] = verts[(size_t)tris[(size_t)id * 3 + k] * 5];
  } else {
    id = -1;
  }
  out[i * 3 + 0] = make_float4(p[0].x, p[0].y, p[0].z, __int_as_float(id));
  out[i * 3 + 1] = make_float4(p[1].x, p[1].y, p[1].z, 0.0f);
  out[i * 3 + 2] = make_float4(p[2].x, p[2].y, p[2].z, 0.0f);
}

// Child codes of the engine's node array (see child_code in dev_scene.hpp):
// one thread per node below `limit` (the instance records of a two-level
// scene follow the nodes and are not nodes). Only a.w / b.w of inner nodes
// are written; the count/first fields read here are never written.
__global__ void k_encode_children(float4* __restrict__ nodes, int32_t N, int32_t limit) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= limit) return;
  if ((int)nodes[(size_t)n * 4 + 2].w > 0) return;  // a leaf node keeps its fields
  for (int side = 0; side < 2; side++) {
    float4 e = nodes[(size_t)n * 4 + side];
    const int child = (int)e.w;
    int code;
    if (child < 0 || child >= N || child == n) {
      code = kChildInvalid;
    } else if (child < limit) {
      const int cnt = (int)nodes[(size_t)child * 4 + 2].w, first = (int)nodes[(size_t)child * 4 + 3].w;
      code = (cnt > 0 && first >= 0 && first < (1 << 24)) ? -(1 + (first << 6) + (min(cnt, kMaxLeafTris) - 1)) : child;
    } else {
      code = child;
    }
    e.w = __int_as_float(code);
    nodes[(size_t)n * 4 + side] = e;
  }
}

// Material index of each triangle: the first vertex's pos.w, as
// fetch_hitpoint reads it (the shading order's key).
__global__ void k_tri_mat(const uint32_t* __restrict__ tris, const float4* __restrict__ verts, int32_t T,
                          uint32_t* __restrict__ out) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  out[t] = (uint32_t)verts[(size_t)tris[(size_t)t * 3] * 5].w;
}

// The three vertex records of each triangle, in corner order (15 float4
// per triangle), so shading reads them without the index indirection.
__global__ void k_tri_verts(const uint32_t* __restrict__ tris, const float4* __restrict__ verts, int32_t T,
                            float4* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // one float4 of the output
  if (i >= (int64_t)T * 15) return;
  const int64_t t = i / 15;
  const int r = (int)(i - t * 15), corner = r / 5, field = r - corner * 5;
  out[i] = verts[(size_t)tris[t * 3 + corner] * 5 + field];
}

// Clears *ok when an inner node below `limit` has a non-finite child box
// corner (the walk's slab tests read a.xyz, b.xyz, c.xyz, d.xyz of inner
// nodes only); *ok starts at 1.
__global__ void k_boxes_finite(const float4* __restrict__ nodes, int32_t limit, int32_t* __restrict__ ok) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= limit) return;
  const float4* nd = nodes + (size_t)n * 4;
  const float4 c = nd[2];
  if ((int)c.w > 0) return;  // a leaf
  const float4 a = nd[0], b = nd[1], d = nd[3];
  const bool fin = isfinite(a.x) && isfinite(a.y) && isfinite(a.z) && isfinite(b.x) && isfinite(b.y) && isfinite(b.z) &&
                   isfinite(c.x) && isfinite(c.y) && isfinite(c.z) && isfinite(d.x) && isfinite(d.y) && isfinite(d.z);
  if (!fin) atomicAnd(ok, 0);
}

// --------------------------------------------------------------------------
// IBL importance-sampling tables, renderer.ts:777-838, bit-identical: the
// luminance and the per-row sums are double (JS numbers), pdf is rounded to
// f32, the cdf is a sequential f32 scan. One lane per row (rows are
// independent; the scan order inside a row is the reference's).
__global__ void k_ibl_luminance(const float4* __restrict__ env, int64_t n, float* __restrict__ f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 c = env[i];
  f[i] = (float)((double)c.x * 0.299 + (double)c.y * 0.587 + (double)c.z * 0.114);
}
__device__ double cdf_row(const float* __restrict__ f, float* __restrict__ pdf, float* __restrict__ cdf, int num) {
  double sum = 0.0;
  for (int i = 0; i < num; i++) sum += (double)f[i];
  if (sum == 0.0) sum = 1.0;
  float acc = 0.0f;
  for (int i = 0; i < num; i++) {
    float p = (float)((double)f[i] / sum);
    pdf[i] = p;
    // the reference adds two f32 values in double and rounds the sum on the
    // Float32Array store; a 53-bit sum of 24-bit operands rounded to 24 bits
    // equals the correctly rounded f32 sum (double rounding is innocuous for
    // + when 53 >= 2*24 + 2, Figueroa 1995), so the f32 add is that result
    acc = (i == 0) ? p : acc + p;
    cdf[i] = acc;
  }
  cdf[num - 1] = 1.0f;
  return sum;
}
__global__ void k_ibl_rows(const float* __restrict__ f, int w, int h, float* __restrict__ pdf, float* __restrict__ cdf,
                           float* __restrict__ rowsum) {
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= h) return;
  size_t o = (size_t)y * w;
  rowsum[y] = (float)cdf_row(f + o, pdf + o, cdf + o, w);
}
__global__ void k_ibl_marginal(const float* __restrict__ rowsum, int h, float* __restrict__ ypdf, float* __restrict__ ycdf,
                               double* __restrict__ total) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *total = cdf_row(rowsum, ypdf, ycdf, h);
}

// --------------------------------------------------------------------------
// Guide tables for the IBL sampling searches: row r < rows is the cdf row r
// (n = w entries), row `rows` is the marginal (n = h entries).
__global__ void k_ibl_guides(const float* __restrict__ cdf, const float* __restrict__ ycdf, int w, int h,
                             int32_t* __restrict__ guide, int32_t* __restrict__ yguide) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nrow = (int64_t)h * (w + 1);
  if (i < nrow) {
    const int r = (int)(i / (w + 1)), g = (int)(i - (int64_t)r * (w + 1));
    guide[i] = lower_bound_f(cdf + (size_t)r * w, w, (float)g / (float)w);
  } else if (i < nrow + h + 1) {
    const int g = (int)(i - nrow);
    yguide[g] = lower_bound_f(ycdf, h, (float)g / (float)h);
  }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int n, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f4(float4* __restrict__ p, int64_t n, float4 v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_stream_init(StreamCtl* __restrict__ ctl) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    StreamCtl z{};
    *ctl = z;
  }
}

// One thread: sizes this bounce's refill (free positions vs scheduled
// items) and resets the queues the bounce appends to.
__global__ void k_plan(StreamCtl* __restrict__ ctl, const StreamSched* __restrict__ sch, int a, int32_t P,
                       uint32_t* __restrict__ order_bins, uint32_t* __restrict__ shade_bins) {
  if (order_bins)  // cleared here rather than by a memset, which a captured graph must not hold
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) order_bins[i] = 0u;
  if (shade_bins)
    for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) shade_bins[i] = 0u;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int Q = ctl->q[a].paths;
  const int F = P - Q;
  int64_t avail = sch->limit - ctl->next_item;
  if (avail < 0) avail = 0;
  const int n = (int)(avail < (int64_t)F ? avail : (int64_t)F);
  ctl->regen_n = n;
  ctl->item0 = ctl->next_item;
  ctl->q_base = Q;
  ctl->next_item += n;
  ctl->q[a].paths = Q + n;
  ctl->q[a ^ 1].paths = 0;
  ctl->q[a ^ 1].shadow = 0;
  ctl->ext_next = 0;
  ctl->sh_next = 0;
}

// ---- ray order of the extension queue -----------------------------------
// Bucket key of a continuing path: the Morton code of its origin's cell in a
// grid of 2^12 cells over the root box of its job's scene, then the sign of
// dir.y (up or down). The grid's 12 bits go to the axes by halving the
// longest cell side each time (scene_key_bits), so a flat terrain gets its
// cells in x and z. Rays of one bucket start close together, walk the same
// part of the tree and mostly hit the same objects, which keeps both the
// traversal's node fetches and the shading's material branches coherent.
// (Measured alternatives: 16 cells per axis; direction octant above an 8^3
// grid, 8^3 with the octant below, 16^3 alone, 16^3 with two direction bits.)
DI uint32_t ray_key(const SceneView& S, float4 o, float4 d) {
  f3 lo = sp3(0.0f), hi = sp3(0.0f);
  if (S.N > 0) {
    const float4* nd = S.nodes + (size_t)(S.top >= 0 ? S.top : 0) * 4;
    const float4 a = nd[0], b = nd[1], c = nd[2], e = nd[3];
    lo = f3{fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z)};
    hi = f3{fmaxf(b.x, e.x), fmaxf(b.y, e.y), fmaxf(b.z, e.z)};
  }
  const int bx = S.key_bits & 15, by = (S.key_bits >> 4) & 15, bz = (S.key_bits >> 8) & 15;
  auto cell = [](float x, float l, float h, int b) -> uint32_t {
    const float w = h - l;
    if (!(w > 0.0f) || b == 0) return 0u;
    const uint32_t nmax = (1u << b) - 1u;
    const float t = (x - l) * ((float)(1u << b) / w);
    return t >= (float)nmax ? nmax : (t > 0.0f ? (uint32_t)t : 0u);  // NaN -> 0
  };
  const uint32_t cx = cell(o.x, lo.x, hi.x, bx), cy = cell(o.y, lo.y, hi.y, by), cz = cell(o.z, lo.z, hi.z, bz);
  uint32_t m = 0u;
  for (int lvl = 11; lvl >= 0; lvl--) {  // interleaved from each axis's top bit down
    if (lvl < bx) m = (m << 1) | ((cx >> lvl) & 1u);
    if (lvl < by) m = (m << 1) | ((cy >> lvl) & 1u);
    if (lvl < bz) m = (m << 1) | ((cz >> lvl) & 1u);
  }
  return (m << 1) | (d.y < 0.0f ? 1u : 0u);
}
static_assert(kOrderBins == 8192, "the key has 13 bits");

// Keys of [0, q_base) and the bucket sizes (block-local counts in LDS, one
// global add per non-empty bucket and block). With several scenes in the
// pool the scene id takes the key's top sch->jbits bits (the cell's low bits
// give way), so a bucket holds one scene's rays.
__global__ void k_ray_keys(const DevJob* __restrict__ jobs, PathState ps, const StreamCtl* __restrict__ ctl,
                           uint16_t* __restrict__ keys, uint32_t* __restrict__ bins, const StreamSched* __restrict__ sch) {
  __shared__ uint32_t h[kOrderBins];
  const int n = ctl->q_base;
  const int jb = sch->jbits;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;  // small queues leave most blocks idle
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)q;
    const float4 o = ps.org[p];
    const DevJob* J = jobs + __float_as_int(o.w);
    uint32_t k = ray_key(J->S, o, ps.dir[p]);
    if (jb > 0) k = (((uint32_t)J->scene & ((1u << jb) - 1u)) << (13 - jb)) | (k >> jb);
    keys[p] = (uint16_t)k;
    atomicAdd(&h[k], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x)
    if (h[i]) atomicAdd(&bins[i], h[i]);
}

// Exclusive scan of the bucket sizes into bucket starts (one block of 1024).
__global__ void k_bin_scan(uint32_t* __restrict__ bins) {
  __shared__ uint32_t part[1024];
  const int t = threadIdx.x;
  constexpr int per = kOrderBins / 1024;
  uint32_t v[per], sum = 0;
  for (int k = 0; k < per; k++) { v[k] = bins[t * per + k]; sum += v[k]; }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
    const uint32_t x = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += x;
    __syncthreads();
  }
  uint32_t run = part[t] - sum;
  for (int k = 0; k < per; k++) { bins[t * per + k] = run; run += v[k]; }
}

// perm[start of bucket + rank] = position, for [0, *n_ptr): each block takes
// a contiguous range, counts it per bucket in LDS, reserves its slots with
// one global add per non-empty bucket, then places its positions. With src,
// the range walks the positions in the order src gives for entries below
// *ns_ptr (the ray order), so each bucket holds its paths in about that
// order (block by block).
__global__ void k_bin_scatter(const uint16_t* __restrict__ keys, const int32_t* __restrict__ n_ptr,
                              uint32_t* __restrict__ bins, int32_t* __restrict__ perm, const int32_t* __restrict__ src,
                              const int32_t* __restrict__ ns_ptr) {
  __shared__ uint32_t h[kOrderBins];
  const int n = *n_ptr;
  const int ns = src ? *ns_ptr : 0;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int lo = (int)min((int64_t)blockIdx.x * per, (int64_t)n), hi = min(lo + per, n);
  if (lo >= hi) return;  // block-uniform: no work for this block
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) h[i] = 0u;
  __syncthreads();
  for (int q = lo + threadIdx.x; q < hi; q += blockDim.x) atomicAdd(&h[keys[q < ns ? src[q] : q]], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x)
    if (h[i]) h[i] = atomicAdd(&bins[i], h[i]);  // count -> this block's first slot in the bucket
  __syncthreads();
  for (int q = lo + threadIdx.x; q < hi; q += blockDim.x) {
    const int p = q < ns ? src[q] : q;
    perm[atomicAdd(&h[keys[p]], 1u)] = p;
  }
}

// Shading order (mode DSPT_ORDER_SHADE): the whole queue [0, qcount) keyed
// by the material of each path's hit (misses in bucket 0), so that a wave
// shades one material's lobes together.
// sch->solo >= 0: every path in the pool belongs to that job, so the job
// index is not read per path (half the bytes this pass reads); the triangle
// range check keeps the lookup inside that job's table whatever the hit.
__global__ void k_hit_keys(const DevJob* __restrict__ jobs, PathState ps, const int32_t* __restrict__ n_ptr,
                           uint16_t* __restrict__ keys, uint32_t* __restrict__ bins, const StreamSched* __restrict__ sch) {
  __shared__ uint32_t h[kOrderBins];
  const int n = *n_ptr;
  const int solo = sch->solo;
  if ((int64_t)blockIdx.x * blockDim.x >= n) return;
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x) h[i] = 0u;
  const uint32_t* __restrict__ solo_mat = solo >= 0 ? jobs[solo].S.tri_mat : nullptr;
  const int32_t solo_t = solo >= 0 ? jobs[solo].S.T : 0;
  __syncthreads();
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)q;
    const int tri = __float_as_int(ps.hit[p].x);
    uint32_t k = 0u;
    if (tri >= 0) {
      if (solo_mat) k = tri < solo_t ? 1u + solo_mat[tri] % (uint32_t)(kOrderBins - 1) : 0u;
      else {  // several jobs: material indices offset by scene, so scenes do not share buckets
        const DevJob* J = jobs + __float_as_int(ps.org[p].w);
        k = 1u + ((uint32_t)J->scene * 509u + J->S.tri_mat[tri]) % (uint32_t)(kOrderBins - 1);
      }
    }
    keys[p] = (uint16_t)k;
    atomicAdd(&h[k], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kOrderBins; i += blockDim.x)
    if (h[i]) atomicAdd(&bins[i], h[i]);
}

// Camera rays for the items of this refill, renderer.frag:129-142 +
// rng_init (rng.glsl:23-29). Item -> task by binary search over the task
// table (offsets ascending).
__global__ void k_regen(PathState ps, const DevJob* __restrict__ jobs, const PassTask* __restrict__ tasks,
                        const StreamSched* __restrict__ sch, const StreamCtl* __restrict__ ctl, unsigned long long* __restrict__ cnt) {
  const int n = ctl->regen_n;
  const int64_t item0 = ctl->item0;
  const int qb = ctl->q_base;
  const int ntasks = sch->ntasks;
  unsigned issued = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int64_t item = item0 + i;
    int lo = 0, hi = ntasks - 1;  // last task with offset <= item
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tasks[mid].offset <= item) lo = mid; else hi = mid - 1;
    }
    const PassTask T = tasks[lo];
    const DevJob* J = jobs + T.job;
    int fslot, rpix;  // frame of the task, pixel of the render region
    item_frame_pixel(item - T.offset, T.npix, T.nframes, fslot, rpix);
    int W = J->W, H = J->H;
    int px, py;
    region_tile_pixel(rpix, J->rw, J->rn / J->rw, px, py);
    px += J->rx0;
    py += J->ry0;
    int pix = py * W + px;
    int frame = T.first_frame + fslot;
    const FrameParams& F = J->F;
    Rng rng = rng_seed(frame * (int)F.max_bounces, px, py);
    float r0 = rng.uniform(), r1 = rng.uniform();
    float vux = (2.0f * (float(px) + 0.5f)) / float(W) - 1.0f;
    float vuy = (2.0f * (float(py) + 0.5f)) / float(H) - 1.0f;
    float offx = r0 * F.inv_res[0], offy = r1 * F.inv_res[1];
    float aspect = F.inv_res[1] / F.inv_res[0];
    float ux = (vux * aspect + offx) * F.cam[3];
    float uy = (vuy * 1.0f + offy) * F.cam[3];
    f3 d = nrm3(f3{ux, uy, -F.focal_length});
    const float* m = F.view;
    f3 wd = f3{m[0] * d.x + m[4] * d.y + m[8] * d.z, m[1] * d.x + m[5] * d.y + m[9] * d.z, m[2] * d.x + m[6] * d.y + m[10] * d.z};
    const int p = qb + i;
    ps.org[p] = make_float4(F.cam[0], F.cam[1], F.cam[2], __int_as_float(T.job));
    ps.dir[p] = make_float4(wd.x, wd.y, wd.z, __int_as_float(0));
    ps.pw[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    ps.L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    ps.meta[p] = make_uint4(rng.x, rng.y, (uint32_t)frame, (uint32_t)pix);
    if (J->rec_L) ps.sig[p] = 2166136261u;
    issued++;
  }
  stat_add(cnt, 12, issued);
}

// Lanes of a traversal wave take new rays from the queue whenever at least
// that many of them are idle (one atomic per wave), so a wave is not held by
// its slowest ray; every ray still runs the same step sequence. Per kernel:
// closest hit (EXT) and shadow rays (SH).
#ifndef DSPT_REFILL_EXT
#define DSPT_REFILL_EXT 48
#endif
#ifndef DSPT_REFILL_SH
#define DSPT_REFILL_SH 48
#endif
// Inner-node steps a wave takes before it tests the leaves its lanes reached.
#ifndef DSPT_INNER_RUN_EXT
#define DSPT_INNER_RUN_EXT 4
#endif
#ifndef DSPT_INNER_RUN_SH
#define DSPT_INNER_RUN_SH 4
#endif

// Wave-uniform refill: idle lanes get consecutive queue entries. With
// CHUNK > 0 a wave claims at least CHUNK positions per atomic on the shared
// queue counter and keeps what it did not hand out as a reserve
// [res_lo, res_hi) for its next refills, so the contended counter sees
// fewer atomics; every position is still taken once (CHUNK = 0: exactly the
// idle lanes, no reserve). Returns the queue index for this lane (or -1);
// `exhausted` is set once the queue is fully claimed, and the wave ends when
// that holds and its reserve is empty. Measured: chunks of 128 speed up the
// one-level closest-hit trace (−3 %) but slow the shadow trace and the
// two-level traces (+3 to +5 %), which keep CHUNK = 0; and a small queue
// (a one-frame pass of a small job) would sit in the reserves of a few waves
// while the rest idle, so chunks apply only from four per wave of the grid.
#ifndef DSPT_REFILL_CHUNK
#define DSPT_REFILL_CHUNK 128
#endif
DD int refill_chunk(int count, int chunk) {
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  return (int64_t)count >= waves * chunk * 4 ? chunk : 0;
}
template <int MIN_IDLE>
DD int refill(bool active, int count, int32_t* next_ptr, bool& exhausted, int& res_lo, int& res_hi, int chunk) {
  const unsigned long long idle = __ballot(!active);
  const int n_idle = __popcll(idle);
  const int have = res_hi - res_lo;  // wave-uniform
  if (n_idle < MIN_IDLE || (exhausted && have == 0)) return -1;
  const int lane = __lane_id();
  const int rank = __popcll(idle & ((1ull << lane) - 1ull));
  int q;
  if (have >= n_idle || exhausted) {  // from the reserve alone
    const int take = min(have, n_idle);
    q = rank < take ? res_lo + rank : -1;
    res_lo += take;
  } else {  // the reserve first, then a new claim
    const int leader = __ffsll((long long)idle) - 1;
    const int want = max(chunk, n_idle - have);
    int base = 0;
    if (lane == leader) base = atomicAdd(next_ptr, want);
    base = __builtin_amdgcn_readfirstlane(__shfl(base, leader));
    if ((int64_t)base + want >= count) exhausted = true;
    const int got = (int)max((int64_t)0, min((int64_t)base + want, (int64_t)count) - base);
    const int r2 = rank - have;
    q = rank < have ? res_lo + rank : (r2 < got ? base + r2 : -1);
    const int used = min(got, n_idle - have);
    res_lo = base + used;
    res_hi = base + got;
  }
  return active ? -1 : q;
}

// The loads of a ray start (its record and its scene's node and triangle
// arrays and sizes) complete at the start: otherwise the compiler treats
// them as possibly in flight at every traversal step and waits on vmcnt(0)
// there, which also waits for the wave's outstanding stores before the
// next node fetch issues.
#ifndef DSPT_START_WAIT_ON
#define DSPT_START_WAIT_ON 1
#endif
#define DSPT_START_WAIT()                                      \
  do {                                                         \
    if (DSPT_START_WAIT_ON) __builtin_amdgcn_s_waitcnt(0x0F70); \
  } while (0)

// Per-lane ray walker: one-level scenes, or (INST) scenes whose jobs may be
// two-level (a per-lane choice by the scene's top-level BVH).
template <bool INST> struct Walker;
template <> struct Walker<false> {
  Geo G;
  RayState r;
  bool fast;  // this ray may take slab_fast (a lane mask, no vector register)
  DD void start(const SceneView& S, LaneStack& st, f3 o, f3 d, float tfar) {
    G = geo_of(S);
    ray_init(r, st, o, d, tfar);
    fast = slab_fast_ok(S.fast_slab, r.org, r.inv);
  }
  // FAST: the caller checked that every busy lane of the wave has `fast`
  template <bool FAST = false> DD int inner(LaneStack& st, unsigned& nv) { return trav_inner<FAST>(G, r, st, nv); }
  template <bool ANY> DD bool leaf(unsigned& nt) { return trav_leaf<ANY>(G, r, nt); }
  DD void restart(LaneStack& st, f3 o, f3 d, float tfar) {
    ray_init(r, st, o, d, tfar);
    fast = false;  // the scene's flag is not kept here: the exact test
  }
  DD int inst() const { return -1; }
};
template <> struct Walker<true> {
  Geo2 G;
  RayState r;
  InstState q;
  DD void start(const SceneView& S, LaneStack& st, f3 o, f3 d, float tfar) {
    G = geo2_of(S);
    restart(st, o, d, tfar);
  }
  static constexpr bool fast = false;  // object-space rays may have an infinite 1/dir: the exact slab test
  template <bool FAST = false> DD int inner(LaneStack& st, unsigned& nv) {
    return G.top >= 0 ? trav_inner2(G, r, q, st, nv) : trav_inner(G.g, r, st, nv);
  }
  template <bool ANY> DD bool leaf(unsigned& nt) { return G.top >= 0 ? trav_leaf2<ANY>(G, r, q, nt) : trav_leaf<ANY>(G.g, r, nt); }
  DD void restart(LaneStack& st, f3 o, f3 d, float tfar) {
    ray_init(r, st, o, d, tfar);
    if (G.top >= 0) r.next = G.top;
    inst_init(q);
  }
  DD int inst() const { return q.hit; }
};

template <bool INST>
__global__ TRACE_BOUNDS void k_trace_ext(const DevJob* __restrict__ jobs, PathState ps, const int32_t* __restrict__ count_ptr,
                                         int32_t* __restrict__ next_ptr, int* __restrict__ spill,
                                         unsigned long long* __restrict__ cnt, const int32_t* __restrict__ perm,
                                         const int32_t* __restrict__ nsorted_ptr) {
  __shared__ int smem[kLdsStack * kTraceBlock];
  const int count = *count_ptr;
  const int nsorted = perm ? *nsorted_ptr : 0;  // queue entries [0, nsorted) are read through perm
  const int chunk = INST ? 0 : refill_chunk(count, DSPT_REFILL_CHUNK);
  LaneStack st;
  st.lds = smem + threadIdx.x;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  st.spill = spill + gtid * (kStackSize - kLdsStack);
  unsigned nv = 0, nt = 0, nr = 0;
  bool active = false, exhausted = false;
  int res_lo = 0, res_hi = 0;  // the wave's reserve of claimed queue positions
  int p = 0;
  Walker<INST> w;
  auto finish_ray = [&]() {
    st_stream(ps.hit + p, make_float4(__int_as_float(w.r.h.tri), w.r.h.t, w.r.h.u, w.r.h.v));  // read once, by the next passes
    if (INST) ps.hinst[p] = w.inst();
    nr++;
    active = false;
  };
  for (;;) {
    const int q = refill<DSPT_REFILL_EXT>(active, count, next_ptr, exhausted, res_lo, res_hi, chunk);
    if (q >= 0) {
      p = q < nsorted ? perm[q] : q;
      const float4 o = ps.org[p], d = ps.dir[p];
      w.start(jobs[__float_as_int(o.w)].S, st, xyz(o), xyz(d), kTFAR_MAX);
      active = true;
      DSPT_START_WAIT();
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }
    // inner nodes until every busy lane has reached its next leaf (or at
    // most DSPT_INNER_RUN_EXT steps), then the leaves together; the slab test runs on
    // hardware min/max when every busy lane allows it (a uniform branch)
    const bool wfast = !INST && __ballot(active && !w.fast) == 0;
    bool leaf = false;
    for (int k = 0; k < DSPT_INNER_RUN_EXT; k++) {
      const bool inner = active && !leaf;
      if (__ballot(inner) == 0) break;
      if (inner) {
        const int res = wfast ? w.template inner<true>(st, nv) : w.template inner<false>(st, nv);
        if (res == 1) leaf = true;
        else if (res == 2) finish_ray();
      }
    }
    if (leaf) {
      w.template leaf<false>(nt);
      if (ray_done(w.r)) finish_ray();
    }
  }
  stat_add(cnt, 4, nv);
  stat_add(cnt, 5, nt);
  stat_add(cnt, 9, nr);
}

// Shadow rays of a path: the IBL ray (flag 1), then the point-light ray
// (flag 2); the unoccluded contributions are summed in that order and added
// as L += (L_ibl + L_point) * path_weight (misptdl.glsl:72-74).
template <bool INST>
__global__ TRACE_BOUNDS void k_trace_shadow(const DevJob* __restrict__ jobs, PathState ps, const int32_t* __restrict__ queue,
                                            const int32_t* __restrict__ count_ptr, int32_t* __restrict__ next_ptr,
                                            int* __restrict__ spill, unsigned long long* __restrict__ cnt) {
  __shared__ int smem[kLdsStack * kTraceBlock];
  const int count = *count_ptr;
  LaneStack st;
  st.wait_spill = !INST;
  st.lds = smem + threadIdx.x;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  st.spill = spill + gtid * (kStackSize - kLdsStack);
  unsigned nv = 0, nt = 0, nr = 0;
  bool active = false, exhausted = false;
  int res_lo = 0, res_hi = 0;  // the wave's reserve of claimed queue positions
  int p = 0, flags = 0, phase = 0;  // p: the lane's shadow queue slot
  Walker<INST> w;
  f3 org = sp3(0.0f), contrib = sp3(0.0f);
  for (;;) {
    const int q = refill<DSPT_REFILL_SH>(active, count, next_ptr, exhausted, res_lo, res_hi, 0);
    if (q >= 0) {
      p = q;
      const float4 so = ps.sh_org[p];
      flags = __float_as_int(so.w) & 3;
      org = xyz(so);
      contrib = sp3(0.0f);
      phase = (flags & 1) ? 0 : 1;
      const float4 sd = phase == 0 ? ps.sh_dir[p] : ps.sh_pdir[p];
      w.start(jobs[__float_as_int(so.w) >> 2].S, st, org, xyz(sd), sd.w);
      active = true;
      if (!INST) DSPT_START_WAIT();
    }
    if (__ballot(active) == 0) {
      if (exhausted) break;
      continue;
    }
    // the exact slab test only: the fast one measured no gain here (with and
    // without a register spill)
    bool leaf = false, fin = false;
    for (int k = 0; k < DSPT_INNER_RUN_SH; k++) {
      const bool inner = active && !leaf && !fin;
      if (__ballot(inner) == 0) break;
      if (inner) {
        const int res = w.inner(st, nv);
        if (res == 1) leaf = true;
        else if (res == 2) fin = true;
      }
    }
    if (leaf) fin = w.template leaf<true>(nt) || ray_done(w.r);
    if (active && fin) {
      nr++;
      if (phase == 0) {
        if (w.r.h.tri < 0) contrib = xyz(ps.sh_c[p]);
      } else {
        if (w.r.h.tri < 0) contrib = contrib + xyz(ps.sh_pc[p]);
      }
      if (phase == 0 && (flags & 2)) {
        phase = 1;
        const float4 pd = ps.sh_pdir[p];
        w.restart(st, org, xyz(pd), pd.w);
      } else {
        const int path = queue[p];
        const float4 L = ps.L[path];
        const f3 wt = xyz(ps.sh_pw[p]);
        const f3 nl = xyz(L) + contrib * wt;
        ps.L[path] = make_float4(nl.x, nl.y, nl.z, L.w);
        active = false;
      }
    }
  }
  stat_add(cnt, 10, nv);
  stat_add(cnt, 11, nt);
  stat_add(cnt, 3, nr);
}

template __global__ void k_trace_ext<false>(const DevJob*, PathState, const int32_t*, int32_t*, int*, unsigned long long*,
                                              const int32_t*, const int32_t*);
template __global__ void k_trace_ext<true>(const DevJob*, PathState, const int32_t*, int32_t*, int*, unsigned long long*,
                                              const int32_t*, const int32_t*);
template __global__ void k_trace_shadow<false>(const DevJob*, PathState, const int32_t*, const int32_t*, int32_t*, int*,
                                               unsigned long long*);
template __global__ void k_trace_shadow<true>(const DevJob*, PathState, const int32_t*, const int32_t*, int32_t*, int*,
                                              unsigned long long*);

DI void sig_mix(uint32_t& h, uint32_t v) { h = (h ^ v) * 16777619u; }

// Finished sample: the reference returns early (no clamp) on a primary miss
// and on an absorbed BSDF sample; every other exit clamps (misptdl.glsl:92).
DI float4 finish(f3 L, float thr, bool do_clamp) {
  if (do_clamp && thr > 0.0f) L = vmin(vmax(L, sp3(0.0f)), sp3(thr));
  return make_float4(L.x, L.y, L.z, 1.0f);
}

// atomicAdd(addr, 1) for every active lane, one atomic per distinct address
// in the wave (lanes of one frame share a counter).
DD void atomic_inc_grouped(int32_t* addr) {
  const uint64_t a = (uint64_t)(uintptr_t)addr;
  for (;;) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const bool mine = a == (((uint64_t)hi << 32) | lo);
    const unsigned long long m = __ballot(mine);
    if (mine) {
      if ((int)__lane_id() == __ffsll((long long)m) - 1) atomicAdd(addr, (int32_t)__popcll(m));
      break;
    }
  }
}

// The debug program (integrator/debug.glsl:1-83), which the renderer runs
// whenever debugMode is not 'None' (renderer.ts:1020), u_debug_mode being
// the mode's index in debugModes (renderer.ts:181, 406; -1 for an unknown
// name): the camera ray's first hit shows one attribute of its render state;
// a miss shows the environment, or its pdf (x 10) or cdf table, when the
// background comes from the IBL, and vec4(0) otherwise. Out of line, in the
// shading kernel's DEBUG instance only.
__device__ __noinline__ float4 debug_sample(const SceneView& S, const FrameParams& F, f3 org, f3 dir, int tri, float4 hit,
                                            int inst) {
  const int mode = (int)F.debug_mode;
  if (tri < 0) {
    if (!(F.bg_from_ibl > 0.0f)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float th;
    const float2 uv = dir_uv(rotate_phi(dir, F.ibl_rotation, false), th);
    const bool have = S.has_ibl && S.ew > 0 && S.eh > 0;
    if (mode == 10) {  // vec4(texture(pdf).xyz, 1.0) * 10.0
      const float p = have ? table_lookup(S.pdf, S.ew, S.eh, uv) : 0.0f;
      return make_float4(p * 10.0f, 0.0f * 10.0f, 0.0f * 10.0f, 1.0f * 10.0f);
    }
    if (mode == 11) return make_float4(have ? table_lookup(S.cdf, S.ew, S.eh, uv) : 0.0f, 0.0f, 0.0f, 1.0f);
    const f3 e = have ? env_lookup(S, uv) : sp3(0.0f);
    return make_float4(e.x, e.y, e.z, 1.0f);
  }
  const Hitpoint hp = fetch_hitpoint(S, org, dir, tri, hit.y, hit.z, hit.w, inst);
  const Shading sh = configure(S, hp, -dir);
  const bsdf::GltfInputs& m = sh.in;
  f3 c = sp3(0.0f);
  switch (mode) {
    case 1: c = m.baseColor; break;                          // D_ALBEDO
    case 2: c = sp3(m.metallic); break;                      // D_METAL
    case 3: c = sp3(m.roughness); break;                     // D_ROUGHNESS
    case 4: c = sh.n; break;                                 // D_NORMAL
    case 5: c = sh.tanA; break;                              // D_TANGENT (after the anisotropy rotation)
    case 6: c = cross3(sh.n, sh.tanA) * hp.tw; break;        // D_BITANGENT: calculateBasis(c.n, c.t).b
    case 7: c = sp3(m.transmission); break;                  // D_TRANSPARENCY
    case 8: c = f3{hp.uv0.x, hp.uv0.y, 0.0f}; break;         // D_UV0
    case 9: c = sp3(m.clearcoat); break;                     // D_CLEARCOAT
    case 12: c = sp3(m.specular); break;                     // D_SPECULAR
    case 13: c = m.specularColor; break;                     // D_SPECULAR_TINT
    case 14: c = m.diffuseTransmission * m.diffuseTransmissionColor; break;  // D_TRANSLUCENCY
    default: break;
  }
  return make_float4(c.x, c.y, c.z, 1.0f);
}

template <bool DEBUG>
__global__ SHADE_BOUNDS void k_shade(const DevJob* __restrict__ jobs, PathState ps, PathState nx,
                                   const int32_t* __restrict__ count_in, QueuePair* __restrict__ qout,
                                   int32_t* __restrict__ shq,
                                   unsigned long long* __restrict__ cnt, const int32_t* __restrict__ perm,
                                   const int32_t* __restrict__ nsorted_ptr) {
  const int count = *count_in;
  const int nsorted = perm ? *nsorted_ptr : 0;  // shade in the traversal's order
  unsigned n_paths = 0, n_hits = 0, n_miss = 0, n_nee = 0, n_ext = 0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
    const int p = q < nsorted ? perm[q] : (int)q;
    const float4 o4 = ld_stream(ps.org + p), d4 = ld_stream(ps.dir + p);
    const uint4 meta = ld_stream(ps.meta + p);
    const int jid = __float_as_int(o4.w);
    const DevJob* J = jobs + jid;
    const bool rec = J->rec_L != nullptr;
    uint32_t sig = rec ? ps.sig[p] : 0u;
    float4 out;
    bool term = true;
    do {
      const SceneView& S = J->S;
      const FrameParams& F = J->F;
      const int maxB = (int)F.max_bounces;
      const int bounce = __float_as_int(d4.w);
      float4 hit = ld_stream(ps.hit + p);
      int tri = __float_as_int(hit.x);
      f3 org = xyz(o4), dir = xyz(d4);
      float4 pw4 = ld_stream(ps.pw + p);
      f3 pw = xyz(pw4);
      float last_pdf = pw4.w;
      float4 L4 = ld_stream(ps.L + p);
      f3 L = xyz(L4);

      if (DEBUG && F.debug_mode != 0.0f) {  // the debug program: one hit, no bounce
        if (rec) { sig_mix(sig, tri < 0 ? 0xFFFFFFFFu : (uint32_t)tri); ps.prim[p] = tri; }
        out = debug_sample(S, F, org, dir, tri, hit, (tri >= 0 && S.top >= 0) ? ps.hinst[p] : -1);
        break;
      }
      if (tri < 0) {
        if (rec) sig_mix(sig, 0xFFFFFFFFu);
        if (bounce == 0) {
          // primary miss: background, returned unclamped (misptdl.glsl:41-46)
          if (F.bg_from_ibl > 0.0f) { f3 e = ibl_eval(S, F, dir) * pw; out = make_float4(e.x, e.y, e.z, 1.0f); }
          else out = make_float4(pm::ppow(F.bg[0], 2.2f), pm::ppow(F.bg[1], 2.2f), pm::ppow(F.bg[2], 2.2f), F.bg[3]);
          if (rec) ps.prim[p] = -1;
        } else {
          n_miss++;
          float ip = ibl_pdf(S, F, dir);
          float mis = last_pdf / (last_pdf + ip);
          L = L + ibl_eval(S, F, dir) * pw * mis * F.use_ibl;
          out = finish(L, F.clamp_threshold, true);
        }
        break;
      }
      if (rec) { sig_mix(sig, (uint32_t)tri); if (bounce == 0) ps.prim[p] = tri; }
      if (bounce >= maxB) {
        out = finish(L, F.clamp_threshold, true);
        break;
      }
      Rng rng{meta.x, meta.y};
      // Russian roulette (renderer.frag:119-123)
      pw = pw * (1.0f / (1.0f - 0.1f));
      if (bounce > 2 && rng.uniform() <= 0.1f) {
        if (rec) sig_mix(sig, 1u);
        out = finish(L, F.clamp_threshold, true);
        break;
      }
      n_hits++;
      Hitpoint hp = fetch_hitpoint(S, org, dir, tri, hit.y, hit.z, hit.w, S.top >= 0 ? ps.hinst[p] : -1);
      f3 wi = -dir;
      Shading sh = configure(S, hp, wi);
      // Beer-Lambert absorption inside thick media (misptdl.glsl:60-63)
      if (sh.backside && !sh.thin) {
        f3 sigma = -vlog(sh.attColor) / sh.attDist;
        pw = pw * vexp(-sigma * hit.y);
      }
      L = L + sh.emissive * 1.0f * pw;

      // BSDF bounce (renderer.frag:94-116)
      f3 wo, bw = sp3(1.0f);
      float pdf = 1.0f;
      // the closure is built once per hit, before the sample (the reference
      // builds it for the sample and again for NEE; both builds are the same
      // pure function of the material), so the glTF inputs are dead before
      // the BSDF code and its registers
      const bsdf::Closure cl = bsdf::build_closure(sh.in, sh.tanA, S.full_profile != 0);
      const bsdf::Frame3 fr = bsdf::aniso_frame(sh.n, cl.tanA);
      if (rng.uniform() > sh.cutout) {
        wo = -wi;
        if (rec) sig_mix(sig, 2u);
      } else {
        float lx = rng.uniform(), ly = rng.uniform(), bz = rng.uniform();
        float comp = rng.uniform();
        bsdf::Sample bs = bsdf::sample(cl, fr, sh.n, wi, comp, lx, ly, bz);
        pdf = bs.pdf;
        bw = bs.w;
        wo = bs.dir;
        if (rec) sig_mix(sig, 3u + (uint32_t)(int)bs.specular);
        if (pdf < kEPS_PDF) {
          // absorbed: returned unclamped
          if (rec) sig_mix(sig, 7u);
          out = make_float4(L.x, L.y, L.z, 1.0f);
          break;
        }
      }
      last_pdf = pdf;
      f3 gn = dot3(hp.ng_raw, wo) < 0.0f ? -hp.ng_raw : hp.ng_raw;
      f3 pos = hp.pos + gn * F.ray_eps;

      // next-event estimation (misptdl.glsl:2-32); radiance added after the
      // shadow test
      float n0 = rng.uniform(), n1 = rng.uniform();
      int flags = 0;
      f3 c_ibl = sp3(0.0f), c_pt = sp3(0.0f);
      f3 ldir = sp3(0.0f);
      float4 pdir = make_float4(0, 0, 0, 0);
      f3 nn = sh.backside ? -sh.n : sh.n;
      if (F.use_ibl > 0.0f) {
        float ipdf;
        f3 d = ibl_sample(S, F, n0, n1, ipdf);
        float cosNL = sat(dot3(d, nn));
        if (cosNL > kEPS_COS && ipdf > kEPS_PDF) {
          f3 e = bsdf::eval(cl, fr, sh.n, wi, d) / gmax(fabsf(dot3(sh.n, d)), kEPS);
          float bp = bsdf::pdf(cl, fr, sh.n, wi, d);
          f3 Li = ibl_eval(S, F, d) * e * cosNL / ipdf;
          if (bp > kEPS_PDF) Li = Li * (ipdf / (ipdf + bp));
          c_ibl = Li;
          ldir = nrm3(d);
          flags |= 1;
          n_nee++;
        }
      }
      if (F.light_pos[3] > 0.5f) {
        f3 lp = f3{F.light_pos[0], F.light_pos[1], F.light_pos[2]};
        f3 ld = lp - pos;
        float dist2 = dot3(ld, ld);
        ld = nrm3(ld);
        float cosNL = sat(dot3(ld, nn));
        f3 dd = lp - pos;
        pdir = make_float4(0, 0, 0, len3(dd));
        f3 dn = nrm3(dd);
        pdir.x = dn.x; pdir.y = dn.y; pdir.z = dn.z;
        flags |= 2;
        if (cosNL > kEPS_COS) {
          f3 e = bsdf::eval(cl, fr, sh.n, wi, ld) / gmax(fabsf(dot3(sh.n, ld)), kEPS);
          c_pt = e * (f3{F.light_emission[0], F.light_emission[1], F.light_emission[2]} / dist2) * cosNL;
        }
      }
      const f3 pw_nee = pw;  // path weight at the time of NEE
      pw = pw * bw;
      // the path continues at position o of the next state buffer and,
      // with a shadow ray, takes shadow-queue slot k: one 64-bit atomic per
      // wave claims both runs (lane order within the wave, as before)
      const unsigned long long act = __ballot(1), shm = __ballot(flags != 0);
      const int lane = __lane_id(), leader = __ffsll((long long)act) - 1;
      const unsigned long long below = (1ull << lane) - 1ull;
      unsigned long long base = 0;
      if (lane == leader)
        base = atomicAdd(reinterpret_cast<unsigned long long*>(qout),
                         (unsigned long long)__popcll(act) | ((unsigned long long)__popcll(shm) << 32));
      base = __shfl(base, leader);
      const int o = (int)(uint32_t)base + __popcll(act & below);
      n_ext++;
      if (flags) {
        const int k = (int)(base >> 32) + __popcll(shm & below);
        // the shadow records at the dense queue slot k (full-line stores,
        // read in order by the shadow trace); shq[k] names the path
        shq[k] = o;
        ps.sh_org[k] = make_float4(pos.x, pos.y, pos.z, __int_as_float(flags | (jid << 2)));
        ps.sh_dir[k] = make_float4(ldir.x, ldir.y, ldir.z, kTFAR_MAX);
        ps.sh_c[k] = make_float4(c_ibl.x, c_ibl.y, c_ibl.z, 0.0f);
        ps.sh_pw[k] = make_float4(pw_nee.x, pw_nee.y, pw_nee.z, 0.0f);
        if (flags & 2) { ps.sh_pdir[k] = pdir; ps.sh_pc[k] = make_float4(c_pt.x, c_pt.y, c_pt.z, 0.0f); }
      }
      st_stream(nx.org + o, make_float4(pos.x, pos.y, pos.z, __int_as_float(jid)));
      st_stream(nx.dir + o, make_float4(wo.x, wo.y, wo.z, __int_as_float(bounce + 1)));
      st_stream(nx.pw + o, make_float4(pw.x, pw.y, pw.z, last_pdf));
      st_stream(nx.L + o, make_float4(L.x, L.y, L.z, L4.w));
      st_stream(nx.meta + o, make_uint4(rng.x, rng.y, meta.z, meta.w));
      if (rec) { nx.sig[o] = sig; nx.prim[o] = ps.prim[p]; }
      term = false;
    } while (false);
    if (term) {
      // finished sample -> the job's frame ring (+ records)
      n_paths++;
      const int2 fp = make_int2((int)meta.z, (int)meta.w);
      const int rs = fp.x % J->ring_frames;
      J->ring[(int64_t)rs * J->npix + fp.y] = out;
      if (rec) {
        int64_t r = (int64_t)(fp.x - J->rec_first) * J->npix + fp.y;
        J->rec_L[r] = out;
        J->rec_sig[r] = sig;
        J->rec_prim[r] = ps.prim[p];
      }
      atomic_inc_grouped(J->done + rs);
    }
  }
  stat_add(cnt, 0, n_paths);
  stat_add(cnt, 6, n_hits);
  stat_add(cnt, 7, n_miss);
  stat_add(cnt, 8, n_nee);
  stat_add(cnt, 2, n_ext);
}
template __global__ void k_shade<false>(const DevJob*, PathState, PathState, const int32_t*, QueuePair*, int32_t*,
                                        unsigned long long*, const int32_t*, const int32_t*);
template __global__ void k_shade<true>(const DevJob*, PathState, PathState, const int32_t*, QueuePair*, int32_t*,
                                       unsigned long long*, const int32_t*, const int32_t*);

// Running mean (renderer.frag:153-156) over every frame that is complete,
// in frame order, starting at the job's next frame to fold (grid.y = job),
// for the pixels of the job's render region (one thread each).
__global__ void k_accum_ready(const DevJob* __restrict__ jobs, const int32_t* __restrict__ job_list,
                              const int32_t* __restrict__ prog) {
  const int jid = job_list[blockIdx.y];
  const DevJob* J = jobs + jid;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;  // pixel of the render region
  const int npix = J->npix, rn = J->rn;
  if (r >= rn) return;
  const int R = J->ring_frames, end = J->frame_end;
  int f = prog[jid];
  if (f >= end || J->done[f % R] != rn) return;
  const int i = (J->ry0 + r / J->rw) * J->W + J->rx0 + r % J->rw;
  float4 a = J->accum[i];
  // at most R frames: the ring slot after them still holds the first one
  for (int k = 0; k < R && f < end && J->done[f % R] == rn; k++) {
    float4 c = J->ring[(int64_t)(f % R) * npix + i];
    float N = float(f);
    a.x = (a.x * (N - 1.0f) + c.x) / N;
    a.y = (a.y * (N - 1.0f) + c.y) / N;
    a.z = (a.z * (N - 1.0f) + c.z) / N;
    a.w = (a.w * (N - 1.0f) + c.w) / N;
    f++;
  }
  J->accum[i] = a;
}

// After k_accum_ready: advances each job's next frame to fold and releases
// the ring slots of the folded frames.
__global__ void k_advance(const DevJob* __restrict__ jobs, const int32_t* __restrict__ job_list, int n, int32_t* __restrict__ prog) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int jid = job_list[k];
    const DevJob* J = jobs + jid;
    const int R = J->ring_frames, end = J->frame_end, rn = J->rn;
    int f = prog[jid];
    for (int k = 0; k < R && f < end && J->done[f % R] == rn; k++) {
      J->done[f % R] = 0;
      f++;
    }
    prog[jid] = f;
  }
}

// Standalone traversal on caller rays {org.xyz, tfar, dir.xyz, pad}.
__global__ __launch_bounds__(kTraceBlock) void k_trace_rays(SceneView S, const float4* __restrict__ rays, int64_t n, int any,
                                                            int32_t* __restrict__ tri, float* __restrict__ t, float* __restrict__ u,
                                                            float* __restrict__ v, int32_t* __restrict__ inst,
                                                            int* __restrict__ spill, unsigned long long* __restrict__ cnt) {
  __shared__ int smem[kLdsStack * kTraceBlock];
  LaneStack st;
  st.lds = smem + threadIdx.x;
  int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  st.spill = spill + gtid * (kStackSize - kLdsStack);
  unsigned nv = 0, nt = 0;
  for (int64_t i = gtid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float4 a = rays[i * 2], b = rays[i * 2 + 1];
    RayHit h;
    int k = -1;
    if (S.top >= 0) {
      const Geo2 G = geo2_of(S);
      h = any ? traverse2<true>(G, xyz(a), xyz(b), a.w, st, nv, nt, k) : traverse2<false>(G, xyz(a), xyz(b), a.w, st, nv, nt, k);
    } else {
      const Geo G = geo_of(S);
      h = any ? traverse<true>(G, xyz(a), xyz(b), a.w, st, nv, nt) : traverse<false>(G, xyz(a), xyz(b), a.w, st, nv, nt);
    }
    tri[i] = h.tri;
    if (inst) inst[i] = k;
    if (t) { t[i] = h.t; u[i] = h.u; v[i] = h.v; }
  }
  stat_add(cnt, 4, nv);
  stat_add(cnt, 5, nt);
}

// --------------------------------------------------------------------------
// Display chain (renderer.ts:665-700), one thread per canvas pixel: the
// display resample (display.frag) picks a texel of the display buffer; FXAA
// (fxaa.frag) or the tonemap pass alone (tonemap.frag) produces it from the
// accumulation; the 8-bit canvas value and/or the float display value are
// written. ix, iy = fxaa.frag's u_inv_res.
__global__ __launch_bounds__(256) void k_postprocess(post::Source S, int fxaa_on, float ix, float iy, int ow, int oh,
                                                     uint32_t* __restrict__ out8, float4* __restrict__ out32) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)ow * oh) return;
  const int ox = (int)(i % ow), oy = (int)(i / ow);
  const int sx = post::nearest((post::vuv(ox, ow) + 1.0f) * 0.5f, S.W);
  const int sy = post::nearest((post::vuv(oy, oh) + 1.0f) * 0.5f, S.H);
  const float4 c = fxaa_on ? post::fxaa(S, sx, sy, ix, iy) : S.texel(sx, sy);
  if (out32) out32[i] = c;
  if (out8)
    out8[i] = post::unorm8(c.x) | (post::unorm8(c.y) << 8) | (post::unorm8(c.z) << 16) | (post::unorm8(c.w) << 24);
}

// --------------------------------------------------------------------------
// Self-test of the pinned transcendentals (pin_math.hpp) on the device: the
// same function table as the CPU oracle's float build.
__global__ __launch_bounds__(256) void k_pin_math(int fn, const float* __restrict__ x, const float* __restrict__ y,
                                                  float* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float r = 0.0f, s, c;
  switch (fn) {
    case 0: r = pm::psin(x[i]); break;
    case 1: r = pm::pcos(x[i]); break;
    case 2: r = pm::pexp(x[i]); break;
    case 3: r = pm::plog(x[i]); break;
    case 4: r = pm::ppow(x[i], y[i]); break;
    case 5: r = pm::patan2(x[i], y[i]); break;
    case 6: r = pm::pacos(x[i]); break;
    case 7: pm::psincos(x[i], s, c); r = s; break;
    case 8: pm::psincos(x[i], s, c); r = c; break;
    default: break;
  }
  out[i] = r;
}

// Plain device copy (16-byte loads and stores, grid-stride): the calibration
// kernel that relates the memory counters to bytes actually moved.
__global__ __launch_bounds__(256) void k_copy_calib(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}
