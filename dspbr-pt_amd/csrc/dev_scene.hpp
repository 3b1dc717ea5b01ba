// Device-side scene/IBL views and the traversal, IBL and shading-point
// routines shared by the wavefront kernels.
#pragma once
#include <climits>
#include "dev_bsdf.hpp"

// Device view of one uploaded scene + IBL. Layouts in HBM:
//   nodes  float4[4*N]  TinyBVH bvh_gpu nodes, uploaded unchanged
//   tri    float4[3*M]  per BVH index slot: (p0, tri id as int bits | -1), (p1), (p2)
//   verts  float4[5*V]  20-float vertex stream, unchanged
//   tris   u32[3*T], mats float4[17*K]
//   env    float4[w*h], pdf/cdf f32[w*h], ypdf/ycdf f32[h]
// One RGBA8 texture array: layers of w x h texels, row 0 first, packed
// R | G << 8 | B << 16 | A << 24 (scene_data_webgl2.ts:233-300 uploads the
// same bytes with texImage3D).
struct TexArrayDesc {
  int64_t offset;  // first texel in SceneView::texels
  int32_t w, h, layers, pad;
};

// One instance of a two-level scene, stored in SceneView::nodes as two
// 64-byte node slots: the world -> object rows and the root of its mesh's
// bottom-level BVH first (what the walk reads on entering the instance, one
// node-sized fetch), the object -> world rows after (read by shading).
struct InstanceDesc {
  float4 w2o[3];  // rows of the 3x4 affine
  int32_t root, mesh, pad0, pad1;
  float4 o2w[3];
  float4 pad2;
};
static_assert(sizeof(InstanceDesc) == 128, "an instance record is two node slots");

struct SceneView {
  const float4* __restrict__ nodes;
  const float4* __restrict__ tri;
  const float4* __restrict__ verts;
  const uint32_t* __restrict__ tris;
  const float4* __restrict__ mats;
  int32_t N, M, T, K;
  const float4* __restrict__ env;
  const float* __restrict__ pdf;
  const float* __restrict__ cdf;
  const float* __restrict__ ypdf;
  const float* __restrict__ ycdf;
  const int32_t* __restrict__ yguide;  // guide tables of ycdf (h+1) and of each cdf row (w+1 per row)
  const int32_t* __restrict__ guide;
  int32_t ew, eh;
  int32_t has_ibl;  // 0: no IBL uploaded; radiance and pdf lookups return 0
  int32_t full_profile;
  const float4* __restrict__ texinfo;       // 2 per texture (TexInfo, material.ts:483-523)
  const TexArrayDesc* __restrict__ texarr;
  const uint32_t* __restrict__ texels;
  int32_t num_tex, num_texarr;
  // two-level scenes (engine extension: the reference flattens instances).
  // `nodes` then holds the meshes' bottom-level BVHs, the top-level BVH over
  // the instances (root `top`) and the InstanceDesc records from node slot
  // `inst_base` on; a top-level child index >= inst_base is an instance
  // record (instance k at inst_base + 2k), so the top level has no leaf
  // nodes. -1: one level.
  int32_t top, inst_base;
  // 1 when every inner node's child boxes are finite (checked at upload):
  // then a ray with finite origin and finite, non-zero 1/dir can take the
  // slab test on hardware min/max (slab_fast)
  int32_t fast_slab;
  // ray order: bits of the origin-cell key per axis (x | y << 4 | z << 8),
  // 12 in all, set at upload from the root box (scene_key_bits)
  int32_t key_bits;
  const uint32_t* __restrict__ tri_mat;  // material index of each triangle (its first vertex's pos.w)
  // the three vertex records of each triangle (5 float4 each), copied at
  // upload so a hit reaches its vertices without the index load; NULL: go
  // through tris and verts
  const float4* __restrict__ tri_verts;
};

// The 52-float uniform block (renderer.frag:10-31).
struct FrameParams {
  float view[16];
  float bg[4];
  float cam[4];
  float inv_res[2];
  float ibl_res[2];
  float frame_count, debug_mode, pad1, use_ibl;
  float ibl_rotation, bg_from_ibl, max_bounces, focal_length;
  float ibl_total_sum, ray_eps, clamp_threshold, pad0;
  float scene_counts[4];
  float light_pos[4];
  float light_emission[4];
};

#ifndef DSPT_SLAB_SELECT
#define DSPT_SLAB_SELECT 1
#endif
#ifndef DSPT_TRI_SELECT
#define DSPT_TRI_SELECT 0
#endif
constexpr int kStackSize = 64;       // BVH_STACK_SIZE, SH/bvh.glsl:8
constexpr int kMaxSteps = 4096;      // MAX_BVH_TRAVERSAL_STEPS
constexpr int kMaxLeafTris = 64;     // MAX_BVH_LEAF_TRIANGLES
constexpr int kLdsStack = 16;        // stack entries held in LDS per lane
constexpr int kTraceBlock = 256;

struct RayHit {
  int tri;
  float t, u, v;
};

DI f3 xf_point(const float4* m, f3 p) {
  return f3{m[0].x * p.x + m[0].y * p.y + m[0].z * p.z + m[0].w, m[1].x * p.x + m[1].y * p.y + m[1].z * p.z + m[1].w,
            m[2].x * p.x + m[2].y * p.y + m[2].z * p.z + m[2].w};
}
DI f3 xf_vec(const float4* m, f3 v) {
  return f3{m[0].x * v.x + m[0].y * v.y + m[0].z * v.z, m[1].x * v.x + m[1].y * v.y + m[1].z * v.z,
            m[2].x * v.x + m[2].y * v.y + m[2].z * v.z};
}
// transpose(m) * v: normals with the inverse-transpose of the other direction
DI f3 xf_tvec(const float4* m, f3 v) {
  return f3{m[0].x * v.x + m[1].x * v.y + m[2].x * v.z, m[0].y * v.x + m[1].y * v.y + m[2].y * v.z,
            m[0].z * v.x + m[1].z * v.y + m[2].z * v.z};
}

#ifndef DSPT_SPILL_WAIT
#define DSPT_SPILL_WAIT 1
#endif
// Per-lane traversal stack: the first kLdsStack entries live in LDS
// (lane-interleaved, conflict-free), the rest in a global spill area that
// is private to the lane.
struct LaneStack {
  int* lds;        // &smem[threadIdx.x], stride kTraceBlock
  int* spill;      // per-lane global area of kStackSize - kLdsStack ints
  bool wait_spill = true;  // a constant per kernel (see get)
  DI void put(int i, int v) {
    if (i < kLdsStack) lds[i * kTraceBlock] = v;
    else spill[i - kLdsStack] = v;
  }
  // The spill read is a non-temporal global load so the compiler cannot
  // merge the two reads into one generic (FLAT) load of a selected pointer.
  // It is waited for inside its own branch: otherwise the popped value may
  // come from vector memory, and the common LDS pop would wait for every
  // outstanding vector memory operation (vmcnt counts stores too, e.g. the
  // hit records of rays that just finished) before the next node fetch.
  // The two-level shadow trace, whose walks spill more often, measured
  // slower with it and clears wait_spill.
  DI int get(int i) const {
    if (i < kLdsStack) return lds[i * kTraceBlock];
#if defined(__HIP_DEVICE_COMPILE__)
    const int v = __builtin_nontemporal_load((const __attribute__((address_space(1))) int*)(spill + (i - kLdsStack)));
    if (DSPT_SPILL_WAIT && wait_spill) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt not waited
    return v;
#else
    return spill[i - kLdsStack];
#endif
  }
};

// Slab test, SH/bvh.glsl:31-41 (GLSL min/max select forms, NaN guarded).
DI float slab(f3 bmin, f3 bmax, f3 org, f3 inv, float tfar) {
  f3 t0 = (bmin - org) * inv;
  f3 t1 = (bmax - org) * inv;
  f3 lo = vmin(t0, t1), hi = vmax(t0, t1);
  float tmin = gmax(gmax(lo.x, lo.y), lo.z);
  float tmax = gmin(gmin(hi.x, hi.y), hi.z);
#if DSPT_SLAB_SELECT
  // the reference's test `isnan(tmin) || isnan(tmax) || tmin > tmax ||
  // tmax < 0 || tmin >= tfar` as one predicate without short-circuit
  // branches: (tmin <= tmax) is false exactly when either is NaN or
  // tmin > tmax, and tfar is never NaN
  const bool ok = (tmin <= tmax) & (tmax >= 0.0f) & (tmin < tfar);
  return ok ? tmin : kBVH_FAR;
#else
  if (isnan(tmin) || isnan(tmax) || tmin > tmax || tmax < 0.0f || tmin >= tfar) return kBVH_FAR;
  return tmin;
#endif
}

// The slab test on hardware min/max (v_min_f32 / v_max_f32): one
// instruction per min/max instead of a compare, a select and a hazard wait.
// Equal to slab() whenever no t value can be NaN: with finite box corners,
// a finite origin and finite, non-zero 1/dir, every (b - o) * inv is finite
// or infinite, never 0 * inf. GLSL's select min/max and the hardware's then
// pick the same value up to the sign of a zero, which no comparison below
// (nor the near/far order of the caller) distinguishes.
DI float slab_fast(f3 bmin, f3 bmax, f3 org, f3 inv, float tfar) {
  const f3 t0 = (bmin - org) * inv;
  const f3 t1 = (bmax - org) * inv;
  const float tmin = fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z));
  const float tmax = fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z));
  const bool ok = (tmin <= tmax) & (tmax >= 0.0f) & (tmin < tfar);
  return ok ? tmin : kBVH_FAR;
}
DI bool finite_nz(float x) { return isfinite(x) && x != 0.0f; }
// Whether a ray may use slab_fast in a scene (see SceneView::fast_slab).
DI bool slab_fast_ok(int32_t scene_ok, f3 org, f3 inv) {
  return scene_ok != 0 && isfinite(org.x) && isfinite(org.y) && isfinite(org.z) && finite_nz(inv.x) && finite_nz(inv.y) &&
         finite_nz(inv.z);
}

// Moeller-Trumbore, SH/bvh.glsl:44-62.
DI bool tri_test(f3 org, f3 dir, f3 p0, f3 p1, f3 p2, float tfar, float& t, float& u, float& v) {
#if DSPT_TRI_SELECT
  // the same arithmetic and the same rejections without early exits: each
  // reference test `if (c) return false` becomes `!c` in one predicate, so
  // NaN comparisons behave as in the branching form
  f3 e0 = p1 - p0, e1 = p2 - p0;
  f3 pv = cross3(dir, e1);
  float det = dot3(e0, pv);
  float f = 1.0f / det;
  f3 s = org - p0;
  float uu = f * dot3(s, pv);
  f3 q = cross3(s, e0);
  float vv = f * dot3(dir, q);
  float tt = f * dot3(e1, q);
  const bool ok = !(fabsf(det) < kEPS) & !(uu < 0.0f) & !(uu > 1.0f) & !(vv < 0.0f) & !(uu + vv > 1.0f) & !(tt < kEPS) &
                  !(tt >= tfar);
  if (ok) { t = tt; u = uu; v = vv; }
  return ok;
#else
  f3 e0 = p1 - p0, e1 = p2 - p0;
  f3 pv = cross3(dir, e1);
  float det = dot3(e0, pv);
  if (fabsf(det) < kEPS) return false;
  float f = 1.0f / det;
  f3 s = org - p0;
  float uu = f * dot3(s, pv);
  if (uu < 0.0f || uu > 1.0f) return false;
  f3 q = cross3(s, e0);
  float vv = f * dot3(dir, q);
  if (vv < 0.0f || uu + vv > 1.0f) return false;
  float tt = f * dot3(e1, q);
  if (tt < kEPS || tt >= tfar) return false;
  t = tt; u = uu; v = vv;
  return true;
#endif
}

// Geometry a traversal reads, copied into registers once per ray.
struct Geo {
  const float4* __restrict__ nodes;
  const float4* __restrict__ tri;
  int32_t N, M;
};
DI Geo geo_of(const SceneView& S) { return Geo{S.nodes, S.tri, S.N, S.M}; }

// Geometry of a two-level traversal: every level's nodes and the instance
// records share one array, so a lane carries two more indices, not pointers.
struct Geo2 {
  Geo g;
  int32_t top, inst_base;
};
DI Geo2 geo2_of(const SceneView& S) { return Geo2{geo_of(S), S.top, S.inst_base}; }

// Child references inside the engine's copy of the node array. At upload
// each inner node's child fields (a.w, b.w) are rewritten from float-encoded
// node indices to int codes (k_encode_children): code >= 0 is the child's
// node index; a child that is a well-formed leaf becomes
// -(1 + first * 64 + min(count, 64) - 1), so the walk reaches its triangles
// without fetching the leaf node (whose box its parent already tested);
// kChildInvalid marks an index the reference would not follow (out of
// range or the node itself). Leaf nodes keep their float fields, so a leaf
// reached by index (the root, or one with an out-of-range range) is read as
// before. Visit order, step counts and the node/triangle counters are those
// of the original array.
constexpr int kChildInvalid = INT_MIN;
DI int child_code(float w) { return __float_as_int(w); }
DI bool leaf_code(int code) { return code < 0; }  // kChildInvalid is never pushed
DI void leaf_range(int code, int M, int& first, int& end) {
  const int v = -code - 1;
  first = v >> 6;
  end = min(first + (v & 63) + 1, M);
}

// Resumable traversal state of one ray (the loop variables of
// SH/bvh.glsl:149-212). The entry on top of the reference's stack is held in
// `next` when it was pushed by the previous step (has_next); the stack then
// logically holds top + has_next entries, so capacity, visit order and step
// count are exactly those of the reference loop.
struct RayState {
  f3 org, dir, inv;
  RayHit h;
  int top, step, next;
  int leaf_first, leaf_end;  // leaf reached by trav_inner, tested by trav_leaf
  bool has_next;
};

DI void ray_init(RayState& r, LaneStack& st, f3 org, f3 dir, float tfar) {
  (void)st;
  r.org = org;
  r.dir = dir;
  r.inv = f3{1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
  r.h = RayHit{-1, tfar, 0.0f, 0.0f};
  r.top = 0;
  r.next = 0;  // the root
  r.has_next = true;
  r.step = 0;
}

DI bool ray_done(const RayState& r) { return !(r.step < kMaxSteps && (r.top > 0 || r.has_next)); }

// One iteration of the stack traversal, in the exact visit order of
// SH/bvh.glsl:149-212: pop a node; a leaf tests its triangles, an inner
// node pushes the children it overlaps, far child first (ties: right child
// popped first). Returns true when the ray is finished. With ANY_HIT it
// finishes at the first accepted triangle; the occlusion boolean equals the
// reference's closest-hit based isOccluded() because both visit the same
// nodes with the same tfar until the first hit.
template <bool ANY_HIT>
DI bool trav_step(const Geo& G, RayState& r, LaneStack& st, unsigned& nv, unsigned& nt) {
  if (ray_done(r)) return true;
  r.step++;
  int ni;
  if (r.has_next) {
    ni = r.next;
    r.has_next = false;
  } else {
    ni = st.get(--r.top);
  }
  int first = 0, end = 0;
  bool leaf;
  float4 a, b, c, d;
  if (leaf_code(ni)) {
    leaf = true;
    leaf_range(ni, G.M, first, end);
  } else {
    if (ni >= G.N) return ray_done(r);
    const float4* nd = G.nodes + (size_t)ni * 4;
    a = ldg(nd); b = ldg(nd + 1); c = ldg(nd + 2); d = ldg(nd + 3);
    issued(a); issued(b); issued(c); issued(d);
    const int cnt = (int)c.w;
    leaf = cnt > 0;
    if (leaf) {
      first = (int)d.w;
      end = min(first + min(cnt, kMaxLeafTris), G.M);
    }
  }
  nv++;
  if (leaf) {
    for (int i = first; i < end; i++) {
      if (i < 0) continue;
      const float4* tp = G.tri + (size_t)i * 3;
      const float4 q0 = ldg(tp), q1 = ldg(tp + 1), q2 = ldg(tp + 2);
      issued(q0); issued(q1); issued(q2);
      const int id = __float_as_int(q0.w);
      if (id < 0) continue;
      nt++;
      float t, u, v;
      if (tri_test(r.org, r.dir, xyz(q0), xyz(q1), xyz(q2), r.h.t, t, u, v)) {
        r.h.t = t; r.h.tri = id; r.h.u = u; r.h.v = v;
        if (ANY_HIT) return true;
      }
    }
  } else {
    const int left = child_code(a.w), right = child_code(b.w);
    const float d0 = slab(xyz(a), xyz(b), r.org, r.inv, r.h.t);
    const float d1 = slab(xyz(c), xyz(d), r.org, r.inv, r.h.t);
    const bool h0 = d0 < kBVH_FAR, h1 = d1 < kBVH_FAR;
    const bool both = h0 && h1;
    const bool near_left = d0 < d1;
    // pushes in reference order: c1 first (the far child when both overlap,
    // else the only one), then c2 (the near child); the last push is the
    // next node visited, so it stays in a register
    const int c1 = both ? (near_left ? right : left) : (h0 ? left : right);
    const int c2 = near_left ? left : right;
    const bool p1 = (h0 || h1) && r.top < kStackSize && c1 != kChildInvalid;
    if (both) {
      if (p1) st.put(r.top++, c1);
      if (r.top < kStackSize && c2 != kChildInvalid) { r.next = c2; r.has_next = true; }
    } else if (p1) {
      r.next = c1;
      r.has_next = true;
    }
  }
  return ray_done(r);
}

// The same step split in two for the wavefront loop: trav_inner pops one
// node; an inner node pushes its children (as trav_step), a leaf is handed
// to trav_leaf, which tests its triangles before the ray pops again, so the
// per-ray sequence of node visits, triangle tests and steps is exactly that
// of trav_step. Returns 0 = continue, 1 = leaf pending, 2 = ray finished.
// The caller never steps a finished ray: a ray starts unfinished, and the
// wavefront loops retire it as soon as a step or a leaf test finishes it
// (ray_done after every leaf), so the step opens without that test.
template <bool FAST = false>
DI int trav_inner(const Geo& G, RayState& r, LaneStack& st, unsigned& nv) {
  r.step++;
  int ni;
  if (r.has_next) {
    ni = r.next;
    r.has_next = false;
  } else {
    ni = st.get(--r.top);
  }
  if (leaf_code(ni)) {  // a leaf child: its range is in the code, no fetch
    nv++;
    leaf_range(ni, G.M, r.leaf_first, r.leaf_end);
    return 1;
  }
  if (ni >= G.N) return ray_done(r) ? 2 : 0;
  nv++;
  const float4* nd = G.nodes + (size_t)ni * 4;
  const float4 a = ldg(nd), b = ldg(nd + 1), c = ldg(nd + 2), d = ldg(nd + 3);
  issued(a); issued(b); issued(c); issued(d);
  const int cnt = (int)c.w;
  if (cnt > 0) {
    const int first = (int)d.w;
    r.leaf_first = first;
    r.leaf_end = min(first + min(cnt, kMaxLeafTris), G.M);
    return 1;
  }
  const int left = child_code(a.w), right = child_code(b.w);
  const float d0 = FAST ? slab_fast(xyz(a), xyz(b), r.org, r.inv, r.h.t) : slab(xyz(a), xyz(b), r.org, r.inv, r.h.t);
  const float d1 = FAST ? slab_fast(xyz(c), xyz(d), r.org, r.inv, r.h.t) : slab(xyz(c), xyz(d), r.org, r.inv, r.h.t);
  const bool h0 = d0 < kBVH_FAR, h1 = d1 < kBVH_FAR;
  const bool both = h0 && h1;
  const bool near_left = d0 < d1;
  const int c1 = both ? (near_left ? right : left) : (h0 ? left : right);
  const int c2 = near_left ? left : right;
  const bool p1 = (h0 || h1) && r.top < kStackSize && c1 != kChildInvalid;
  if (both) {
    if (p1) st.put(r.top++, c1);
    if (r.top < kStackSize && c2 != kChildInvalid) { r.next = c2; r.has_next = true; }
  } else if (p1) {
    r.next = c1;
    r.has_next = true;
  }
  return ray_done(r) ? 2 : 0;
}

// Triangle tests of the pending leaf; with ANY_HIT returns true at the first
// accepted triangle.
template <bool ANY_HIT>
DI bool trav_leaf(const Geo& G, RayState& r, unsigned& nt) {
  for (int i = r.leaf_first; i < r.leaf_end; i++) {
    if (i < 0) continue;
    const float4* tp = G.tri + (size_t)i * 3;
    const float4 q0 = ldg(tp), q1 = ldg(tp + 1), q2 = ldg(tp + 2);
    issued(q0); issued(q1); issued(q2);
    const int id = __float_as_int(q0.w);
    if (id < 0) continue;
    nt++;
    float t, u, v;
    if (tri_test(r.org, r.dir, xyz(q0), xyz(q1), xyz(q2), r.h.t, t, u, v)) {
      r.h.t = t; r.h.tri = id; r.h.u = u; r.h.v = v;
      if (ANY_HIT) return true;
    }
  }
  return false;
}

// Two-level traversal (scenes with instances). The top-level BVH is walked
// like a mesh BVH; at a top-level leaf the ray is moved into the instance's
// object space (the direction is not renormalised, so t keeps its meaning)
// and the instance's mesh BVH is walked on the same stack above `base`;
// when the stack is back at `base` the world ray is restored. Leaf
// triangles are tested by trav_leaf as for one level; `inst` records the
// instance of the closest hit.
struct InstState {
  f3 worg, wdir;       // the world ray while inside an instance
  int base, cur, hit;  // base >= 0: inside instance `cur`, whose walk sits above stack depth base; hit = instance of the closest hit
};

DI void inst_init(InstState& q) {
  q.base = -1; q.cur = -1; q.hit = -1;
  q.worg = q.wdir = f3{0.0f, 0.0f, 0.0f};
}

DI int trav_inner2(const Geo2& G, RayState& r, InstState& q, LaneStack& st, unsigned& nv) {
  if (q.base >= 0 && !r.has_next && r.top == q.base) {
    // back at the top level: the world ray (its inverse recomputed exactly as ray_init does)
    q.base = -1;
    r.org = q.worg; r.dir = q.wdir;
    r.inv = f3{1.0f / r.dir.x, 1.0f / r.dir.y, 1.0f / r.dir.z};
  }
  if (ray_done(r)) return 2;
  r.step++;
  int ni;
  if (r.has_next) {
    ni = r.next;
    r.has_next = false;
  } else {
    ni = st.get(--r.top);
  }
  const int N = G.g.N;
  if (leaf_code(ni)) {  // a bottom-level leaf child (the top level has none)
    nv++;
    if (q.base < 0) return ray_done(r) ? 2 : 0;
    leaf_range(ni, G.g.M, r.leaf_first, r.leaf_end);
    return 1;
  }
  if (ni >= N) return ray_done(r) ? 2 : 0;
  nv++;
  if (q.base < 0 && ni >= G.inst_base) {
    // a top-level child that is an instance record (its world box was the
    // parent's slab test): enter the instance
    if (ni + 1 >= N) return ray_done(r) ? 2 : 0;
    const float4* D = G.g.nodes + (size_t)ni * 4;
    const float4 w2o[3] = {ldg(D), ldg(D + 1), ldg(D + 2)};
    const float4 meta = ldg(D + 3);
    q.worg = r.org; q.wdir = r.dir;
    q.base = r.top;
    q.cur = (ni - G.inst_base) >> 1;
    r.org = xf_point(w2o, q.worg);
    r.dir = xf_vec(w2o, q.wdir);
    r.inv = f3{1.0f / r.dir.x, 1.0f / r.dir.y, 1.0f / r.dir.z};
    r.next = __float_as_int(meta.x);
    r.has_next = true;
    return 0;
  }
  const float4* nd = G.g.nodes + (size_t)ni * 4;
  const float4 a = ldg(nd), b = ldg(nd + 1), c = ldg(nd + 2), d = ldg(nd + 3);
  issued(a); issued(b); issued(c); issued(d);
  const int cnt = (int)c.w;
  if (cnt > 0) {
    if (q.base < 0) return ray_done(r) ? 2 : 0;  // the top level has no leaves
    const int first = (int)d.w;
    r.leaf_first = first;
    r.leaf_end = min(first + min(cnt, kMaxLeafTris), G.g.M);
    return 1;
  }
  const int left = child_code(a.w), right = child_code(b.w);
  const float d0 = slab(xyz(a), xyz(b), r.org, r.inv, r.h.t);
  const float d1 = slab(xyz(c), xyz(d), r.org, r.inv, r.h.t);
  const bool h0 = d0 < kBVH_FAR, h1 = d1 < kBVH_FAR;
  const bool both = h0 && h1;
  const bool near_left = d0 < d1;
  const int c1 = both ? (near_left ? right : left) : (h0 ? left : right);
  const int c2 = near_left ? left : right;
  const bool p1 = (h0 || h1) && r.top < kStackSize && c1 != kChildInvalid;
  if (both) {
    if (p1) st.put(r.top++, c1);
    if (r.top < kStackSize && c2 != kChildInvalid) { r.next = c2; r.has_next = true; }
  } else if (p1) {
    r.next = c1;
    r.has_next = true;
  }
  return ray_done(r) ? 2 : 0;
}

// trav_leaf for the two-level walk: a closer hit also records the instance.
template <bool ANY_HIT>
DI bool trav_leaf2(const Geo2& G, RayState& r, InstState& q, unsigned& nt) {
  const float t0 = r.h.t;
  const bool hit = trav_leaf<ANY_HIT>(G.g, r, nt);
  if (r.h.t != t0 || hit) q.hit = q.cur;
  return hit;
}

// Whole two-level traversal of one ray (standalone probes).
template <bool ANY_HIT>
DI RayHit traverse2(const Geo2& G, f3 org, f3 dir, float tfar, LaneStack& st, unsigned& nv, unsigned& nt, int& inst) {
  RayState r;
  InstState q;
  ray_init(r, st, org, dir, tfar);
  r.next = G.top;
  inst_init(q);
  for (;;) {
    const int res = trav_inner2(G, r, q, st, nv);
    if (res == 2) break;
    if (res == 1) {
      if (trav_leaf2<ANY_HIT>(G, r, q, nt)) break;
    }
  }
  inst = q.hit;
  return r.h;
}

// Whole traversal of one ray.
template <bool ANY_HIT>
DI RayHit traverse(const Geo G, f3 org, f3 dir, float tfar, LaneStack& st, unsigned& nv, unsigned& nt) {
  RayState r;
  ray_init(r, st, org, dir, tfar);
  while (!trav_step<ANY_HIT>(G, r, st, nv, nt)) {
  }
  return r.h;
}

// ---------------------------------------------------------------- IBL
// lower_bound, SH/utils.glsl:196-217
DI int lower_bound_f(const float* __restrict__ row, int size, float value) {
  int count = size, first = 0;
  while (count > 0) {
    int step = count / 2;
    int idx = first + step;
    if (row[idx] < value) { first = idx + 1; count -= step + 1; }
    else count = step;
  }
  return first;
}
// lower_bound over [lo, hi) only: equals lower_bound_f over the whole
// non-decreasing row whenever the answer lies in [lo, hi].
DI int lower_bound_range(const float* __restrict__ row, int lo, int hi, float value) {
  int count = hi - lo, first = lo;
  while (count > 0) {
    int step = count / 2;
    int idx = first + step;
    if (row[idx] < value) { first = idx + 1; count -= step + 1; }
    else count = step;
  }
  return first;
}
// Guide table of a non-decreasing CDF row of n entries: guide[g] =
// lower_bound(row, g / n) for g = 0..n. For a value in bucket g the answer
// lies between guide[g - 2] and guide[g + 3] (two buckets of margin cover
// any rounding of value * n), so the search runs over a few entries and
// returns exactly lower_bound_f(row, n, value).
DI int lower_bound_guided(const float* __restrict__ row, const int32_t* __restrict__ guide, int n, float value) {
  int g = (int)(value * (float)n);
  g = g < 0 ? 0 : (g >= n ? n - 1 : g);
  const int lo = guide[g >= 2 ? g - 2 : 0];
  const int hi = guide[g + 3 <= n ? g + 3 : n];
  return lower_bound_range(row, lo, hi, value);
}
DI f3 rotate_phi(f3 d, float rot, bool inverse) {
  float a = inverse ? -rot : rot;
  float c, s;
  pm::psincos(a, s, c);
  return f3{c * d.x + 0.0f * d.y + (-s) * d.z, 0.0f * d.x + 1.0f * d.y + 0.0f * d.z, s * d.x + 0.0f * d.y + c * d.z};
}
DI float sin_pdf(float theta) {
  return 1.0f / (2.0f * kPI * kPI * gmax(kEPS_COS, pm::psin(theta)));
}
// dir_to_uv, SH/utils.glsl:141-147
DI float2 dir_uv(f3 d, float& theta) {
  theta = pm::pacos(gmax(-1.0f, gmin(1.0f, d.y)));
  float ph = pm::patan2(d.z, d.x);
  if (ph < 0.0f) ph = kTWO_PI + ph;
  return make_float2(ph / kTWO_PI, theta / kPI);
}
// texture(u_sampler_env_map, uv): GL LINEAR + REPEAT on RGBA32F
DI f3 env_lookup(const SceneView& S, float2 uv) {
  int w = S.ew, h = S.eh;
  float x = uv.x * float(w) - 0.5f, y = uv.y * float(h) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = (int)fx, iy = (int)fy;
  int x0 = ix % w; if (x0 < 0) x0 += w;
  int x1 = (ix + 1) % w; if (x1 < 0) x1 += w;
  int y0 = iy % h; if (y0 < 0) y0 += h;
  int y1 = (iy + 1) % h; if (y1 < 0) y1 += h;
  f3 t00 = xyz(S.env[(size_t)y0 * w + x0]), t10 = xyz(S.env[(size_t)y0 * w + x1]);
  f3 t01 = xyz(S.env[(size_t)y1 * w + x0]), t11 = xyz(S.env[(size_t)y1 * w + x1]);
  f3 top = t00 * (1.0f - a) + t10 * a;
  f3 bot = t01 * (1.0f - a) + t11 * a;
  return top * (1.0f - b) + bot * b;
}
// texture() of a one-channel R32F importance table (renderer.ts:858-859:
// the pdf and cdf textures, LINEAR + REPEAT by gl_utils.ts:203-216): the red
// channel of the bilinear fetch, with the arithmetic of env_lookup.
DI float table_lookup(const float* __restrict__ t, int w, int h, float2 uv) {
  float x = uv.x * float(w) - 0.5f, y = uv.y * float(h) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  float a = x - fx, b = y - fy;
  int ix = (int)fx, iy = (int)fy;
  int x0 = ix % w; if (x0 < 0) x0 += w;
  int x1 = (ix + 1) % w; if (x1 < 0) x1 += w;
  int y0 = iy % h; if (y0 < 0) y0 += h;
  int y1 = (iy + 1) % h; if (y1 < 0) y1 += h;
  float top = t[(size_t)y0 * w + x0] * (1.0f - a) + t[(size_t)y0 * w + x1] * a;
  float bot = t[(size_t)y1 * w + x0] * (1.0f - a) + t[(size_t)y1 * w + x1] * a;
  return top * (1.0f - b) + bot * b;
}
// evaluateMaterialTextureValue (scene_data_webgl2.ts:325-340):
// texture(array, vec3(uv * scale + offset, layer)) with GL LINEAR + REPEAT on
// RGBA8 UNORM (texel / 255, bilinear in f32, as the environment lookup). The
// layer is round-to-nearest and clamped; an array index out of range gives
// vec4(1) like the generated if-chain. uv set 1 is never written by the
// reference's render state (renderer.frag:72-91), taken as (0, 0).
DI float unorm8(uint32_t v, int k) { return float((v >> (8 * k)) & 0xFFu) / 255.0f; }
DI float4 unorm8x4(uint32_t v) { return make_float4(unorm8(v, 0), unorm8(v, 1), unorm8(v, 2), unorm8(v, 3)); }
DI float4 tex_value(const SceneView& S, float id, float2 uv0) {
  const int idx = (int)id;
  const float4 t0 = ldg(S.texinfo + 2 * (size_t)idx), t1 = ldg(S.texinfo + 2 * (size_t)idx + 1);
  const int ai = (int)t0.z;
  if (ai < 0 || ai >= S.num_texarr) return make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  const float2 uv = t1.z > 0.5f ? make_float2(0.0f, 0.0f) : uv0;
  const float tu = uv.x * t1.x + t0.x, tv = uv.y * t1.y + t0.y;
  const TexArrayDesc A = S.texarr[ai];
  const int w = A.w, h = A.h;
  float lf = floorf(t0.w + 0.5f);
  lf = gmin(gmax(lf, 0.0f), float(A.layers - 1));
  const int layer = (int)lf;
  const float x = tu * float(w) - 0.5f, y = tv * float(h) - 0.5f;
  float fx = floorf(x), fy = floorf(y);
  const float a = x - fx, b = y - fy;
  fx = gmin(gmax(fx, -1.0e9f), 1.0e9f);  // keeps the integer texel index in range
  fy = gmin(gmax(fy, -1.0e9f), 1.0e9f);
  const int ix = (int)fx, iy = (int)fy;
  int x0 = ix % w; if (x0 < 0) x0 += w;
  int x1 = (ix + 1) % w; if (x1 < 0) x1 += w;
  int y0 = iy % h; if (y0 < 0) y0 += h;
  int y1 = (iy + 1) % h; if (y1 < 0) y1 += h;
  const uint32_t* L = S.texels + A.offset + (int64_t)layer * w * h;
  const float4 t00 = unorm8x4(L[(int64_t)y0 * w + x0]), t10 = unorm8x4(L[(int64_t)y0 * w + x1]);
  const float4 t01 = unorm8x4(L[(int64_t)y1 * w + x0]), t11 = unorm8x4(L[(int64_t)y1 * w + x1]);
  const float ia = 1.0f - a, ib = 1.0f - b;
  const float4 top = make_float4(t00.x * ia + t10.x * a, t00.y * ia + t10.y * a, t00.z * ia + t10.z * a, t00.w * ia + t10.w * a);
  const float4 bot = make_float4(t01.x * ia + t11.x * a, t01.y * ia + t11.y * a, t01.z * ia + t11.z * a, t01.w * ia + t11.w * a);
  return make_float4(top.x * ib + bot.x * b, top.y * ib + bot.y * b, top.z * ib + bot.z * b, top.w * ib + bot.w * b);
}
DI f3 ibl_eval(const SceneView& S, const FrameParams& F, f3 d) {
  if (!S.has_ibl) return sp3(0.0f);
  float th;
  float2 uv = dir_uv(rotate_phi(d, F.ibl_rotation, false), th);
  return env_lookup(S, uv);
}
DI float ibl_pdf(const SceneView& S, const FrameParams& F, f3 d) {
  if (!S.has_ibl || F.ibl_res[0] <= 0.0f || F.ibl_res[1] <= 0.0f) return 0.0f;
  float th;
  float2 uv = dir_uv(rotate_phi(d, F.ibl_rotation, false), th);
  float pw = sin_pdf(th);
  float w = F.ibl_res[0], h = F.ibl_res[1];
  float x = gmin(uv.x * w, w - 1.0f);
  float y = gmin(uv.y * h, h - 1.0f);
  int ix = min(max((int)x, 0), S.ew - 1), iy = min(max((int)y, 0), S.eh - 1);  // memory guard only
  return w * h * S.ypdf[iy] * S.pdf[(size_t)iy * S.ew + ix] * pw;
}
DI f3 ibl_sample(const SceneView& S, const FrameParams& F, float r0, float r1, float& pdf) {
  int w = (int)F.ibl_res[0], h = (int)F.ibl_res[1];
  if (!S.has_ibl || w <= 0 || h <= 0) { pdf = 0.0f; return f3{0.0f, 1.0f, 0.0f}; }
  int y = lower_bound_guided(S.ycdf, S.yguide, h, r1);
  float py = y < h ? S.ypdf[y] : 0.0f;
  y = y >= h ? h - 1 : y;
  const float* crow = S.cdf + (size_t)y * w;
  int x = lower_bound_guided(crow, S.guide + (size_t)y * (w + 1), w, r0);
  float px = x < w ? S.pdf[(size_t)y * w + x] : 0.0f;
  x = x >= w ? w - 1 : x;
  float spdf = F.ibl_res[0] * F.ibl_res[1] * py * px;
  float u = float(x) / F.ibl_res[0], v = float(y) / F.ibl_res[1];
  float theta = v * kPI, phi = u * kTWO_PI;
  float st, ct, sp, cp;
  pm::psincos(theta, st, ct);
  pm::psincos(phi, sp, cp);
  float pw = 1.0f / (2.0f * kPI * kPI * gmax(kEPS_COS, st));  // sin_pdf(theta)
  f3 d = f3{st * cp, ct, st * sp};
  pdf = spdf * pw;
  return rotate_phi(d, F.ibl_rotation, true);
}
