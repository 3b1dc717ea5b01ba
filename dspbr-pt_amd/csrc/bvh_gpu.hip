// GPU BVH build in the bvh_gpu node layout (the format the traversal reads,
// packages/lib/shader/bvh.glsl:1-6): a linear BVH (Karras 2012, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees") over
// 63-bit Morton codes of the triangle centroids, with every subtree of at
// most `max_leaf` triangles collapsed into one leaf. It replaces the host
// build (renderer.ts:872-912 -> tinybvh_builder.ts:142-215) where build time
// matters more than tree quality, e.g. admitting new scenes while serving.
//
// Passes (all O(n), HBM-bound, one thread per triangle or node):
//   k_bvh_boxes   triangle boxes + centroids; centroid bounds by wave
//                 reduction + ordered-integer atomics
//   k_bvh_morton  63-bit keys (21 bits per axis), value = triangle id
//   sort          radix sort of (key, id) pairs (rocPRIM via hipCUB)
//   k_bvh_karras  internal node i: its key range and split (longest common
//                 prefix, ties broken by position), children, parents
//   k_bvh_refit   bottom-up boxes: the second child to arrive at a parent
//                 (atomic counter) unions the two boxes and climbs on
//   k_bvh_flags   which Karras nodes become output nodes: inner nodes over
//                 more than max_leaf triangles, and the largest subtrees of
//                 at most max_leaf triangles below them (leaves)
//   scan          exclusive sum of the flags = output slots (root -> 0)
//   k_bvh_emit    16 floats per output node; leaf = contiguous range of the
//                 sorted ids, which are the BVH index array
// Karras node ids: internal 0 .. n-2 (root 0), leaves n-1 .. 2n-2 (leaf
// n-1+k = sorted position k). Included by dspt_host.hip inside its
// anonymous namespace (after <hipcub/hipcub.hpp>, Allocs and grid_for).

namespace gbvh {

DD unsigned ordered(float f) {  // float -> unsigned with the same order
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
DD float unordered(unsigned u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u); }

DD unsigned long long spread21(unsigned long long v) {  // 21 bits -> every third bit
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

}  // namespace gbvh

// triangle boxes (lo.xyz, hi.xyz) and centroid bounds (6 ordered ints: min xyz, max xyz)
__global__ __launch_bounds__(256) void k_bvh_boxes(const float4* __restrict__ verts, const uint32_t* __restrict__ tris,
                                                   int32_t n, float4* __restrict__ tlo, float4* __restrict__ thi,
                                                   unsigned* __restrict__ cb) {
  const int32_t t = blockIdx.x * 256 + threadIdx.x;
  float c[3] = {INFINITY, INFINITY, INFINITY}, d[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (t < n) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < 3; k++) {
      const float4 p = verts[(size_t)tris[(size_t)t * 3 + k] * 5];
      lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
      hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
    }
    tlo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    thi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    for (int k = 0; k < 3; k++) c[k] = d[k] = 0.5f * (lo[k] + hi[k]);
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int k = 0; k < 3; k++) {
      c[k] = fminf(c[k], __shfl_xor(c[k], off));
      d[k] = fmaxf(d[k], __shfl_xor(d[k], off));
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; k++) {
      atomicMin(&cb[k], gbvh::ordered(c[k]));
      atomicMax(&cb[3 + k], gbvh::ordered(d[k]));
    }
}

__global__ __launch_bounds__(256) void k_bvh_morton(const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                    const unsigned* __restrict__ cb, int32_t n,
                                                    unsigned long long* __restrict__ keys, int32_t* __restrict__ vals) {
  const int32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const float4 a = tlo[t], b = thi[t];
  const float c[3] = {0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z)};
  unsigned long long key = 0;
  for (int k = 0; k < 3; k++) {
    const float mn = gbvh::unordered(cb[k]), mx = gbvh::unordered(cb[3 + k]);
    const float ext = mx - mn;
    float u = ext > 0.0f ? (c[k] - mn) / ext : 0.0f;
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    const unsigned q = (unsigned)fminf(u * 2097152.0f, 2097151.0f);
    key |= gbvh::spread21(q) << (2 - k);
  }
  keys[t] = key;
  vals[t] = t;
}

namespace gbvh {
DD float4 ld3(const float4* p) {
  const float* f = reinterpret_cast<const float*>(p);
  return make_float4(__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(f + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0.0f);
}
// longest common prefix of sorted keys i and j (-1 outside [0, n)), ties by position
DD int delta(const unsigned long long* __restrict__ keys, int32_t n, int32_t i, int32_t j) {
  if (j < 0 || j >= n) return -1;
  const unsigned long long a = keys[i], b = keys[j];
  if (a == b) return 64 + __clz((unsigned)(i ^ j));
  return __clzll(a ^ b);
}
}  // namespace gbvh

__global__ __launch_bounds__(256) void k_bvh_karras(const unsigned long long* __restrict__ keys, int32_t n,
                                                    int32_t* __restrict__ parent, int2* __restrict__ range,
                                                    int2* __restrict__ child) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (gbvh::delta(keys, n, i, i + 1) - gbvh::delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = gbvh::delta(keys, n, i, i - d);
  int lmax = 2;
  while (gbvh::delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (gbvh::delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = gbvh::delta(keys, n, i, j);
  int s = 0;
  for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
    if (gbvh::delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    if (t == 1) break;
  }
  const int split = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const int cl = (lo == split) ? (n - 1 + split) : split;
  const int cr = (hi == split + 1) ? (n - 1 + split + 1) : split + 1;
  range[i] = make_int2(lo, hi);
  child[i] = make_int2(cl, cr);
  parent[cl] = i;
  parent[cr] = i;
  if (i == 0) parent[0] = -1;
}

__global__ __launch_bounds__(256) void k_bvh_refit(const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                   const int32_t* __restrict__ sorted, int32_t n,
                                                   const int32_t* __restrict__ parent, const int2* __restrict__ child,
                                                   float4* blo, float4* bhi, int32_t* __restrict__ arrivals) {
  // blo / bhi are read and written by many lanes: no __restrict__
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  int node = n - 1 + k;
  const int t = sorted[k];
  blo[node] = tlo[t];
  bhi[node] = thi[t];
  for (;;) {
    const int p = parent[node];
    if (p < 0) break;
    __threadfence();
    if (atomicAdd(&arrivals[p], 1) == 0) break;  // the sibling is not done yet: it continues
    __threadfence();
    const int2 c = child[p];
    // the sibling's box was written by another lane (maybe on another CU)
    // before its atomic: read both boxes with device-scope atomic loads,
    // which bypass the non-coherent per-CU cache
    const float4 a0 = gbvh::ld3(blo + c.x), a1 = gbvh::ld3(bhi + c.x);
    const float4 b0 = gbvh::ld3(blo + c.y), b1 = gbvh::ld3(bhi + c.y);
    blo[p] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), 0.0f);
    bhi[p] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), 0.0f);
    node = p;
  }
}

namespace gbvh {
DD int size_of(int id, int32_t n, const int2* __restrict__ range) {
  if (id >= n - 1) return 1;
  const int2 r = range[id];
  return r.y - r.x + 1;
}
}  // namespace gbvh

// flag[id] = 1 when Karras node id becomes an output node
__global__ __launch_bounds__(256) void k_bvh_flags(int32_t n, int32_t max_leaf, const int32_t* __restrict__ parent,
                                                   const int2* __restrict__ range, int32_t* __restrict__ flag) {
  const int32_t id = blockIdx.x * 256 + threadIdx.x;
  if (id >= 2 * n - 1) return;
  const int p = n == 1 ? -1 : parent[id];
  const bool parent_inner = p < 0 || gbvh::size_of(p, n, range) > max_leaf;
  flag[id] = parent_inner ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_bvh_emit(int32_t n, int32_t max_leaf, const int2* __restrict__ range,
                                                  const int2* __restrict__ child, const int32_t* __restrict__ flag,
                                                  const int32_t* __restrict__ slot, const float4* __restrict__ blo,
                                                  const float4* __restrict__ bhi, float4* __restrict__ out) {
  const int32_t id = blockIdx.x * 256 + threadIdx.x;
  if (id >= 2 * n - 1 || !flag[id]) return;
  float4* o = out + (size_t)slot[id] * 4;
  const int sz = gbvh::size_of(id, n, range);
  if (id < n - 1 && sz > max_leaf) {
    const int2 c = child[id];
    const float4 l0 = blo[c.x], l1 = bhi[c.x], r0 = blo[c.y], r1 = bhi[c.y];
    o[0] = make_float4(l0.x, l0.y, l0.z, (float)slot[c.x]);
    o[1] = make_float4(l1.x, l1.y, l1.z, (float)slot[c.y]);
    o[2] = make_float4(r0.x, r0.y, r0.z, 0.0f);
    o[3] = make_float4(r1.x, r1.y, r1.z, 0.0f);
  } else {
    const int first = id < n - 1 ? range[id].x : id - (n - 1);
    const float4 b0 = blo[id], b1 = bhi[id];
    o[0] = make_float4(b0.x, b0.y, b0.z, 0.0f);
    o[1] = make_float4(b1.x, b1.y, b1.z, 0.0f);
    o[2] = make_float4(b0.x, b0.y, b0.z, (float)sz);
    o[3] = make_float4(b1.x, b1.y, b1.z, (float)first);
  }
}

// ---------------------------------------------------------------- PLOC
// Parallel locally-ordered clustering (Meister & Bittner 2018, "Parallel
// Locally-Ordered Clustering for Bounding Volume Hierarchy Construction"):
// clusters start as the Morton-sorted triangles; each round every cluster
// finds the neighbour within kPlocRadius positions whose merged box has the
// least surface area, mutual nearest neighbours merge into a new node (the
// lower position keeps the merged cluster), and the survivors are compacted
// in order. New node ids come from a scan over the merging positions, so the
// build is deterministic. Node ids: leaves 0 .. n-1 (= sorted positions),
// internal n .. 2n-2 in creation order (root last).
constexpr int kPlocRadius = 16;

namespace gbvh {
DD float half_area(float4 lo, float4 hi) {
  const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
  return dx * dy + dy * dz + dz * dx;
}
DD float4 vmin4(float4 a, float4 b) { return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f); }
DD float4 vmax4(float4 a, float4 b) { return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f); }
}  // namespace gbvh

__global__ __launch_bounds__(256) void k_ploc_init(const float4* __restrict__ tlo, const float4* __restrict__ thi,
                                                   const int32_t* __restrict__ sorted, int32_t n, float4* __restrict__ nlo,
                                                   float4* __restrict__ nhi, int32_t* __restrict__ nsize,
                                                   int32_t* __restrict__ cnode, float4* __restrict__ clo,
                                                   float4* __restrict__ chi) {
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int t = sorted[k];
  const float4 a = tlo[t], b = thi[t];
  nlo[k] = a; nhi[k] = b; nsize[k] = 1;
  cnode[k] = k; clo[k] = a; chi[k] = b;
}

// nearest neighbour (least merged half-area; ties -> the lower position) in [i-R, i+R]
__global__ __launch_bounds__(256) void k_ploc_nn(const float4* __restrict__ clo, const float4* __restrict__ chi, int32_t c,
                                                 int32_t* __restrict__ nn) {
  __shared__ float4 slo[256 + 2 * kPlocRadius], shi[256 + 2 * kPlocRadius];
  const int32_t base = blockIdx.x * 256 - kPlocRadius;
  for (int k = threadIdx.x; k < 256 + 2 * kPlocRadius; k += 256) {
    const int32_t g = base + k;
    if (g >= 0 && g < c) { slo[k] = clo[g]; shi[k] = chi[g]; }
  }
  __syncthreads();
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= c) return;
  const int li = i - base;
  const float4 a0 = slo[li], a1 = shi[li];
  float best = INFINITY;
  int bj = -1;
  const int j0 = max(0, i - kPlocRadius), j1 = min(c - 1, i + kPlocRadius);
  for (int j = j0; j <= j1; j++) {
    if (j == i) continue;
    const int lj = j - base;
    const float sa = gbvh::half_area(gbvh::vmin4(a0, slo[lj]), gbvh::vmax4(a1, shi[lj]));
    if (sa < best) { best = sa; bj = j; }
  }
  nn[i] = bj;
}

// merge[i] = 1 if position i starts a mutual pair (i < nn[i]); keep[i] = 0 if i is absorbed
__global__ __launch_bounds__(256) void k_ploc_flags(const int32_t* __restrict__ nn, int32_t c, int32_t* __restrict__ merge,
                                                    int32_t* __restrict__ keep) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= c) return;
  const int j = nn[i];
  const bool mutual = j >= 0 && nn[j] == i;
  merge[i] = (mutual && i < j) ? 1 : 0;
  keep[i] = (mutual && i > j) ? 0 : 1;
}

__global__ __launch_bounds__(256) void k_ploc_merge(const int32_t* __restrict__ nn, int32_t c, int32_t n, int32_t next_id,
                                                    int32_t max_leaf, int32_t* __restrict__ esize,
                                                    const int32_t* __restrict__ merge, const int32_t* __restrict__ mscan,
                                                    const int32_t* __restrict__ keep, const int32_t* __restrict__ kscan,
                                                    const int32_t* __restrict__ cnode, const float4* __restrict__ clo,
                                                    const float4* __restrict__ chi, int32_t* __restrict__ cnode2,
                                                    float4* __restrict__ clo2, float4* __restrict__ chi2,
                                                    float4* __restrict__ nlo, float4* __restrict__ nhi,
                                                    int2* __restrict__ child, int32_t* __restrict__ parent,
                                                    int32_t* __restrict__ nsize) {
  const int32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= c || !keep[i]) return;
  int node = cnode[i];
  float4 lo = clo[i], hi = chi[i];
  if (merge[i]) {
    const int j = nn[i];
    const int a = node, b = cnode[j];
    node = next_id + mscan[i];
    lo = gbvh::vmin4(lo, clo[j]);
    hi = gbvh::vmax4(hi, chi[j]);
    nlo[node] = lo; nhi[node] = hi;
    child[node - n] = make_int2(a, b);
    parent[a] = node; parent[b] = node;
    const int sz = nsize[a] + nsize[b];
    nsize[node] = sz;
    // output nodes in this subtree: a leaf when collapsed, else itself + its children's
    esize[node] = sz > max_leaf ? 1 + (nsize[a] > max_leaf ? esize[a] : 1) + (nsize[b] > max_leaf ? esize[b] : 1) : 1;
  }
  const int o = kscan[i];
  cnode2[o] = node; clo2[o] = lo; chi2[o] = hi;
}

// first primitive position of every node's subtree (the sum of the sizes of
// the left siblings on the way to the root), whether it is an output node,
// and its output slot = its preorder position among the output nodes
// (1 + the output nodes of the left sibling's subtree for every step up),
// so a parent's children follow it in memory as in a depth-first build
__global__ __launch_bounds__(256) void k_ploc_layout(int32_t n, int32_t max_leaf, const int32_t* __restrict__ parent,
                                                     const int2* __restrict__ child, const int32_t* __restrict__ nsize,
                                                     const int32_t* __restrict__ esize, int32_t* __restrict__ offset,
                                                     int32_t* __restrict__ flag, int32_t* __restrict__ slot) {
  const int32_t id = blockIdx.x * 256 + threadIdx.x;
  const int32_t m = 2 * n - 1;
  if (id >= m) return;
  int off = 0, pre = 0, node = id;
  for (;;) {
    const int p = parent[node];
    if (p < 0) break;
    const int2 c = child[p - n];
    pre += 1;
    if (c.y == node) {
      off += nsize[c.x];
      pre += nsize[c.x] > max_leaf ? esize[c.x] : 1;
    }
    node = p;
  }
  offset[id] = off;
  const int p = parent[id];
  flag[id] = (p < 0 || nsize[p] > max_leaf) ? 1 : 0;
  slot[id] = pre;
}

__global__ __launch_bounds__(256) void k_ploc_order(int32_t n, const int32_t* __restrict__ offset,
                                                    const int32_t* __restrict__ sorted, int32_t* __restrict__ bvh_idx) {
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  bvh_idx[offset[k]] = sorted[k];  // leaf k = sorted position k
}

__global__ __launch_bounds__(256) void k_ploc_emit(int32_t n, int32_t max_leaf, const int2* __restrict__ child,
                                                   const int32_t* __restrict__ nsize, const int32_t* __restrict__ offset,
                                                   const int32_t* __restrict__ flag, const int32_t* __restrict__ slot,
                                                   const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                   float4* __restrict__ out) {
  const int32_t id = blockIdx.x * 256 + threadIdx.x;
  const int32_t m = 2 * n - 1;
  if (id >= m || !flag[id]) return;
  float4* o = out + (size_t)slot[id] * 4;
  const int sz = nsize[id];
  if (id >= n && sz > max_leaf) {
    const int2 c = child[id - n];
    const float4 l0 = nlo[c.x], l1 = nhi[c.x], r0 = nlo[c.y], r1 = nhi[c.y];
    o[0] = make_float4(l0.x, l0.y, l0.z, (float)slot[c.x]);
    o[1] = make_float4(l1.x, l1.y, l1.z, (float)slot[c.y]);
    o[2] = make_float4(r0.x, r0.y, r0.z, 0.0f);
    o[3] = make_float4(r1.x, r1.y, r1.z, 0.0f);
  } else {
    const float4 b0 = nlo[id], b1 = nhi[id];
    o[0] = make_float4(b0.x, b0.y, b0.z, 0.0f);
    o[1] = make_float4(b1.x, b1.y, b1.z, 0.0f);
    o[2] = make_float4(b0.x, b0.y, b0.z, (float)sz);
    o[3] = make_float4(b1.x, b1.y, b1.z, (float)offset[id]);
  }
}

// Host driver: builds into caller-provided device buffers `nodes` (capacity
// 2T-1 nodes) and `bvh_idx` (T), on `stream`; returns the node count.
// method 0: linear BVH (Karras); 1: PLOC (better trees, a few times slower).
int gpu_bvh_build(hipStream_t stream, const float4* verts, const uint32_t* tris, int32_t n, int32_t max_leaf, int32_t method,
                  float4* nodes, int32_t* bvh_idx, int64_t* num_nodes, std::string& err) {
  if (n <= 0) { err = "bvh build: no triangles"; return DSPT_ERR_ARG; }
  if (method != 0 && method != 1) { err = "bvh build: unknown method"; return DSPT_ERR_ARG; }
  max_leaf = std::max(1, std::min(max_leaf, 64));
  const int64_t m = 2 * (int64_t)n - 1;
  if (m >= (1LL << 24)) { err = "bvh build: more than 2^24 nodes (float-encoded indices)"; return DSPT_ERR_UNSUPPORTED; }
  Allocs A;
  float4 *tlo, *thi, *blo, *bhi;
  unsigned long long *k0, *k1;
  int32_t *v0, *v1, *parent, *arrivals, *flag, *slot;
  int2 *range, *child;
  unsigned* cb;
  int r = 0;
  if ((r = A.get(&tlo, n)) || (r = A.get(&thi, n)) || (r = A.get(&blo, m)) || (r = A.get(&bhi, m)) || (r = A.get(&k0, n)) ||
      (r = A.get(&k1, n)) || (r = A.get(&v0, n)) || (r = A.get(&v1, n)) || (r = A.get(&parent, m)) ||
      (r = A.get(&arrivals, m)) || (r = A.get(&flag, m)) || (r = A.get(&slot, m)) ||
      (r = A.get(&range, std::max<int64_t>(n - 1, 1))) || (r = A.get(&child, std::max<int64_t>(n - 1, 1))) ||
      (r = A.get(&cb, 6))) {
    A.release();
    err = "bvh build: device memory";
    return r;
  }
  const unsigned init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
  void* tmp = nullptr;
  auto fail_hip = [&](hipError_t e, const char* what) {
    if (tmp) (void)hipFree(tmp);
    A.release();
    err = std::string("bvh build: ") + what + ": " + hipGetErrorString(e);
    return DSPT_ERR_DEVICE;
  };
  hipError_t e;
  if ((e = hipMemcpyAsync(cb, init, sizeof(init), hipMemcpyHostToDevice, stream)) != hipSuccess) return fail_hip(e, "copy");
  k_bvh_boxes<<<grid_for(n, 256), 256, 0, stream>>>(verts, tris, n, tlo, thi, cb);
  k_bvh_morton<<<grid_for(n, 256), 256, 0, stream>>>(tlo, thi, cb, n, k0, v0);
  size_t sort_bytes = 0, scan_bytes = 0;
  if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, k0, k1, v0, v1, n, 0, 63, stream)) != hipSuccess)
    return fail_hip(e, "sort sizing");
  if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, slot, (int)m, stream)) != hipSuccess)
    return fail_hip(e, "scan sizing");
  size_t tmp_bytes = std::max(sort_bytes, scan_bytes);
  if ((e = hipMalloc(&tmp, tmp_bytes)) != hipSuccess) return fail_hip(e, "sort storage");
  if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, k0, k1, v0, v1, n, 0, 63, stream)) != hipSuccess)
    return fail_hip(e, "sort");
  k_fill_i32<<<grid_for(m, 256), 256, 0, stream>>>(parent, (int)m, -1);
  int64_t count = 0;
  if (method == 0) {
    // ---- linear BVH
    k_fill_i32<<<grid_for(m, 256), 256, 0, stream>>>(arrivals, (int)m, 0);
    if (n > 1) k_bvh_karras<<<grid_for(n - 1, 256), 256, 0, stream>>>(k1, n, parent, range, child);
    k_bvh_refit<<<grid_for(n, 256), 256, 0, stream>>>(tlo, thi, v1, n, parent, child, blo, bhi, arrivals);
    k_bvh_flags<<<grid_for(m, 256), 256, 0, stream>>>(n, max_leaf, parent, range, flag);
    if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, flag, slot, (int)m, stream)) != hipSuccess)
      return fail_hip(e, "scan");
    k_bvh_emit<<<grid_for(m, 256), 256, 0, stream>>>(n, max_leaf, range, child, flag, slot, blo, bhi, nodes);
    if ((e = hipMemcpyAsync(bvh_idx, v1, (size_t)n * 4, hipMemcpyDeviceToDevice, stream)) != hipSuccess)
      return fail_hip(e, "copy");
  } else {
    // ---- PLOC: cluster arrays double-buffered in (k0 as int32 node ids | blo/bhi halves are the node boxes)
    int32_t *cn[2], *nn, *mflag, *mscan, *kscan, *nsize, *esize, *offset;
    float4 *clo[2], *chi[2];
    Allocs B;
    if ((r = B.get(&cn[0], n)) || (r = B.get(&cn[1], n)) || (r = B.get(&clo[0], n)) || (r = B.get(&clo[1], n)) ||
        (r = B.get(&chi[0], n)) || (r = B.get(&chi[1], n)) || (r = B.get(&nn, n)) || (r = B.get(&mflag, n)) ||
        (r = B.get(&mscan, n)) || (r = B.get(&kscan, n)) || (r = B.get(&nsize, m)) || (r = B.get(&esize, m)) ||
        (r = B.get(&offset, m))) {
      B.release();
      if (tmp) (void)hipFree(tmp);
      A.release();
      err = "bvh build: device memory";
      return r;
    }
    int32_t* keep = arrivals;  // reused: n entries
    k_ploc_init<<<grid_for(n, 256), 256, 0, stream>>>(tlo, thi, v1, n, blo, bhi, nsize, cn[0], clo[0], chi[0]);
    int32_t c = n, next_id = n, cur = 0;
    int32_t tot[2] = {0, 0};
    int32_t* pin = nullptr;
    if ((e = hipHostMalloc(&pin, 4 * sizeof(int32_t))) != hipSuccess) { B.release(); return fail_hip(e, "pinned"); }
    while (c > 1) {
      k_ploc_nn<<<grid_for(c, 256), 256, 0, stream>>>(clo[cur], chi[cur], c, nn);
      k_ploc_flags<<<grid_for(c, 256), 256, 0, stream>>>(nn, c, mflag, keep);
      size_t sb = tmp_bytes;
      if ((e = hipcub::DeviceScan::ExclusiveSum(tmp, sb, mflag, mscan, c, stream)) != hipSuccess ||
          (e = hipcub::DeviceScan::ExclusiveSum(tmp, sb, keep, kscan, c, stream)) != hipSuccess) {
        (void)hipHostFree(pin);
        B.release();
        return fail_hip(e, "scan");
      }
      k_ploc_merge<<<grid_for(c, 256), 256, 0, stream>>>(nn, c, n, next_id, max_leaf, esize, mflag, mscan, keep, kscan,
                                                          cn[cur], clo[cur],
                                                          chi[cur], cn[cur ^ 1], clo[cur ^ 1], chi[cur ^ 1], blo, bhi, child,
                                                          parent, nsize);
      (void)hipMemcpyAsync(pin, mscan + c - 1, 4, hipMemcpyDeviceToHost, stream);
      (void)hipMemcpyAsync(pin + 1, mflag + c - 1, 4, hipMemcpyDeviceToHost, stream);
      (void)hipMemcpyAsync(pin + 2, kscan + c - 1, 4, hipMemcpyDeviceToHost, stream);
      (void)hipMemcpyAsync(pin + 3, keep + c - 1, 4, hipMemcpyDeviceToHost, stream);
      if ((e = hipStreamSynchronize(stream)) != hipSuccess) { (void)hipHostFree(pin); B.release(); return fail_hip(e, "sync"); }
      tot[0] = pin[0] + pin[1];
      tot[1] = pin[2] + pin[3];
      if (tot[0] <= 0 || tot[1] >= c) {
        (void)hipHostFree(pin);
        B.release();
        if (tmp) (void)hipFree(tmp);
        A.release();
        err = "bvh build: clustering made no progress (non-finite triangle bounds?)";
        return DSPT_ERR_ARG;
      }
      next_id += tot[0];
      c = tot[1];
      cur ^= 1;
    }
    (void)hipHostFree(pin);
    k_ploc_layout<<<grid_for(m, 256), 256, 0, stream>>>(n, max_leaf, parent, child, nsize, esize, offset, flag, slot);
    k_ploc_order<<<grid_for(n, 256), 256, 0, stream>>>(n, offset, v1, bvh_idx);
    k_ploc_emit<<<grid_for(m, 256), 256, 0, stream>>>(n, max_leaf, child, nsize, offset, flag, slot, blo, bhi, nodes);
    // output node count: the root's (the last node created; a lone triangle is its own root)
    int32_t rs = 1;
    if (n > 1) {
      int32_t rsz = 0;
      (void)hipMemcpyAsync(&rsz, nsize + (m - 1), 4, hipMemcpyDeviceToHost, stream);
      (void)hipMemcpyAsync(&rs, esize + (m - 1), 4, hipMemcpyDeviceToHost, stream);
      if ((e = hipStreamSynchronize(stream)) != hipSuccess) { B.release(); return fail_hip(e, "sync"); }
      if (rsz <= max_leaf) rs = 1;
    }
    if ((e = hipStreamSynchronize(stream)) != hipSuccess) { B.release(); return fail_hip(e, "sync"); }
    B.release();
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e, "launch");
    (void)hipFree(tmp);
    tmp = nullptr;
    A.release();
    *num_nodes = rs;
    return 0;
  }
  if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e, "launch");
  int32_t last_slot = 0, last_flag = 0;
  (void)hipMemcpyAsync(&last_slot, slot + m - 1, 4, hipMemcpyDeviceToHost, stream);
  (void)hipMemcpyAsync(&last_flag, flag + m - 1, 4, hipMemcpyDeviceToHost, stream);
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail_hip(e, "sync");
  (void)hipFree(tmp);
  tmp = nullptr;
  A.release();
  count = (int64_t)last_slot + last_flag;
  *num_nodes = count;
  return 0;
}
