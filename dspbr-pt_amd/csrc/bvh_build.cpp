// Engine-side BVH builder emitting the TinyBVH "bvh_gpu" node layout that
// the reference's traversal consumes (packages/lib/shader/bvh.glsl:1-6):
// node n = 16 floats (lmin.xyz, left) (lmax.xyz, right) (rmin.xyz, count)
// (rmax.xyz, first); count > 0 marks a leaf over bvhIndex[first, first+count).
// Indices are stored as float values (exact below 2^24).
//
// Binned SAH (16 bins per axis), leaves of at most `max_leaf` triangles,
// subtrees built in parallel on host threads. The tree topology is an
// input to the path tracer (both the engine and the oracle traverse the
// same array), not a parity-relevant computation.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <future>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float* p) {
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
  }
  void grow(const Box& b) {
    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
  }
  float area() const {
    float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    if (d0 < 0 || d1 < 0 || d2 < 0) return 0.0f;
    return 2.0f * (d0 * d1 + d1 * d2 + d2 * d0);
  }
};

struct BNode {
  Box box;
  int32_t left = -1, right = -1;  // tree node ids (interior)
  int32_t first = 0, count = 0;   // leaf range in the index array
};

struct Builder {
  const Box* tbox;
  const float* cent;  // 3 per triangle
  std::vector<int32_t>& idx;
  std::vector<BNode>& nodes;
  std::atomic<int32_t>& next;
  int max_leaf;

  int32_t alloc() { return next.fetch_add(1); }

  int32_t build(int32_t first, int32_t count, int depth) {
    int32_t id = alloc();
    BNode& nd = nodes[id];
    Box b, cb;
    for (int32_t i = first; i < first + count; i++) {
      b.grow(tbox[idx[i]]);
      cb.grow(cent + 3 * (size_t)idx[i]);
    }
    nd.box = b;
    if (count <= 1) { nd.first = first; nd.count = count; return id; }
    // binned SAH over the centroid bounds
    const int NB = 16;
    int best_axis = -1, best_split = -1;
    float best_cost = INFINITY;
    for (int ax = 0; ax < 3; ax++) {
      float ext = cb.hi[ax] - cb.lo[ax];
      if (!(ext > 0.0f)) continue;
      Box bb[NB];
      int bc[NB] = {0};
      float scale = NB / ext;
      for (int32_t i = first; i < first + count; i++) {
        int t = idx[i];
        int bi = std::min(NB - 1, (int)((cent[3 * (size_t)t + ax] - cb.lo[ax]) * scale));
        bc[bi]++;
        bb[bi].grow(tbox[t]);
      }
      float la[NB - 1];
      int lc[NB - 1];
      Box acc;
      int c = 0;
      for (int k = 0; k < NB - 1; k++) { acc.grow(bb[k]); c += bc[k]; la[k] = acc.area(); lc[k] = c; }
      Box racc;
      int rc = 0;
      for (int k = NB - 1; k > 0; k--) {
        racc.grow(bb[k]); rc += bc[k];
        int nl = lc[k - 1];
        if (nl == 0 || rc == 0) continue;
        float cost = la[k - 1] * nl + racc.area() * rc;
        if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = k; }
      }
    }
    float leaf_cost = b.area() * count;
    float split_cost = b.area() * 1.0f + best_cost;  // one node visit + both children
    if (best_axis < 0) {
      if (count <= max_leaf) { nd.first = first; nd.count = count; return id; }
      // coincident centroids: split the range in halves
      int32_t h = count / 2;
      int32_t l = build(first, h, depth + 1);
      int32_t r = build(first + h, count - h, depth + 1);
      nodes[id].left = l;
      nodes[id].right = r;
      return id;
    }
    if (count <= max_leaf && split_cost >= leaf_cost) { nd.first = first; nd.count = count; return id; }
    float ext = cb.hi[best_axis] - cb.lo[best_axis];
    float scale = NB / ext;
    auto mid_it = std::partition(idx.begin() + first, idx.begin() + first + count, [&](int32_t t) {
      int bi = std::min(NB - 1, (int)((cent[3 * (size_t)t + best_axis] - cb.lo[best_axis]) * scale));
      return bi < best_split;
    });
    int32_t nl = (int32_t)(mid_it - (idx.begin() + first));
    if (nl == 0 || nl == count) nl = count / 2;
    int32_t l, r;
    if (depth < 10 && count > 65536) {
      auto fut = std::async(std::launch::async, [&, first, nl, depth]() { return build(first, nl, depth + 1); });
      r = build(first + nl, count - nl, depth + 1);
      l = fut.get();
    } else {
      l = build(first, nl, depth + 1);
      r = build(first + nl, count - nl, depth + 1);
    }
    nodes[id].left = l;
    nodes[id].right = r;
    return id;
  }
};

}  // namespace

// Returns 0 on success. out_nodes: 16 floats per node (DFS order, root 0).
namespace {

// Shared back end: binned SAH over per-item boxes and centroids, then the
// bvh_gpu layout in DFS order (root 0). out_idx lists item ids per leaf.
int build_items(std::vector<Box>& tb, std::vector<float>& cent, int64_t nt, int max_leaf, std::vector<float>& out_nodes,
                std::vector<int32_t>& out_idx) {
  std::vector<int32_t> idx(nt);
  for (int64_t t = 0; t < nt; t++) idx[t] = (int32_t)t;
  std::vector<BNode> nodes(2 * (size_t)nt);
  std::atomic<int32_t> next{0};
  Builder b{tb.data(), cent.data(), idx, nodes, next, max_leaf};
  b.build(0, (int32_t)nt, 0);
  int32_t total = next.load();
  // DFS renumbering, root = 0
  std::vector<int32_t> order;
  order.reserve(total);
  std::vector<int32_t> remap(total, -1);
  std::vector<int32_t> stack{0};
  while (!stack.empty()) {
    int32_t n = stack.back();
    stack.pop_back();
    remap[n] = (int32_t)order.size();
    order.push_back(n);
    if (nodes[n].count == 0 && nodes[n].left >= 0) {
      stack.push_back(nodes[n].right);
      stack.push_back(nodes[n].left);
    }
  }
  out_nodes.assign((size_t)order.size() * 16, 0.0f);
  for (size_t i = 0; i < order.size(); i++) {
    const BNode& n = nodes[order[i]];
    float* o = &out_nodes[i * 16];
    if (n.count > 0 || n.left < 0) {
      // leaf: own bounds in the left slot, count/first in the w fields
      for (int k = 0; k < 3; k++) { o[k] = n.box.lo[k]; o[4 + k] = n.box.hi[k]; o[8 + k] = n.box.lo[k]; o[12 + k] = n.box.hi[k]; }
      o[3] = 0.0f; o[7] = 0.0f;
      o[11] = (float)n.count;
      o[15] = (float)n.first;
    } else {
      const BNode& L = nodes[n.left];
      const BNode& R = nodes[n.right];
      for (int k = 0; k < 3; k++) { o[k] = L.box.lo[k]; o[4 + k] = L.box.hi[k]; o[8 + k] = R.box.lo[k]; o[12 + k] = R.box.hi[k]; }
      o[3] = (float)remap[n.left];
      o[7] = (float)remap[n.right];
      o[11] = 0.0f;
      o[15] = 0.0f;
    }
  }
  out_idx = std::move(idx);
  return 0;
}

}  // namespace

int dspt_bvh_build_impl(const float* pos, int64_t nv, const uint32_t* ind, int64_t nt, int max_leaf,
                        std::vector<float>& out_nodes, std::vector<int32_t>& out_idx, std::string& err) {
  if (nt <= 0) { err = "bvh build: no triangles"; return -1; }
  if (nt > (1 << 24)) { err = "bvh build: more than 2^24 triangles (float-encoded indices)"; return -1; }
  if (max_leaf < 1) max_leaf = 1;
  if (max_leaf > 64) max_leaf = 64;
  std::vector<Box> tb(nt);
  std::vector<float> cent(3 * (size_t)nt);
  for (int64_t t = 0; t < nt; t++) {
    for (int k = 0; k < 3; k++) {
      uint32_t v = ind[3 * t + k];
      if ((int64_t)v >= nv) { err = "bvh build: vertex index out of range"; return -1; }
      tb[t].grow(pos + 3 * (size_t)v);
    }
    for (int k = 0; k < 3; k++) cent[3 * t + k] = 0.5f * (tb[t].lo[k] + tb[t].hi[k]);
  }
  return build_items(tb, cent, nt, max_leaf, out_nodes, out_idx);
}

// Same layout over axis-aligned boxes (f32[6n]: lo.xyz, hi.xyz), e.g. the
// world bounds of instances for a top-level BVH.
int dspt_bvh_build_boxes_impl(const float* boxes, int64_t n, int max_leaf, std::vector<float>& out_nodes,
                              std::vector<int32_t>& out_idx, std::string& err) {
  if (n <= 0) { err = "bvh build: no boxes"; return -1; }
  if (n > (1 << 24)) { err = "bvh build: more than 2^24 boxes"; return -1; }
  if (max_leaf < 1) max_leaf = 1;
  if (max_leaf > 64) max_leaf = 64;
  std::vector<Box> tb(n);
  std::vector<float> cent(3 * (size_t)n);
  for (int64_t i = 0; i < n; i++) {
    tb[i].grow(boxes + 6 * i);
    tb[i].grow(boxes + 6 * i + 3);
    for (int k = 0; k < 3; k++) {
      if (!std::isfinite(boxes[6 * i + k]) || !std::isfinite(boxes[6 * i + 3 + k])) { err = "bvh build: non-finite box"; return -1; }
      cent[3 * i + k] = 0.5f * (tb[i].lo[k] + tb[i].hi[k]);
    }
  }
  return build_items(tb, cent, n, max_leaf, out_nodes, out_idx);
}
