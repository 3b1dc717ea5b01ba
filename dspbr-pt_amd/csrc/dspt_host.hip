// Host runtime and C ABI (include/dspt.h) of the gfx950 path-tracing engine.
//
// Resources: resident scenes and IBLs (ids), render jobs (scene + IBL +
// frame block + resolution + frame range + accumulation). All rendering
// goes through one path pool: work items (job, frame, pixel) are issued in
// the order of a host-maintained task list, and a slot freed by a finished
// path is refilled on the next bounce. The host runs the pool in chunks of
// a few bounces; after each chunk the device folds every complete frame
// into its job's running mean and the host reads back the progress one
// chunk late (the GPU always has the next chunk queued). Two schedules:
//   STREAM  frames are appended round-robin over the listed jobs ahead of
//           the pool, so the pool stays full across frames and jobs;
//   PASS    frame-synchronous passes: up to a pool's worth of frames is
//           issued only when the pool is idle (the plain loop).
// A job's frame f is scheduled only when frame f - ring_frames has been
// folded, which bounds the ring of finished-but-unfolded frames.
// The single-image API (dspt_upload_scene / dspt_upload_ibl /
// dspt_set_frame / dspt_render) drives an implicit job on scene 0 / IBL 0.
// Kernels are in dspt_kernels.hip (same translation unit).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dspt.h"
#include "dspt_kernels.hip"

int dspt_bvh_build_impl(const float* pos, int64_t nv, const uint32_t* ind, int64_t nt, int max_leaf,
                        std::vector<float>& out_nodes, std::vector<int32_t>& out_idx, std::string& err);
int dspt_bvh_build_boxes_impl(const float* boxes, int64_t n, int max_leaf, std::vector<float>& out_nodes,
                              std::vector<int32_t>& out_idx, std::string& err);

// Scratch memory of the allocation-free ray query (dspt_trace_closest_noalloc):
// the traversal-stack spill (kStackSize - kLdsStack ints per lane) of at most
// kNoAllocBlocks blocks of kTraceBlock lanes, 24 MiB. Every GPU the library
// runs on holds its own copy until the process ends. Host code never uses the
// array's name as an address: dspt_create looks up each device's address
// with hipGetSymbolAddress and the entry point passes that to the kernel.
constexpr int kNoAllocBlocks = 512;
constexpr size_t kNoAllocScratchInts = (size_t)kNoAllocBlocks * kTraceBlock * (kStackSize - kLdsStack);
static __device__ __align__(256) int scratch[kNoAllocScratchInts];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                                       \
  do {                                                                                                     \
    hipError_t e_ = (expr);                                                                                \
    if (e_ != hipSuccess) return fail(DSPT_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxJobs = 4096;
constexpr int kMaxTasks = 8192;
constexpr int64_t kStallChunks = 4096;

// Device-memory limit of a context (dspt_set_memory_limit): every Allocs of
// the context counts against it while it holds memory.
struct MemLimit {
  int64_t cap = 0;  // 0 = no limit
  int64_t used = 0;
};

struct Allocs {
  std::vector<void*> ptrs;
  size_t bytes = 0;
  MemLimit* lim = nullptr;
  template <class T>
  int get(T** out, size_t n) {
    void* p = nullptr;
    size_t sz = std::max<size_t>(n, 1) * sizeof(T);
    if (lim && lim->cap > 0 && lim->used + (int64_t)sz > lim->cap)
      return fail(DSPT_ERR_MEMORY, "device memory limit: " + std::to_string(sz) + " B more would exceed " +
                                       std::to_string(lim->cap) + " B (" + std::to_string(lim->used) + " B in use)");
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess)
      return fail(DSPT_ERR_MEMORY, std::string("hipMalloc(") + std::to_string(sz) + "): " + hipGetErrorString(e));
    ptrs.push_back(p);
    bytes += sz;
    if (lim) lim->used += (int64_t)sz;
    *out = (T*)p;
    return 0;
  }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
    if (lim) lim->used -= (int64_t)bytes;
    bytes = 0;
  }
};

// Allocations that live for one call: released on every return path.
struct ScopedAllocs : Allocs {
  ScopedAllocs() = default;
  ScopedAllocs(const ScopedAllocs&) = delete;
  ScopedAllocs& operator=(const ScopedAllocs&) = delete;
  ~ScopedAllocs() { release(); }
};

struct SceneRes {
  Allocs mem;
  float4 *nodes = nullptr, *tri = nullptr, *verts = nullptr, *mats = nullptr;
  uint32_t* tris = nullptr;
  int32_t N = 0, M = 0, T = 0, K = 0;
  int64_t V = 0;
  int users = 0;
  int32_t max_tex_id = -1;  // largest texture id a material references
  Allocs tex_mem;           // material textures (replaced as a whole)
  float4* texinfo = nullptr;
  TexArrayDesc* texarr = nullptr;
  uint32_t* texels = nullptr;
  int32_t num_tex = 0, num_texarr = 0;
  // two-level scenes: root of the top level and first instance record in `nodes`
  int32_t top = -1, inst_base = -1, I = 0;
  int32_t fast_slab = 0;  // every inner node's child boxes are finite
  int32_t key_bits = 0;   // ray-order grid bits per axis (SceneView::key_bits)
  int32_t* d_flag = nullptr;
  uint32_t* tri_mat = nullptr;  // material of each triangle (resident scenes)
  float4* tri_verts = nullptr;  // vertex records per triangle corner (resident scenes)
};

struct IblRes {
  Allocs mem;
  float4* env = nullptr;
  float *pdf = nullptr, *cdf = nullptr, *ypdf = nullptr, *ycdf = nullptr;
  int32_t *guide = nullptr, *yguide = nullptr;
  int w = 0, h = 0;
  double total = 0.0;
  int users = 0;
};

struct JobRes {
  Allocs mem;
  int scene = -1, ibl = -1;
  FrameParams F{};
  int W = 0, H = 0;
  float4* accum = nullptr;
  float4* ring = nullptr;
  int32_t* done = nullptr;
  int ring_frames = 1;
  int frame_begin = 1, frame_end = 1;  // current frame range [begin, end)
  int next_sched = 1;                  // next frame to schedule
  int acc_known = 1;                   // frames below are folded (latest readback)
  int64_t epoch = 0;                   // first chunk whose readback describes this range
  bool record = false;
  float4* rec_L = nullptr;
  uint32_t* rec_sig = nullptr;
  int32_t* rec_prim = nullptr;
  int rec_first = 1, rec_count = 0;
  bool alive = false, cancel = false;
  // render region (rw == 0: the whole image), set per call by the
  // single-image API; its pixels are the work items of each frame
  int rx0 = 0, ry0 = 0, rw = 0, rh = 0;
  int64_t npix() const { return (int64_t)W * H; }
  int64_t items() const { return rw > 0 ? (int64_t)rw * rh : npix(); }
  bool complete() const { return acc_known >= frame_end; }
  bool in_flight() const { return next_sched > acc_known; }
};

struct Readback {
  StreamCtl ctl;
  int32_t prog[kMaxJobs];
};

// Device address of `scratch` per device, looked up once per device with that
// device current.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(int device, int** out) {
  if (device < 0 || device >= kMaxDevices) return fail(DSPT_ERR_UNSUPPORTED, "device index above the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (int*)p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace

struct dspt_ctx {
  int device = 0;
  int profile = 0;
  hipStream_t stream = nullptr;
  std::vector<SceneRes*> scenes;
  std::vector<IblRes*> ibls;
  std::vector<JobRes*> jobs;
  // single-image API state (implicit job on scene 0 / IBL 0)
  FrameParams F{};
  bool have_frame = false;
  int W = 0, H = 0;
  int legacy_job = -1;
  bool record = false;
  int last_frames = 0;
  // job table
  DevJob* d_jobs = nullptr;
  DevJob* h_jobs = nullptr;  // pinned staging
  bool jobs_dirty = true;
  MemLimit lim;  // device-memory limit over every Allocs below
  // path pool
  Allocs state_mem;
  int64_t cap = 0;
  PathState ps[2] = {};  // alternate per bounce; hit and shadow fields shared
  int32_t* shq = nullptr;
  int* spill = nullptr;
  int trace_blocks = 0;
  int trace_blocks_max = 2048;
  int shade_blocks_max = 16384;  // blocks of 256 threads, 64 per CU (8192 and 32768 measured slower)
  int64_t max_paths = 16 << 20;  // about 5 GB of path state
  unsigned long long* cnt = nullptr;
  // scheduler
  StreamCtl* d_ctl = nullptr;
  StreamSched* d_sched = nullptr;
  int32_t* d_prog = nullptr;
  int32_t* d_joblist = nullptr;
  PassTask* d_tasks = nullptr;
  Readback* h_rb = nullptr;        // pinned [2]
  StreamSched* h_sched = nullptr;  // pinned [2]
  PassTask* h_tasks = nullptr;     // pinned [2][kMaxTasks]
  int32_t* h_joblist = nullptr;    // pinned [2][kMaxJobs]
  hipEvent_t ev_chunk[2] = {nullptr, nullptr};
  int64_t limit_at[2] = {0, 0};
  int qsel_at[2] = {0, 0};  // state buffer holding the live paths after the chunk
  int64_t chunk_no = 0;
  bool have_prev = false;
  int qsel = 0;
  int64_t sched_limit = 0, next_known = 0, limit_known = 0, alive_known = 0;
  // items below this were offered to an enqueued chunk: the tasks holding
  // them may be partly issued, so they are never extended (the item order
  // within a task depends on its frame count)
  int64_t exposed_limit = 0;
  std::deque<PassTask> tasks;
  int mode = DSPT_SCHED_STREAM;
  int policy = DSPT_POLICY_RR;
  int pass_frames = 0;
  int chunk_iters = 2;
  float ring_factor = 8.0f;
  size_t rr = 0;
  int64_t budget = 0;  // admission budget for job memory (0 = unlimited)
  // kernel timing
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  double kstat[DSPT_K_COUNT][2] = {};
  // canvas output of the display chain, grown on demand
  char* post_buf = nullptr;
  size_t post_cap = 0;
  // the bounce loop of a chunk captured as a HIP graph, one per (state
  // buffer holding the live paths at chunk start, two-level kernels, debug
  // shading instance); rebuilt when the pool, the grids or the chunk length
  // change
  bool use_graphs = true;
  hipGraphExec_t bounce_graph[2][2][2] = {};
  std::array<int64_t, 6> graph_key{};  // (pool, pool address, grids, chunk length, ray order) of the graphs
  // ray order of the extension queue (allocated with the pool)
  int ray_order = DSPT_ORDER_SHADE;
  int* scratch = nullptr;  // this device's address of the no-alloc scratch array
  uint16_t* order_keys = nullptr;
  int32_t* order_perm = nullptr;
  uint32_t* order_bins = nullptr;
  int32_t* shade_perm = nullptr;
  uint32_t* shade_bins = nullptr;
};

namespace {

int grid_for(int64_t n, int block) { return (int)std::max<int64_t>(1, (n + block - 1) / block); }

// Sums the kCntSlots rows of the work counters into h[kNumCounters].
int read_counters(dspt_ctx* c, unsigned long long* h) {
  std::vector<unsigned long long> rows((size_t)kCntSlots * kNumCounters);
  HIPCHK(hipMemcpy(rows.data(), c->cnt, rows.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int i = 0; i < kNumCounters; i++) h[i] = 0;
  for (int r = 0; r < kCntSlots; r++)
    for (int i = 0; i < kNumCounters; i++) h[i] += rows[(size_t)r * kNumCounters + i];
  return 0;
}

void drop_graphs(dspt_ctx* c) {
  for (auto& plane : c->bounce_graph)
    for (auto& row : plane)
      for (auto& g : row)
        if (g) {
          (void)hipGraphExecDestroy(g);
          g = nullptr;
        }
}

#include "bvh_gpu.hip"

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

// With timing on, brackets the launch with events on the context stream
// (the stream the kernel runs on).
template <class Fn>
void timed(dspt_ctx* c, int kclass, Fn&& launch) {
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

int ring_frames_for(const dspt_ctx* c, int64_t npix, int frames) {
  double want = std::ceil((double)c->ring_factor * (double)c->max_paths / (double)npix) + 2.0;
  int64_t r = (int64_t)std::min<double>(want, 1 << 20);
  r = std::min<int64_t>(r, std::max(frames, 1));
  return (int)std::max<int64_t>(r, 1);
}

int64_t job_bytes(int64_t npix, bool record, int64_t frames, int ring) {
  int64_t b = npix * 16 + (int64_t)ring * npix * 16 + (int64_t)ring * 4;
  if (record) b += frames * npix * (16 + 8);
  return b;
}

int64_t memory_in_use(dspt_ctx* c) {
  int64_t b = (int64_t)c->state_mem.bytes;
  for (auto* s : c->scenes) if (s) b += (int64_t)s->mem.bytes + (int64_t)s->tex_mem.bytes;
  for (auto* i : c->ibls) if (i) b += (int64_t)i->mem.bytes;
  for (auto* j : c->jobs) if (j) b += (int64_t)j->mem.bytes;
  return b;
}

SceneView view_of(dspt_ctx* c, const JobRes* j) {
  SceneView S{};
  const SceneRes* s = c->scenes[j->scene];
  S.nodes = s->nodes; S.tri = s->tri; S.verts = s->verts; S.tris = s->tris; S.mats = s->mats; S.tri_mat = s->tri_mat;
  S.tri_verts = s->tri_verts;
  S.N = s->N; S.M = s->M; S.T = s->T; S.K = s->K;
  S.texinfo = s->texinfo; S.texarr = s->texarr; S.texels = s->texels;
  S.num_tex = s->num_tex; S.num_texarr = s->num_texarr;
  S.top = s->top; S.inst_base = s->inst_base;
  S.fast_slab = s->fast_slab;
  S.key_bits = s->key_bits;
  const IblRes* ib = (j->ibl >= 0 && j->ibl < (int)c->ibls.size()) ? c->ibls[j->ibl] : nullptr;
  if (ib) {
    S.env = ib->env; S.pdf = ib->pdf; S.cdf = ib->cdf; S.ypdf = ib->ypdf; S.ycdf = ib->ycdf;
    S.guide = ib->guide; S.yguide = ib->yguide;
    S.ew = ib->w; S.eh = ib->h; S.has_ibl = 1;
  }
  S.full_profile = c->profile == DSPT_PROFILE_FULL;
  return S;
}

int sync_job_table(dspt_ctx* c) {
  if (!c->jobs_dirty) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));  // staging buffer may be in use by a prior copy
  for (size_t i = 0; i < c->jobs.size(); i++) {
    JobRes* j = c->jobs[i];
    DevJob d{};
    if (j && j->alive) {
      d.S = view_of(c, j);
      d.F = j->F;
      d.W = j->W; d.H = j->H; d.npix = (int32_t)j->npix();
      d.rx0 = j->rw > 0 ? j->rx0 : 0; d.ry0 = j->rw > 0 ? j->ry0 : 0;
      d.rw = j->rw > 0 ? j->rw : j->W; d.rn = (int32_t)j->items();
      d.scene = j->scene;
      d.accum = j->accum;
      d.ring = j->ring; d.done = j->done; d.ring_frames = j->ring_frames; d.frame_end = j->frame_end;
      d.rec_L = j->record ? j->rec_L : nullptr;
      d.rec_sig = j->rec_sig; d.rec_prim = j->rec_prim;
      d.rec_first = j->rec_first;
    }
    c->h_jobs[i] = d;
  }
  if (!c->jobs.empty())
    HIPCHK(hipMemcpyAsync(c->d_jobs, c->h_jobs, c->jobs.size() * sizeof(DevJob), hipMemcpyHostToDevice, c->stream));
  c->jobs_dirty = false;
  return 0;
}

bool job_ok(dspt_ctx* c, int id) { return id >= 0 && id < (int)c->jobs.size() && c->jobs[id] && c->jobs[id]->alive; }

// The pool is idle once the latest readback shows every scheduled item
// issued and no path alive, and nothing was scheduled after that chunk.
bool pool_idle(const dspt_ctx* c) {
  return c->next_known >= c->limit_known && c->alive_known == 0 && c->limit_known == c->sched_limit;
}

// Path pool of c->max_paths slots (reallocated only while idle).
int ensure_pool(dspt_ctx* c) {
  if (c->cap > 0 && (c->cap == c->max_paths || !pool_idle(c))) return 0;  // resize once idle
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t cap = c->max_paths;
  drop_graphs(c);  // they hold the old pool's addresses
  c->graph_key = {};
  c->state_mem.release();
  c->cap = 0;
  PathState& a = c->ps[0];
  PathState& b = c->ps[1];
  a = PathState{};
  b = PathState{};
  Allocs& A = c->state_mem;
  int r = 0;
  for (PathState* ps : {&a, &b})
    if ((r = A.get(&ps->org, cap)) || (r = A.get(&ps->dir, cap)) || (r = A.get(&ps->pw, cap)) || (r = A.get(&ps->L, cap)) ||
        (r = A.get(&ps->meta, cap)) ||
        (r = A.get(&ps->sig, cap)) || (r = A.get(&ps->prim, cap)))
      return r;
  if ((r = A.get(&a.hit, cap)) || (r = A.get(&a.hinst, cap)) || (r = A.get(&a.sh_org, cap)) || (r = A.get(&a.sh_dir, cap)) || (r = A.get(&a.sh_c, cap)) ||
      (r = A.get(&a.sh_pw, cap)) || (r = A.get(&a.sh_pdir, cap)) || (r = A.get(&a.sh_pc, cap)) || (r = A.get(&c->shq, cap)))
    return r;
  b.hit = a.hit; b.hinst = a.hinst; b.sh_org = a.sh_org; b.sh_dir = a.sh_dir; b.sh_c = a.sh_c; b.sh_pw = a.sh_pw; b.sh_pdir = a.sh_pdir;
  b.sh_pc = a.sh_pc;
  c->trace_blocks = (int)std::min<int64_t>(c->trace_blocks_max, grid_for(cap, kTraceBlock));
  if ((r = A.get(&c->spill, (size_t)c->trace_blocks * kTraceBlock * (kStackSize - kLdsStack)))) return r;
  if ((r = A.get(&c->order_keys, cap)) || (r = A.get(&c->order_perm, cap)) || (r = A.get(&c->order_bins, kOrderBins)) ||
      (r = A.get(&c->shade_perm, cap)) || (r = A.get(&c->shade_bins, kOrderBins)))
    return r;
  k_stream_init<<<1, 64, 0, c->stream>>>(c->d_ctl);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  c->cap = cap;
  c->qsel = 0;
  c->sched_limit = c->next_known = c->limit_known = c->alive_known = 0;
  c->exposed_limit = 0;
  c->limit_at[0] = c->limit_at[1] = 0;
  c->tasks.clear();
  c->have_prev = false;
  return 0;
}

bool eligible(const dspt_ctx* c, const JobRes* j) {
  (void)c;
  return j && j->alive && !j->cancel && j->next_sched < j->frame_end && j->next_sched < j->acc_known + j->ring_frames;
}

bool append_frames(dspt_ctx* c, int id, int nf) {
  JobRes* j = c->jobs[id];
  const int64_t npix = j->items();
  if (!c->tasks.empty()) {
    PassTask& b = c->tasks.back();
    if (b.job == id && b.first_frame + b.nframes == j->next_sched && b.offset + (int64_t)b.nframes * b.npix == c->sched_limit &&
        b.offset >= c->exposed_limit) {
      b.nframes += nf;
      c->sched_limit += (int64_t)nf * npix;
      j->next_sched += nf;
      return true;
    }
  }
  if ((int)c->tasks.size() >= kMaxTasks) return false;
  PassTask t;
  t.job = id; t.first_frame = j->next_sched; t.nframes = nf; t.npix = (int32_t)npix; t.offset = c->sched_limit;
  c->tasks.push_back(t);
  c->sched_limit += (int64_t)nf * npix;
  j->next_sched += nf;
  return true;
}

// Appends frames of the listed jobs to the task list (see the file comment).
void schedule(dspt_ctx* c, const std::vector<int>& ids) {
  if (ids.empty()) return;
  if (c->mode == DSPT_SCHED_PASS) {
    if (!pool_idle(c)) return;
    int64_t room = c->cap;
    bool any = false;
    for (int id : ids) {
      if (!job_ok(c, id)) continue;
      JobRes* j = c->jobs[id];
      if (!eligible(c, j)) continue;
      int64_t npix = j->items();
      int64_t nf = std::min<int64_t>(j->frame_end - j->next_sched, j->acc_known + j->ring_frames - j->next_sched);
      if (c->pass_frames > 0) nf = std::min<int64_t>(nf, c->pass_frames);
      nf = std::min<int64_t>(nf, room / npix);
      if (nf <= 0) {
        if (any) continue;
        nf = 1;  // a frame larger than the pool streams through it
      }
      if (!append_frames(c, id, (int)nf)) break;
      any = true;
      room -= nf * npix;
      if (room <= 0) break;
    }
    return;
  }
  const int64_t target = c->next_known + (int64_t)(c->chunk_iters + 2) * c->cap;
  if (c->policy == DSPT_POLICY_RR) {
    const size_t n = ids.size();
    bool progress = true;
    while (c->sched_limit < target && progress) {
      progress = false;
      for (size_t k = 0; k < n && c->sched_limit < target; k++) {
        int id = ids[(c->rr + k) % n];
        if (!job_ok(c, id) || !eligible(c, c->jobs[id])) continue;
        if (!append_frames(c, id, 1)) return;
        progress = true;
      }
      c->rr = (c->rr + 1) % n;
    }
    return;
  }
  // FIFO (listed order) or least remaining work first: whole runs of frames
  // of one job before the next, so few scenes share the caches at a time.
  std::vector<int> order;
  for (int id : ids)
    if (job_ok(c, id)) order.push_back(id);
  if (c->policy == DSPT_POLICY_SRPT)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      const JobRes* x = c->jobs[a];
      const JobRes* y = c->jobs[b];
      return (int64_t)(x->frame_end - x->next_sched) * x->items() < (int64_t)(y->frame_end - y->next_sched) * y->items();
    });
  for (int id : order) {
    JobRes* j = c->jobs[id];
    while (c->sched_limit < target && eligible(c, j)) {
      int64_t want = (target - c->sched_limit + j->items() - 1) / j->items();
      int64_t nf = std::min<int64_t>({want, (int64_t)(j->frame_end - j->next_sched),
                                      (int64_t)(j->acc_known + j->ring_frames - j->next_sched)});
      if (nf <= 0 || !append_frames(c, id, (int)nf)) return;
    }
    if (c->sched_limit >= target) return;
  }
}

void consume_readback(dspt_ctx* c, int s, int64_t chunk) {
  if (c->cap <= 0) return;  // the pool was released after this chunk
  const Readback& rb = c->h_rb[s];
  c->next_known = rb.ctl.next_item;
  c->alive_known = rb.ctl.q[c->qsel_at[s]].paths;
  c->limit_known = c->limit_at[s];
  const int n = std::min<int>((int)c->jobs.size(), kMaxJobs);
  for (int i = 0; i < n; i++) {
    JobRes* j = c->jobs[i];
    if (!j || !j->alive || j->epoch > chunk) continue;
    j->acc_known = std::max(j->acc_known, std::min(rb.prog[i], j->frame_end));
  }
}

// Enqueues one chunk (schedule, chunk_iters bounces, fold), then waits for
// the previous chunk and takes its readback.
int run_chunk(dspt_ctx* c, const std::vector<int>& ids, bool schedule_new) {
  if (int r = ensure_pool(c)) return r;
  if (c->cap <= 0) return fail(DSPT_ERR_STATE, "path pool not allocated");
  if (int r = sync_job_table(c)) return r;
  if (schedule_new) schedule(c, ids);
  while (!c->tasks.empty()) {
    const PassTask& t = c->tasks.front();
    if (t.offset + (int64_t)t.nframes * t.npix > c->next_known) break;
    c->tasks.pop_front();
  }
  const int s = (int)(c->chunk_no & 1);
  const int nt = (int)c->tasks.size();
  PassTask* ht = c->h_tasks + (size_t)s * kMaxTasks;
  for (int i = 0; i < nt; i++) ht[i] = c->tasks[i];
  c->h_sched[s].limit = c->sched_limit;
  c->exposed_limit = c->sched_limit;
  c->qsel_at[s] = (c->qsel + c->chunk_iters) & 1;
  c->h_sched[s].ntasks = nt;
  c->limit_at[s] = c->sched_limit;
  int32_t* hj = c->h_joblist + (size_t)s * kMaxJobs;
  int nj = 0;
  int64_t max_npix = 0;
  for (size_t i = 0; i < c->jobs.size() && nj < kMaxJobs; i++) {
    JobRes* j = c->jobs[i];
    if (j && j->alive && j->in_flight()) {
      hj[nj++] = (int32_t)i;
      max_npix = std::max(max_npix, j->items());
    }
  }
  // the jobs with paths in the pool are those listed (progress is read back
  // late, so a listed job may already be done); one of them: its index
  c->h_sched[s].solo = nj == 1 ? hj[0] : -1;
  int nscenes = 0;  // distinct scenes among them (counted up to 32)
  int32_t sc[32];
  for (int a = 0; a < nj && nscenes < 32; a++) {
    const int32_t id = c->jobs[hj[a]]->scene;
    bool seen = false;
    for (int b = 0; b < nscenes && !seen; b++) seen = sc[b] == id;
    if (!seen) sc[nscenes++] = id;
  }
  int jbits = 0;  // ray-order key bits for the scene id: enough for those scenes, at most 5
  while (jbits < 5 && (1 << jbits) < nscenes) jbits++;
  c->h_sched[s].jbits = jbits;
  if (nt > 0) HIPCHK(hipMemcpyAsync(c->d_tasks, ht, (size_t)nt * sizeof(PassTask), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_sched, c->h_sched + s, sizeof(StreamSched), hipMemcpyHostToDevice, c->stream));
  if (nj > 0) HIPCHK(hipMemcpyAsync(c->d_joblist, hj, (size_t)nj * 4, hipMemcpyHostToDevice, c->stream));
  const int64_t P = c->cap;
  const int shade_blocks = (int)std::min<int64_t>(grid_for(P, 256), c->shade_blocks_max);
  const int trace_blocks = c->trace_blocks;
  StreamCtl* ctl = c->d_ctl;
  const int32_t P32 = (int32_t)P;
  // two-level traversal only while a job on a two-level scene is live
  bool inst = false;
  for (const JobRes* j : c->jobs)
    if (j && j->alive && c->scenes[j->scene] && c->scenes[j->scene]->top >= 0) inst = true;
  // the shading kernel's debug instance only while a job renders a debug view
  bool dbg = false;
  for (const JobRes* j : c->jobs)
    if (j && j->alive && j->F.debug_mode != 0.0f) dbg = true;
  const bool order = c->ray_order == DSPT_ORDER_BUCKETS || c->ray_order == DSPT_ORDER_SHADE;
  const bool shade_order = c->ray_order == DSPT_ORDER_SHADE;
  const int32_t* perm = order ? c->order_perm : nullptr;
  const int order_blocks = (int)std::min<int64_t>(DSPT_ORDER_BLOCKS, grid_for(P, DSPT_ORDER_THREADS));
  // chunk_iters bounces; the state buffers alternate, starting at q0
  auto bounces = [&](int q0) {
    for (int k = 0; k < c->chunk_iters; k++) {
      const int a = (q0 + k) & 1;
      const PathState cur = c->ps[a], nxt = c->ps[a ^ 1];
      timed(c, DSPT_K_RAYGEN, [&] {
        k_plan<<<1, 64, 0, c->stream>>>(ctl, c->d_sched, a, P32, order ? c->order_bins : nullptr,
                                        shade_order ? c->shade_bins : nullptr);
        k_regen<<<shade_blocks, 256, 0, c->stream>>>(cur, c->d_jobs, c->d_tasks, c->d_sched, ctl, c->cnt);
      });
      if (order)
        timed(c, DSPT_K_ORDER, [&] {
          k_ray_keys<<<order_blocks, DSPT_ORDER_THREADS, 0, c->stream>>>(c->d_jobs, cur, ctl, c->order_keys, c->order_bins,
                                                                         c->d_sched);
          k_bin_scan<<<1, 1024, 0, c->stream>>>(c->order_bins);
          k_bin_scatter<<<order_blocks, DSPT_ORDER_THREADS, 0, c->stream>>>(c->order_keys, &ctl->q_base, c->order_bins, c->order_perm,
                                                                            nullptr, nullptr);
        });
      timed(c, DSPT_K_TRACE_EXT, [&] {
        if (inst)
          k_trace_ext<true><<<trace_blocks, kTraceBlock, 0, c->stream>>>(c->d_jobs, cur, &ctl->q[a].paths, &ctl->ext_next,
                                                                        c->spill, c->cnt, perm, &ctl->q_base);
        else
          k_trace_ext<false><<<trace_blocks, kTraceBlock, 0, c->stream>>>(c->d_jobs, cur, &ctl->q[a].paths, &ctl->ext_next,
                                                                         c->spill, c->cnt, perm, &ctl->q_base);
      });
      if (shade_order)
        timed(c, DSPT_K_ORDER, [&] {
          k_hit_keys<<<order_blocks, DSPT_ORDER_THREADS, 0, c->stream>>>(c->d_jobs, cur, &ctl->q[a].paths, c->order_keys,
                                                                         c->shade_bins, c->d_sched);
          k_bin_scan<<<1, 1024, 0, c->stream>>>(c->shade_bins);
          k_bin_scatter<<<order_blocks, DSPT_ORDER_THREADS, 0, c->stream>>>(c->order_keys, &ctl->q[a].paths, c->shade_bins,
                                                             c->shade_perm, perm, &ctl->q_base);  // buckets in ray order
        });
      timed(c, DSPT_K_SHADE, [&] {
        const int32_t* sperm = shade_order ? c->shade_perm : perm;
        const int32_t* nsorted = shade_order ? &ctl->q[a].paths : &ctl->q_base;
        if (dbg)
          k_shade<true><<<shade_blocks, 256, 0, c->stream>>>(c->d_jobs, cur, nxt, &ctl->q[a].paths, &ctl->q[a ^ 1],
                                                              c->shq, c->cnt, sperm, nsorted);
        else
          k_shade<false><<<shade_blocks, 256, 0, c->stream>>>(c->d_jobs, cur, nxt, &ctl->q[a].paths, &ctl->q[a ^ 1],
                                                               c->shq, c->cnt, sperm, nsorted);
      });
      timed(c, DSPT_K_TRACE_SHADOW, [&] {
        if (inst)
          k_trace_shadow<true><<<trace_blocks, kTraceBlock, 0, c->stream>>>(c->d_jobs, nxt, c->shq, &ctl->q[a ^ 1].shadow,
                                                                           &ctl->sh_next, c->spill, c->cnt);
        else
          k_trace_shadow<false><<<trace_blocks, kTraceBlock, 0, c->stream>>>(c->d_jobs, nxt, c->shq, &ctl->q[a ^ 1].shadow,
                                                                            &ctl->sh_next, c->spill, c->cnt);
      });
    }
  };
  const int q0 = c->qsel;
  if (c->use_graphs && !c->timing) {
    // every launch argument of the loop is fixed by (pool, grids, chunk
    // length, q0, inst, dbg): the work sizes live in device counters
    const std::array<int64_t, 6> key{P, (int64_t)(uintptr_t)c->ps[0].org, trace_blocks, shade_blocks, c->chunk_iters,
                                     c->ray_order};
    if (key != c->graph_key) {
      HIPCHK(hipStreamSynchronize(c->stream));  // no launch of an old graph in flight
      drop_graphs(c);
      c->graph_key = key;
    }
    hipGraphExec_t& ge = c->bounce_graph[q0][inst ? 1 : 0][dbg ? 1 : 0];
    if (!ge) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      bounces(q0);
      const hipError_t ce = hipStreamEndCapture(c->stream, &g);
      if (ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(DSPT_ERR_DEVICE, std::string("bounce-loop capture: ") + hipGetErrorString(ce));
      }
      const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        ge = nullptr;
        return fail(DSPT_ERR_DEVICE, std::string("bounce-loop graph: ") + hipGetErrorString(ie));
      }
    }
    HIPCHK(hipGraphLaunch(ge, c->stream));
  } else {
    bounces(q0);
  }
  c->qsel = (q0 + c->chunk_iters) & 1;
  if (nj > 0) {
    timed(c, DSPT_K_ACCUM, [&] {
      k_accum_ready<<<dim3(grid_for(max_npix, 256), (unsigned)nj), 256, 0, c->stream>>>(c->d_jobs, c->d_joblist, c->d_prog);
      k_advance<<<1, 256, 0, c->stream>>>(c->d_jobs, c->d_joblist, nj, c->d_prog);
    });
  }
  HIPCHK(hipGetLastError());
  const int njobs = std::min<int>((int)c->jobs.size(), kMaxJobs);
  HIPCHK(hipMemcpyAsync(&c->h_rb[s].ctl, ctl, sizeof(StreamCtl), hipMemcpyDeviceToHost, c->stream));
  if (njobs > 0) HIPCHK(hipMemcpyAsync(c->h_rb[s].prog, c->d_prog, (size_t)njobs * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipEventRecord(c->ev_chunk[s], c->stream));
  if (c->have_prev) {
    HIPCHK(hipEventSynchronize(c->ev_chunk[s ^ 1]));
    consume_readback(c, s ^ 1, c->chunk_no - 1);
  }
  c->have_prev = true;
  c->chunk_no++;
  return 0;
}

// Waits for the last enqueued chunk and takes its readback.
int settle(dspt_ctx* c) {
  if (!c->have_prev) return 0;
  const int s = (int)((c->chunk_no - 1) & 1);
  HIPCHK(hipEventSynchronize(c->ev_chunk[s]));
  consume_readback(c, s, c->chunk_no - 1);
  return 0;
}

// Progress fingerprint of the pool as of the latest readback.
int64_t progress_mark(const dspt_ctx* c) {
  int64_t m = c->next_known * 1000003 + c->alive_known;
  for (const JobRes* j : c->jobs)
    if (j && j->alive) m = m * 31 + j->acc_known;
  return m;
}

// Runs chunks until every listed job has folded its whole frame range. A
// pool with work always retires a path within max_bounces + 1 bounces, so
// kStallChunks chunks without any change in the readback is an error.
int run_until_complete(dspt_ctx* c, const std::vector<int>& ids) {
  auto all_done = [&] {
    for (int id : ids)
      if (job_ok(c, id) && !c->jobs[id]->cancel && !c->jobs[id]->complete()) return false;
    return true;
  };
  int64_t mark = progress_mark(c), still = 0;
  while (!all_done()) {
    if (int r = run_chunk(c, ids, true)) return r;
    int64_t m = progress_mark(c);
    still = (m == mark) ? still + 1 : 0;
    mark = m;
    if (still > kStallChunks) return fail(DSPT_ERR_STATE, "scheduler made no progress");
  }
  return 0;
}

// Runs chunks without scheduling until job `id` has nothing in flight.
int drain_job(dspt_ctx* c, int id) {
  JobRes* j = c->jobs[id];
  if (j->in_flight()) {
    if (int r = settle(c)) return r;
  }
  int64_t mark = progress_mark(c), still = 0;
  while (j->in_flight()) {
    if (int r = run_chunk(c, {}, false)) return r;
    int64_t m = progress_mark(c);
    still = (m == mark) ? still + 1 : 0;
    mark = m;
    if (still > kStallChunks) return fail(DSPT_ERR_STATE, "drain made no progress");
  }
  return 0;
}

#ifndef DSPT_PACK_INNER_NODES
#define DSPT_PACK_INNER_NODES 1
#endif
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
// Checks shared by every scene constructor; returns the largest texture id
// a material references in *max_tex.
int validate_geometry(const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* mats, int32_t K,
                      int32_t* max_tex_out) {
  if (!vertices || !tri_idx || !mats) return fail(DSPT_ERR_ARG, "NULL scene buffer");
  if (V <= 0 || T <= 0 || K <= 0) return fail(DSPT_ERR_ARG, "empty scene");
  if (T >= (1LL << 24)) return fail(DSPT_ERR_UNSUPPORTED, "scene exceeds float-encoded index range");
  for (int64_t i = 0; i < 3 * T; i++)
    if ((int64_t)tri_idx[i] >= V) return fail(DSPT_ERR_ARG, "triangle index " + std::to_string(i) + " out of range");
  for (int64_t v = 0; v < V; v++) {
    float mi = vertices[v * 20 + 3];
    if (!(mi >= 0.0f) || mi >= (float)K) return fail(DSPT_ERR_ARG, "vertex " + std::to_string(v) + ": material index out of range");
  }
  int32_t max_tex = -1;
  for (int32_t k = 0; k < K; k++)
    for (int i : {39, 44, 45, 46, 47, 48, 49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 59, 63}) {
      const float id = mats[(size_t)k * 68 + i];
      if (!std::isfinite(id) || id >= 16777216.0f)
        return fail(DSPT_ERR_ARG, "material " + std::to_string(k) + ": texture id is not a finite value below 2^24");
      if (id >= 0.0f) max_tex = std::max(max_tex, (int32_t)id);
    }
  *max_tex_out = max_tex;
  return 0;
}

// Bits of the ray-order grid per axis from the scene's root box (the node
// at `root`, read back once): each of the 12 bits halves the longest cell
// side (ties: x, then y), so the cells of a flat scene spread over its
// extent. Packed as x | y << 4 | z << 8 for ray_key.
int scene_key_bits(dspt_ctx* c, SceneRes* s, int64_t root) {
  float nd[16];
  HIPCHK(hipMemcpy(nd, s->nodes + (size_t)root * 4, sizeof(nd), hipMemcpyDeviceToHost));
  float ex = std::max(nd[4], nd[12]) - std::min(nd[0], nd[8]);
  float ey = std::max(nd[5], nd[13]) - std::min(nd[1], nd[9]);
  float ez = std::max(nd[6], nd[14]) - std::min(nd[2], nd[10]);
  int bx = 0, by = 0, bz = 0;
  for (int i = 0; i < 12; i++) {
    if (ex >= ey && ex >= ez) { bx++; ex *= 0.5f; }
    else if (ey >= ez) { by++; ey *= 0.5f; }
    else { bz++; ez *= 0.5f; }
  }
  s->key_bits = bx | (by << 4) | (bz << 8);
  return 0;
}

// Scene whose BVH is built on the GPU (gpu_bvh_build) from the uploaded
// vertex stream; no host-side node array. With mats == NULL (K = 0) only the
// geometry is checked and uploaded (dspt_build_bvh_gpu).
int make_scene_gpu_bvh(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* mats,
                       int32_t K, int32_t max_leaf, int32_t method, SceneRes** out, double* build_ms) {
  int32_t max_tex = -1;
  if (mats) {
    if (int r = validate_geometry(vertices, V, tri_idx, T, mats, K, &max_tex)) return r;
  } else {
    if (!vertices || !tri_idx) return fail(DSPT_ERR_ARG, "NULL scene buffer");
    if (V <= 0 || T <= 0) return fail(DSPT_ERR_ARG, "empty scene");
    if (T >= (1LL << 24)) return fail(DSPT_ERR_UNSUPPORTED, "scene exceeds float-encoded index range");
    for (int64_t i = 0; i < 3 * T; i++)
      if ((int64_t)tri_idx[i] >= V) return fail(DSPT_ERR_ARG, "triangle index " + std::to_string(i) + " out of range");
    K = 0;
  }
  HIPCHK(hipSetDevice(c->device));
  SceneRes* s = new SceneRes();
  s->mem.lim = s->tex_mem.lim = &c->lim;
  Allocs& A = s->mem;
  int r = 0;
  float4* big_nodes = nullptr;
  int32_t* d_bvh = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto bail = [&](int code) {
    if (big_nodes) (void)hipFree(big_nodes);
    if (d_bvh) (void)hipFree(d_bvh);
    if (e0) c->ev_pool.push_back(e0);
    if (e1) c->ev_pool.push_back(e1);
    A.release();
    delete s;
    return code;
  };
  auto chk = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? 0 : bail(fail(DSPT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)));
  };
  if ((r = A.get(&s->verts, (size_t)V * 5)) || (r = A.get(&s->tris, (size_t)T * 3)) ||
      (K > 0 && (r = A.get(&s->mats, (size_t)K * 17))) || (r = A.get(&s->tri, (size_t)T * 3)))
    return bail(r);
  if (hipMalloc(&big_nodes, (size_t)(2 * T - 1) * 64) != hipSuccess || hipMalloc(&d_bvh, (size_t)T * 4) != hipSuccess)
    return bail(fail(DSPT_ERR_MEMORY, "bvh build buffers"));
  if ((r = chk(hipMemcpyAsync(s->verts, vertices, (size_t)V * 80, hipMemcpyHostToDevice, c->stream), "vertex upload")) ||
      (r = chk(hipMemcpyAsync(s->tris, tri_idx, (size_t)T * 12, hipMemcpyHostToDevice, c->stream), "index upload")) ||
      (K > 0 && (r = chk(hipMemcpyAsync(s->mats, mats, (size_t)K * 272, hipMemcpyHostToDevice, c->stream), "material upload"))) ||
      (r = chk(hipStreamSynchronize(c->stream), "upload")))
    return r;
  e0 = take_event(c);
  e1 = take_event(c);
  (void)hipEventRecord(e0, c->stream);
  int64_t N = 0;
  std::string err;
  if (int rc = gpu_bvh_build(c->stream, s->verts, s->tris, (int32_t)T, max_leaf, method, big_nodes, d_bvh, &N, err))
    return bail(fail(rc, err));
  (void)hipEventRecord(e1, c->stream);
  if ((r = A.get(&s->nodes, (size_t)N * 4))) return bail(r);
  if ((r = chk(hipMemcpyAsync(s->nodes, big_nodes, (size_t)N * 64, hipMemcpyDeviceToDevice, c->stream), "node copy")))
    return r;
  k_prep_tris<<<grid_for(T, 256), 256, 0, c->stream>>>(d_bvh, s->tris, s->verts, (int32_t)T, (int32_t)T, s->tri);
  if (mats) {  // a resident scene: child codes (the host-readback build keeps the bvh_gpu fields)
    if ((r = A.get(&s->d_flag, 1)) || (r = A.get(&s->tri_mat, T)) || (r = A.get(&s->tri_verts, (size_t)T * 15)))
      return bail(r);
    k_tri_mat<<<grid_for(T, 256), 256, 0, c->stream>>>(s->tris, s->verts, (int32_t)T, s->tri_mat);
    k_tri_verts<<<grid_for(T * 15, 256), 256, 0, c->stream>>>(s->tris, s->verts, (int32_t)T, s->tri_verts);
    k_encode_children<<<grid_for(N, 256), 256, 0, c->stream>>>(s->nodes, (int32_t)N, (int32_t)N);
    k_fill_i32<<<1, 64, 0, c->stream>>>(s->d_flag, 1, 1);
    k_boxes_finite<<<grid_for(N, 256), 256, 0, c->stream>>>(s->nodes, (int32_t)N, s->d_flag);
  }
  if ((r = chk(hipGetLastError(), "scene kernels")) || (r = chk(hipStreamSynchronize(c->stream), "scene build"))) return r;
  if (mats && (r = chk(hipMemcpy(&s->fast_slab, s->d_flag, 4, hipMemcpyDeviceToHost), "scene flag"))) return r;
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  c->ev_pool.push_back(e0);
  c->ev_pool.push_back(e1);
  e0 = e1 = nullptr;
  if (build_ms) *build_ms = ms;
  (void)hipFree(big_nodes);
  (void)hipFree(d_bvh);
  big_nodes = nullptr;
  d_bvh = nullptr;
  s->N = (int32_t)N; s->M = (int32_t)T; s->T = (int32_t)T; s->K = K; s->V = V;
  s->max_tex_id = max_tex;
  if (mats && (r = scene_key_bits(c, s, 0))) return bail(r);
  *out = s;
  return 0;
}

// encode_limit: nodes below it get child codes (k_encode_children); -1 = all
// N (a two-level scene passes the first instance-record slot).
// The engine's copy of a one-level node array with the inner nodes first,
// in their given order, and the leaf nodes behind them. The traversal never
// fetches a well-formed leaf (its parent holds its box and the child field
// becomes a leaf code), so in a depth-first array about half of every
// 128-byte line of fetched nodes was a leaf nobody reads; packed, a line
// holds two inner nodes, and a node is followed by its first inner child.
// A relabelling of node ids: child fields are rewritten as the traversal
// reads them ((int) of the float; out-of-range values unchanged), the root
// stays at 0, leaves keep their fields, so visit order, steps and counters
// are those of the given array, malformed trees included. Skipped for a
// leaf root and for arrays whose ids a float cannot hold exactly.
static bool pack_inner_nodes(const float* nodes, int64_t N, std::vector<float>& out) {
  if (N > (1 << 24)) return false;
  auto is_leaf = [&](int64_t n) { return (int)nodes[n * 16 + 11] > 0; };  // (int) count > 0, as the kernels test it
  if (is_leaf(0)) return false;
  std::vector<int32_t> id((size_t)N);
  int32_t k = 0;
  for (int64_t n = 0; n < N; n++)
    if (!is_leaf(n)) id[n] = k++;
  for (int64_t n = 0; n < N; n++)
    if (is_leaf(n)) id[n] = k++;
  out.resize((size_t)N * 16);
  for (int64_t n = 0; n < N; n++) {
    float* o = &out[(size_t)id[n] * 16];
    std::memcpy(o, nodes + n * 16, 64);
    if (is_leaf(n)) continue;
    for (int f : {3, 7}) {
      const int child = (int)o[f];
      if (child >= 0 && child < N) o[f] = (float)id[child];
    }
  }
  return true;
}

int make_scene(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* nodes, int64_t N,
               const int32_t* bvh_idx, int64_t M, const float* mats, int32_t K, SceneRes** out, bool own_nodes = false,
               int64_t encode_limit = -1) {
  if (!vertices || !tri_idx || !nodes || !bvh_idx || !mats) return fail(DSPT_ERR_ARG, "NULL scene buffer");
  if (N <= 0 || M <= 0) return fail(DSPT_ERR_ARG, "empty BVH");
  if (N >= (1LL << 30) || M >= (1LL << 30)) return fail(DSPT_ERR_UNSUPPORTED, "scene exceeds float-encoded index range");
  int32_t max_tex = -1;
  if (int r = validate_geometry(vertices, V, tri_idx, T, mats, K, &max_tex)) return r;
  if (!own_nodes)
    if (int r = validate_nodes(nodes, N)) return r;
  HIPCHK(hipSetDevice(c->device));
  SceneRes* s = new SceneRes();
  s->mem.lim = s->tex_mem.lim = &c->lim;
  Allocs& A = s->mem;
  int r = 0;
  int32_t* d_bvh = nullptr;
  auto bail = [&](int code) {
    if (d_bvh) (void)hipFree(d_bvh);
    A.release();
    delete s;
    return code;
  };
  auto chk = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? 0 : bail(fail(DSPT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)));
  };
  if ((r = A.get(&s->nodes, (size_t)N * 4)) || (r = A.get(&s->tri, (size_t)M * 3)) || (r = A.get(&s->verts, (size_t)V * 5)) ||
      (r = A.get(&s->tris, (size_t)T * 3)) || (r = A.get(&s->mats, (size_t)K * 17)))
    return bail(r);
  if (hipMalloc(&d_bvh, (size_t)M * sizeof(int32_t)) != hipSuccess) {
    d_bvh = nullptr;
    return bail(fail(DSPT_ERR_MEMORY, "bvh index staging"));
  }
  {
    // the packed copy is uploaded and waited for here, so no copy reads it after it is freed
    std::vector<float> packed;
    const bool pack = encode_limit < 0 && DSPT_PACK_INNER_NODES && pack_inner_nodes(nodes, N, packed);
    if ((r = chk(hipMemcpyAsync(s->nodes, pack ? packed.data() : nodes, (size_t)N * 64, hipMemcpyHostToDevice, c->stream),
                 "node upload")))
      return r;
    if (pack && (r = chk(hipStreamSynchronize(c->stream), "node upload"))) return r;
  }
  if (
      (r = chk(hipMemcpyAsync(s->verts, vertices, (size_t)V * 80, hipMemcpyHostToDevice, c->stream), "vertex upload")) ||
      (r = chk(hipMemcpyAsync(s->tris, tri_idx, (size_t)T * 12, hipMemcpyHostToDevice, c->stream), "index upload")) ||
      (r = chk(hipMemcpyAsync(s->mats, mats, (size_t)K * 272, hipMemcpyHostToDevice, c->stream), "material upload")) ||
      (r = chk(hipMemcpyAsync(d_bvh, bvh_idx, (size_t)M * 4, hipMemcpyHostToDevice, c->stream), "bvh index upload")))
    return r;
  k_prep_tris<<<grid_for(M, 256), 256, 0, c->stream>>>(d_bvh, s->tris, s->verts, (int32_t)M, (int32_t)T, s->tri);
  const int64_t lim = encode_limit < 0 ? N : std::min(encode_limit, N);
  k_encode_children<<<grid_for(lim, 256), 256, 0, c->stream>>>(s->nodes, (int32_t)N, (int32_t)lim);
  if ((r = A.get(&s->d_flag, 1)) || (r = A.get(&s->tri_mat, T)) || (r = A.get(&s->tri_verts, (size_t)T * 15)))
    return bail(r);
  k_fill_i32<<<1, 64, 0, c->stream>>>(s->d_flag, 1, 1);
  k_boxes_finite<<<grid_for(lim, 256), 256, 0, c->stream>>>(s->nodes, (int32_t)lim, s->d_flag);
  k_tri_mat<<<grid_for(T, 256), 256, 0, c->stream>>>(s->tris, s->verts, (int32_t)T, s->tri_mat);
  k_tri_verts<<<grid_for(T * 15, 256), 256, 0, c->stream>>>(s->tris, s->verts, (int32_t)T, s->tri_verts);
  if ((r = chk(hipGetLastError(), "scene kernels")) || (r = chk(hipStreamSynchronize(c->stream), "scene upload"))) return r;
  if ((r = chk(hipMemcpy(&s->fast_slab, s->d_flag, 4, hipMemcpyDeviceToHost), "scene flag"))) return r;
  (void)hipFree(d_bvh);
  d_bvh = nullptr;
  s->N = (int32_t)N; s->M = (int32_t)M; s->T = (int32_t)T; s->K = K; s->V = V;
  s->max_tex_id = max_tex;
  if ((r = scene_key_bits(c, s, 0))) return bail(r);
  *out = s;
  return 0;
}

// Two-level scene: the engine builds one bottom-level BVH per mesh (the
// mesh's triangles, object space), concatenates them (child indices and leaf
// ranges offset), and a top-level BVH over the instances' world bounds with
// one instance per leaf.
int make_scene_instanced(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T,
                         const int64_t* mesh_first, const int64_t* mesh_count, int32_t num_meshes, const float* inst_o2w,
                         const int32_t* inst_mesh, int32_t num_inst, const float* mats, int32_t K, SceneRes** out) {
  if (!vertices || !tri_idx || !mesh_first || !mesh_count || !inst_o2w || !inst_mesh || !mats)
    return fail(DSPT_ERR_ARG, "NULL scene buffer");
  if (V <= 0 || T <= 0 || num_meshes <= 0 || num_inst <= 0 || K <= 0) return fail(DSPT_ERR_ARG, "empty scene");
  if (T >= (1LL << 24)) return fail(DSPT_ERR_UNSUPPORTED, "meshes exceed 2^24 triangles");
  for (int32_t m = 0; m < num_meshes; m++)
    if (mesh_first[m] < 0 || mesh_count[m] <= 0 || mesh_first[m] + mesh_count[m] > T)
      return fail(DSPT_ERR_ARG, "mesh " + std::to_string(m) + ": triangle range out of bounds");
  for (int64_t i = 0; i < 3 * T; i++)
    if ((int64_t)tri_idx[i] >= V) return fail(DSPT_ERR_ARG, "triangle index " + std::to_string(i) + " out of range");
  // bottom-level BVHs, concatenated
  std::vector<float> nodes;
  std::vector<int32_t> bidx;
  std::vector<int32_t> root(num_meshes);
  std::vector<std::vector<uint32_t>> mverts(num_meshes);  // vertices each mesh uses
  std::vector<float> pos(3 * (size_t)V);
  for (int64_t v = 0; v < V; v++)
    for (int k = 0; k < 3; k++) pos[3 * v + k] = vertices[v * 20 + k];
  for (int32_t m = 0; m < num_meshes; m++) {
    std::vector<float> mn;
    std::vector<int32_t> mi;
    std::string err;
    if (dspt_bvh_build_impl(pos.data(), V, tri_idx + 3 * mesh_first[m], mesh_count[m], 4, mn, mi, err))
      return fail(DSPT_ERR_ARG, "mesh " + std::to_string(m) + ": " + err);
    const int64_t noff = (int64_t)nodes.size() / 16, ioff = (int64_t)bidx.size();
    if (noff + (int64_t)mn.size() / 16 >= (1LL << 24) || ioff + (int64_t)mi.size() >= (1LL << 24))
      return fail(DSPT_ERR_UNSUPPORTED, "bottom-level BVHs exceed float-encoded index range");
    root[m] = (int32_t)noff;
    for (size_t i = 0; i < mn.size() / 16; i++) {
      float* o = &mn[i * 16];
      if (o[11] > 0.0f) o[15] += (float)ioff;       // leaf: first
      else { o[3] += (float)noff; o[7] += (float)noff; }  // inner: children
    }
    for (int32_t& t : mi) t += (int32_t)mesh_first[m];
    nodes.insert(nodes.end(), mn.begin(), mn.end());
    bidx.insert(bidx.end(), mi.begin(), mi.end());
    std::vector<uint32_t>& mv = mverts[m];
    mv.assign(tri_idx + 3 * mesh_first[m], tri_idx + 3 * (mesh_first[m] + mesh_count[m]));
    std::sort(mv.begin(), mv.end());
    mv.erase(std::unique(mv.begin(), mv.end()), mv.end());
  }
  // instances: world -> object inverse, world bounds (slightly padded so the
  // top level never culls what the object-space test would hit)
  std::vector<InstanceDesc> inst(num_inst);
  std::vector<float> boxes(6 * (size_t)num_inst);
  for (int32_t i = 0; i < num_inst; i++) {
    const int32_t m = inst_mesh[i];
    if (m < 0 || m >= num_meshes) return fail(DSPT_ERR_ARG, "instance " + std::to_string(i) + ": mesh out of range");
    const float* a = inst_o2w + 12 * (size_t)i;
    double M[3][4];
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 4; k++) {
        if (!std::isfinite(a[4 * r + k])) return fail(DSPT_ERR_ARG, "instance " + std::to_string(i) + ": non-finite transform");
        M[r][k] = a[4 * r + k];
      }
    const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                       M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
    if (!(std::fabs(det) > 1e-12)) return fail(DSPT_ERR_ARG, "instance " + std::to_string(i) + ": singular transform");
    double Inv[3][3];
    Inv[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) / det;
    Inv[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
    Inv[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
    Inv[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) / det;
    Inv[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
    Inv[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
    Inv[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) / det;
    Inv[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
    Inv[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
    InstanceDesc& D = inst[i];
    for (int r = 0; r < 3; r++) {
      const double tr = -(Inv[r][0] * M[0][3] + Inv[r][1] * M[1][3] + Inv[r][2] * M[2][3]);
      D.w2o[r] = make_float4((float)Inv[r][0], (float)Inv[r][1], (float)Inv[r][2], (float)tr);
      D.o2w[r] = make_float4(a[4 * r], a[4 * r + 1], a[4 * r + 2], a[4 * r + 3]);
    }
    D.root = root[m];
    D.mesh = m;
    D.pad0 = D.pad1 = 0;
    D.pad2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // world box of the transformed vertices (tighter than the transformed
    // object box for rotated instances)
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t vi : mverts[m]) {
      const float* p = &pos[3 * (size_t)vi];
      for (int r = 0; r < 3; r++) {
        const double w = M[r][0] * p[0] + M[r][1] * p[1] + M[r][2] * p[2] + M[r][3];
        lo[r] = std::min(lo[r], w);
        hi[r] = std::max(hi[r], w);
      }
    }
    for (int r = 0; r < 3; r++) {
      const double pad = 1e-5 * (hi[r] - lo[r]) + 1e-5 * std::max(std::fabs(lo[r]), std::fabs(hi[r])) + 1e-6;
      boxes[6 * (size_t)i + r] = (float)(lo[r] - pad);
      boxes[6 * (size_t)i + 3 + r] = (float)(hi[r] + pad);
    }
  }
  std::vector<float> tn;
  std::vector<int32_t> ti;
  std::string err;
  if (dspt_bvh_build_boxes_impl(boxes.data(), num_inst, 1, tn, ti, err)) return fail(DSPT_ERR_ARG, "top level: " + err);
  // one node array: bottom levels, the top level's inner nodes (children
  // offset; a child that is a one-instance leaf points at the instance
  // record instead, so entering an instance costs no leaf fetch), instance
  // records
  const int64_t nb = (int64_t)nodes.size() / 16, TN = (int64_t)tn.size() / 16;
  const int64_t N = nb + TN + 2 * (int64_t)num_inst;
  if (N >= (1LL << 24)) return fail(DSPT_ERR_UNSUPPORTED, "two-level node array exceeds float-encoded index range");
  const int64_t inst_base = nb + TN;
  auto remap = [&](int64_t child, int64_t* out) {
    if (child < 0 || child >= TN) return false;
    const float* o = &tn[child * 16];
    if (o[11] > 0.0f) {
      const int first = (int)o[15];
      if (o[11] != 1.0f || first < 0 || first >= (int)ti.size()) return false;
      *out = inst_base + 2 * (int64_t)ti[first];
    } else {
      *out = nb + child;
    }
    return true;
  };
  for (int64_t i = 0; i < TN; i++) {
    float* o = &tn[i * 16];
    if (o[11] > 0.0f) continue;
    int64_t l, r;
    if (!remap((int64_t)o[3], &l) || !remap((int64_t)o[7], &r)) return fail(DSPT_ERR_STATE, "top level: malformed node");
    o[3] = (float)l;
    o[7] = (float)r;
  }
  int64_t top = 0;
  if (!remap(0, &top)) return fail(DSPT_ERR_STATE, "top level: malformed root");
  nodes.insert(nodes.end(), tn.begin(), tn.end());
  const size_t ib = nodes.size();
  nodes.resize(ib + (size_t)num_inst * 32);
  std::memcpy(nodes.data() + ib, inst.data(), (size_t)num_inst * sizeof(InstanceDesc));
  SceneRes* s = nullptr;
  if (int r = make_scene(c, vertices, V, tri_idx, T, nodes.data(), N, bidx.data(), (int64_t)bidx.size(), mats, K, &s, true,
                         inst_base))
    return r;
  s->top = (int32_t)top;
  s->inst_base = (int32_t)inst_base;
  s->I = num_inst;
  if (int r = scene_key_bits(c, s, top)) {
    s->mem.release();
    delete s;
    return r;
  }
  *out = s;
  return 0;
}

int make_ibl(dspt_ctx* c, const float* rgba, int32_t w, int32_t h, IblRes** out) {
  if (!rgba || w <= 0 || h <= 0) return fail(DSPT_ERR_ARG, "invalid IBL image");
  HIPCHK(hipSetDevice(c->device));
  // owns the new IBL until it is handed out; every early return frees it
  struct Owner {
    IblRes* b;
    ~Owner() {
      if (b) { b->mem.release(); delete b; }
    }
  } own{new IblRes()};
  IblRes* b = own.b;
  b->mem.lim = &c->lim;
  size_t n = (size_t)w * h;
  Allocs& A = b->mem;
  int r = 0;
  if ((r = A.get(&b->env, n)) || (r = A.get(&b->pdf, n)) || (r = A.get(&b->cdf, n)) || (r = A.get(&b->ypdf, (size_t)h)) ||
      (r = A.get(&b->ycdf, (size_t)h)) || (r = A.get(&b->guide, (size_t)h * (w + 1))) || (r = A.get(&b->yguide, (size_t)h + 1)))
    return r;
  ScopedAllocs tmp;  // the build's temporaries, freed on every return
  tmp.lim = &c->lim;
  float* lum = nullptr;
  float* rowsum = nullptr;
  double* total = nullptr;
  if ((r = tmp.get(&lum, n)) || (r = tmp.get(&rowsum, (size_t)h)) || (r = tmp.get(&total, 1))) return r;
  HIPCHK(hipMemcpyAsync(b->env, rgba, n * 16, hipMemcpyHostToDevice, c->stream));
  k_ibl_luminance<<<grid_for((int64_t)n, 256), 256, 0, c->stream>>>(b->env, (int64_t)n, lum);
  k_ibl_rows<<<grid_for(h, 64), 64, 0, c->stream>>>(lum, w, h, b->pdf, b->cdf, rowsum);
  k_ibl_marginal<<<1, 64, 0, c->stream>>>(rowsum, h, b->ypdf, b->ycdf, total);
  k_ibl_guides<<<grid_for((int64_t)h * (w + 1) + h + 1, 256), 256, 0, c->stream>>>(b->cdf, b->ycdf, w, h, b->guide, b->yguide);
  HIPCHK(hipGetLastError());
  double tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, total, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  b->w = w; b->h = h; b->total = tot;
  own.b = nullptr;
  *out = b;
  return 0;
}

template <class T>
int put_slot(std::vector<T*>& v, T* x, int want = -1) {
  if (want >= 0) {
    if ((int)v.size() <= want) v.resize(want + 1, nullptr);
    v[want] = x;
    return want;
  }
  for (size_t i = 0; i < v.size(); i++)
    if (!v[i]) { v[i] = x; return (int)i; }
  v.push_back(x);
  return (int)v.size() - 1;
}

int check_job_frame(dspt_ctx* c, const FrameParams& F, int scene, int ibl) {
  if (!(F.max_bounces >= 0.0f) || F.max_bounces > 1024.0f) return fail(DSPT_ERR_ARG, "max_bounces out of range");
  if (scene < 0 || scene >= (int)c->scenes.size() || !c->scenes[scene]) return fail(DSPT_ERR_ARG, "unknown scene id");
  if (c->scenes[scene]->max_tex_id >= c->scenes[scene]->num_tex)
    return fail(DSPT_ERR_STATE, "a material references texture " + std::to_string(c->scenes[scene]->max_tex_id) +
                                    " but the scene has " + std::to_string(c->scenes[scene]->num_tex) +
                                    " textures (dspt_scene_set_textures)");
  bool needs_ibl = F.use_ibl > 0.0f || F.bg_from_ibl > 0.0f;
  const IblRes* b = (ibl >= 0 && ibl < (int)c->ibls.size()) ? c->ibls[ibl] : nullptr;
  if (ibl >= 0 && !b) return fail(DSPT_ERR_ARG, "unknown IBL id");
  if (needs_ibl && !b) return fail(DSPT_ERR_STATE, "IBL enabled but no IBL uploaded");
  if (b && needs_ibl && ((int)F.ibl_res[0] != b->w || (int)F.ibl_res[1] != b->h))
    return fail(DSPT_ERR_ARG, "u_ibl_resolution does not match the IBL");
  return 0;
}

int create_job(dspt_ctx* c, int scene, int ibl, const FrameParams& F, int W, int H, int num_frames, bool record, bool admit,
               int want_id, int* out_id) {
  if (W <= 0 || H <= 0 || (int64_t)W * H > (1LL << 28)) return fail(DSPT_ERR_ARG, "invalid resolution");
  if (num_frames < 0) return fail(DSPT_ERR_ARG, "num_frames < 0");
  if (int r = check_job_frame(c, F, scene, ibl)) return r;
  int64_t npix = (int64_t)W * H;
  int ring = ring_frames_for(c, npix, num_frames);
  int64_t need = job_bytes(npix, record, num_frames, ring);
  if (admit && c->budget > 0 && memory_in_use(c) + need > c->budget)
    return fail(DSPT_ERR_MEMORY, "admission: job needs " + std::to_string(need) + " B over the memory budget");
  if (admit) {  // the device's free HBM gates admission whatever the budget
    constexpr int64_t kAdmitReserve = 256ll << 20;
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipSetDevice(c->device));
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (int64_t)free_b < need + kAdmitReserve)
      return fail(DSPT_ERR_MEMORY, "admission: job needs " + std::to_string(need) + " B, the device has " +
                                       std::to_string(free_b) + " B free");
  }
  int slot_id = want_id;
  if (slot_id < 0) {
    slot_id = (int)c->jobs.size();
    for (size_t i = 0; i < c->jobs.size(); i++)
      if (!c->jobs[i]) { slot_id = (int)i; break; }
  }
  if (slot_id >= kMaxJobs) return fail(DSPT_ERR_MEMORY, "too many jobs");
  JobRes* j = new JobRes();
  j->mem.lim = &c->lim;
  int r = 0;
  if ((r = j->mem.get(&j->accum, (size_t)npix)) || (r = j->mem.get(&j->ring, (size_t)(npix * ring))) ||
      (r = j->mem.get(&j->done, (size_t)ring))) {
    j->mem.release();
    delete j;
    return r;
  }
  if (record && num_frames > 0) {
    if ((r = j->mem.get(&j->rec_L, (size_t)(npix * num_frames))) || (r = j->mem.get(&j->rec_sig, (size_t)(npix * num_frames))) ||
        (r = j->mem.get(&j->rec_prim, (size_t)(npix * num_frames)))) {
      j->mem.release();
      delete j;
      return r;
    }
  }
  k_fill_f4<<<grid_for(npix, 256), 256, 0, c->stream>>>(j->accum, npix, make_float4(0, 0, 0, 0));
  k_fill_i32<<<grid_for(ring, 256), 256, 0, c->stream>>>(j->done, ring, 0);
  k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_prog + slot_id, 1, 1);
  j->scene = scene; j->ibl = ibl; j->F = F; j->W = W; j->H = H;
  j->ring_frames = ring;
  j->frame_begin = 1; j->frame_end = num_frames + 1;
  j->next_sched = 1; j->acc_known = 1;
  j->epoch = c->chunk_no;
  j->record = record; j->rec_first = 1; j->rec_count = num_frames;
  j->alive = true;
  c->scenes[scene]->users++;
  if (ibl >= 0) c->ibls[ibl]->users++;
  *out_id = put_slot(c->jobs, j, slot_id);
  c->jobs_dirty = true;
  return 0;
}

int destroy_job(dspt_ctx* c, int id) {
  JobRes* j = c->jobs[id];
  j->cancel = true;
  if (int r = drain_job(c, id)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (j->scene >= 0 && c->scenes[j->scene]) c->scenes[j->scene]->users--;
  if (j->ibl >= 0 && j->ibl < (int)c->ibls.size() && c->ibls[j->ibl]) c->ibls[j->ibl]->users--;
  j->mem.release();
  delete j;
  c->jobs[id] = nullptr;
  c->jobs_dirty = true;
  return 0;
}

int legacy_job(dspt_ctx* c, int* out) {
  if (c->legacy_job >= 0 && job_ok(c, c->legacy_job)) { *out = c->legacy_job; return 0; }
  return fail(DSPT_ERR_STATE, "no render target (dspt_resize)");
}

std::vector<int> listed_jobs(dspt_ctx* c, const int32_t* job_ids, int32_t n) {
  std::vector<int> ids;
  if (job_ids) {
    for (int i = 0; i < n; i++)
      if (job_ok(c, job_ids[i])) ids.push_back(job_ids[i]);
  } else {
    for (size_t i = 0; i < c->jobs.size(); i++)
      if (job_ok(c, (int)i) && (int)i != c->legacy_job) ids.push_back((int)i);
  }
  return ids;
}

}  // namespace

extern "C" {

const char* dspt_last_error(void) { return g_err.c_str(); }
const char* dspt_version(void) { return "dspt 0.3.0 gfx950"; }

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
  c->state_mem.lim = &c->lim;
  c->profile = profile;
  if (const char* mp = std::getenv("DSPT_MAX_PATHS")) c->max_paths = std::max<int64_t>(65536, std::atoll(mp));
  if (const char* g = std::getenv("DSPT_BOUNCE_GRAPHS")) c->use_graphs = std::atoi(g) != 0;
  if (const char* o = std::getenv("DSPT_RAY_ORDER")) {
    const int m = std::atoi(o);
    c->ray_order = (m == DSPT_ORDER_BUCKETS || m == DSPT_ORDER_SHADE) ? m : DSPT_ORDER_NONE;
  }
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  void* p = nullptr;
  auto dev = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (ok && hipMalloc(&q, bytes) != hipSuccess) { ok = false; q = nullptr; }
    return q;
  };
  auto host = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (ok && hipHostMalloc(&q, bytes, hipHostMallocDefault) != hipSuccess) { ok = false; q = nullptr; }
    return q;
  };
  c->cnt = (unsigned long long*)dev((size_t)kCntSlots * kNumCounters * sizeof(unsigned long long));
  if (ok) ok = hipMemset(c->cnt, 0, (size_t)kCntSlots * kNumCounters * sizeof(unsigned long long)) == hipSuccess;
  c->d_jobs = (DevJob*)dev(kMaxJobs * sizeof(DevJob));
  c->d_tasks = (PassTask*)dev(kMaxTasks * sizeof(PassTask));
  c->d_ctl = (StreamCtl*)dev(sizeof(StreamCtl));
  c->d_sched = (StreamSched*)dev(sizeof(StreamSched));
  c->d_prog = (int32_t*)dev(kMaxJobs * sizeof(int32_t));
  c->d_joblist = (int32_t*)dev(kMaxJobs * sizeof(int32_t));
  c->h_jobs = (DevJob*)host(kMaxJobs * sizeof(DevJob));
  c->h_rb = (Readback*)host(2 * sizeof(Readback));
  c->h_sched = (StreamSched*)host(2 * sizeof(StreamSched));
  c->h_tasks = (PassTask*)host(2 * (size_t)kMaxTasks * sizeof(PassTask));
  c->h_joblist = (int32_t*)host(2 * (size_t)kMaxJobs * sizeof(int32_t));
  for (int i = 0; i < 2 && ok; i++) ok = hipEventCreateWithFlags(&c->ev_chunk[i], hipEventDisableTiming) == hipSuccess;
  (void)p;
  if (!ok) {
    dspt_destroy(c);
    return fail(DSPT_ERR_MEMORY, "context allocation failed");
  }
  if (int r = scratch_address(device, &c->scratch)) {  // hipSetDevice(device) is current
    dspt_destroy(c);
    return r;
  }
  *out = c;
  return 0;
}

int dspt_destroy(dspt_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drain_events(c);
  drop_graphs(c);
  for (size_t i = 0; i < c->jobs.size(); i++)
    if (c->jobs[i]) { c->jobs[i]->mem.release(); delete c->jobs[i]; }
  for (auto* s : c->scenes) if (s) { s->mem.release(); s->tex_mem.release(); delete s; }
  for (auto* b : c->ibls) if (b) { b->mem.release(); delete b; }
  c->state_mem.release();
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : c->ev_chunk) if (e) (void)hipEventDestroy(e);
  for (void* d : {(void*)c->cnt, (void*)c->d_jobs, (void*)c->d_tasks, (void*)c->d_ctl, (void*)c->d_sched, (void*)c->d_prog,
                  (void*)c->d_joblist, (void*)c->post_buf})
    if (d) (void)hipFree(d);
  for (void* h : {(void*)c->h_jobs, (void*)c->h_rb, (void*)c->h_sched, (void*)c->h_tasks, (void*)c->h_joblist})
    if (h) (void)hipHostFree(h);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

// ------------------------------------------------------------------ resources
int dspt_scene_create(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* nodes,
                      int64_t N, const int32_t* bvh_idx, int64_t M, const float* mats, int32_t K, int32_t* scene_id) {
  if (!c || !scene_id) return fail(DSPT_ERR_ARG, "NULL argument");
  SceneRes* s = nullptr;
  if (int r = make_scene(c, vertices, V, tri_idx, T, nodes, N, bvh_idx, M, mats, K, &s)) return r;
  *scene_id = put_slot(c->scenes, s);
  return 0;
}

int dspt_scene_create_instanced(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T,
                                const int64_t* mesh_first_tri, const int64_t* mesh_num_tris, int32_t num_meshes,
                                const float* instance_o2w, const int32_t* instance_mesh, int32_t num_instances,
                                const float* mats, int32_t K, int32_t* scene_id) {
  if (!c || !scene_id) return fail(DSPT_ERR_ARG, "NULL argument");
  SceneRes* s = nullptr;
  if (int r = make_scene_instanced(c, vertices, V, tri_idx, T, mesh_first_tri, mesh_num_tris, num_meshes, instance_o2w,
                                   instance_mesh, num_instances, mats, K, &s))
    return r;
  *scene_id = put_slot(c->scenes, s);
  return 0;
}

int dspt_scene_create_gpu_bvh(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T,
                              const float* mats, int32_t K, int32_t max_leaf, int32_t method, int32_t* scene_id,
                              double* build_ms) {
  if (!c || !scene_id) return fail(DSPT_ERR_ARG, "NULL argument");
  SceneRes* s = nullptr;
  if (int r = make_scene_gpu_bvh(c, vertices, V, tri_idx, T, mats, K, max_leaf, method, &s, build_ms)) return r;
  *scene_id = put_slot(c->scenes, s);
  return 0;
}

int dspt_build_bvh_gpu(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, int32_t max_leaf,
                       int32_t method, float* nodes, int64_t* num_nodes, int32_t* bvh_indices, double* build_ms) {
  if (!c || !nodes || !num_nodes || !bvh_indices) return fail(DSPT_ERR_ARG, "NULL argument");
  SceneRes* s = nullptr;
  if (int r = make_scene_gpu_bvh(c, vertices, V, tri_idx, T, nullptr, 0, max_leaf, method, &s, build_ms)) return r;
  std::vector<float4> tri((size_t)T * 3);
  hipError_t e = hipMemcpy(nodes, s->nodes, (size_t)s->N * 64, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(tri.data(), s->tri, (size_t)T * 48, hipMemcpyDeviceToHost);
  *num_nodes = s->N;
  s->mem.release();
  delete s;
  if (e != hipSuccess) return fail(DSPT_ERR_DEVICE, std::string("bvh readback: ") + hipGetErrorString(e));
  for (int64_t i = 0; i < T; i++) std::memcpy(&bvh_indices[i], &tri[(size_t)i * 3].w, 4);  // id bits in .w
  return 0;
}

int dspt_scene_destroy(dspt_ctx* c, int32_t id) {
  if (!c || id < 0 || id >= (int)c->scenes.size() || !c->scenes[id]) return fail(DSPT_ERR_ARG, "unknown scene id");
  if (c->scenes[id]->users > 0) return fail(DSPT_ERR_STATE, "scene is used by live jobs");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->scenes[id]->mem.release();
  c->scenes[id]->tex_mem.release();
  delete c->scenes[id];
  c->scenes[id] = nullptr;
  return 0;
}

int dspt_ibl_create(dspt_ctx* c, const float* rgba, int32_t w, int32_t h, double* total_sum, int32_t* ibl_id) {
  if (!c || !ibl_id) return fail(DSPT_ERR_ARG, "NULL argument");
  IblRes* b = nullptr;
  if (int r = make_ibl(c, rgba, w, h, &b)) return r;
  if (total_sum) *total_sum = b->total;
  *ibl_id = put_slot(c->ibls, b);
  return 0;
}

int dspt_ibl_destroy(dspt_ctx* c, int32_t id) {
  if (!c || id < 0 || id >= (int)c->ibls.size() || !c->ibls[id]) return fail(DSPT_ERR_ARG, "unknown IBL id");
  if (c->ibls[id]->users > 0) return fail(DSPT_ERR_STATE, "IBL is used by live jobs");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ibls[id]->mem.release();
  delete c->ibls[id];
  c->ibls[id] = nullptr;
  return 0;
}

// Runs chunks without scheduling until nothing is in flight.
static int drain_all(dspt_ctx* c);

int dspt_scene_set_textures(dspt_ctx* c, int32_t scene_id, const float* texinfos, int32_t num_texinfos,
                            const int32_t* array_dims, int32_t num_arrays, const uint8_t* texels, int64_t num_texel_bytes) {
  if (!c || scene_id < 0 || scene_id >= (int)c->scenes.size() || !c->scenes[scene_id])
    return fail(DSPT_ERR_ARG, "unknown scene id");
  if (num_texinfos < 0 || num_arrays < 0 || num_texel_bytes < 0) return fail(DSPT_ERR_ARG, "negative texture count");
  if ((num_texinfos > 0 && !texinfos) || (num_arrays > 0 && (!array_dims || !texels)))
    return fail(DSPT_ERR_ARG, "NULL texture buffer");
  int64_t want = 0;
  std::vector<TexArrayDesc> desc((size_t)std::max(num_arrays, 1));
  for (int32_t a = 0; a < num_arrays; a++) {
    const int32_t w = array_dims[3 * a], h = array_dims[3 * a + 1], l = array_dims[3 * a + 2];
    if (w <= 0 || h <= 0 || l <= 0 || (int64_t)w * h * l > (1LL << 31))
      return fail(DSPT_ERR_ARG, "texture array " + std::to_string(a) + ": invalid dimensions");
    desc[a].offset = want / 4;
    desc[a].w = w; desc[a].h = h; desc[a].layers = l; desc[a].pad = 0;
    want += (int64_t)w * h * l * 4;
  }
  if (want != num_texel_bytes) return fail(DSPT_ERR_ARG, "texel byte count does not match the array dimensions");
  for (int32_t i = 0; i < num_texinfos; i++)
    for (int k = 0; k < 8; k++)
      if (!std::isfinite(texinfos[(size_t)i * 8 + k]))
        return fail(DSPT_ERR_ARG, "texture info " + std::to_string(i) + " holds a non-finite value");
  HIPCHK(hipSetDevice(c->device));
  if (int r = drain_all(c)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  SceneRes* s = c->scenes[scene_id];
  s->tex_mem.release();
  s->texinfo = nullptr; s->texarr = nullptr; s->texels = nullptr;
  s->num_tex = 0; s->num_texarr = 0;
  if (num_texinfos > 0 || num_arrays > 0) {
    Allocs& A = s->tex_mem;
    int r = 0;
    if ((r = A.get(&s->texinfo, (size_t)num_texinfos * 2)) || (r = A.get(&s->texarr, (size_t)std::max(num_arrays, 1))) ||
        (r = A.get(&s->texels, (size_t)std::max<int64_t>(want / 4, 1)))) {
      A.release();
      return r;
    }
    if (num_texinfos > 0)
      HIPCHK(hipMemcpyAsync(s->texinfo, texinfos, (size_t)num_texinfos * 32, hipMemcpyHostToDevice, c->stream));
    if (num_arrays > 0) {
      HIPCHK(hipMemcpyAsync(s->texarr, desc.data(), (size_t)num_arrays * sizeof(TexArrayDesc), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(s->texels, texels, (size_t)want, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    s->num_tex = num_texinfos;
    s->num_texarr = num_arrays;
  }
  c->jobs_dirty = true;
  return 0;
}

static int drain_all(dspt_ctx* c) {
  for (size_t i = 0; i < c->jobs.size(); i++)
    if (job_ok(c, (int)i) && c->jobs[i]->in_flight())
      if (int r = drain_job(c, (int)i)) return r;
  return 0;
}

// Single-image API: scene 0 / IBL 0 are replaced in place (after the work in
// flight has drained); the implicit job is re-pointed at them.
int dspt_upload_scene(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T, const float* nodes,
                      int64_t N, const int32_t* bvh_idx, int64_t M, const float* mats, int32_t K) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  SceneRes* s = nullptr;
  if (int r = make_scene(c, vertices, V, tri_idx, T, nodes, N, bvh_idx, M, mats, K, &s)) return r;
  if (int r = drain_all(c)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  int users = 0;
  if (!c->scenes.empty() && c->scenes[0]) {
    users = c->scenes[0]->users;
    c->scenes[0]->mem.release();
    c->scenes[0]->tex_mem.release();
    delete c->scenes[0];
  }
  s->users = users;
  put_slot(c->scenes, s, 0);
  c->jobs_dirty = true;
  return 0;
}

int dspt_upload_scene_instanced(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T,
                                const int64_t* mesh_first_tri, const int64_t* mesh_num_tris, int32_t num_meshes,
                                const float* instance_o2w, const int32_t* instance_mesh, int32_t num_instances,
                                const float* mats, int32_t K) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  SceneRes* s = nullptr;
  if (int r = make_scene_instanced(c, vertices, V, tri_idx, T, mesh_first_tri, mesh_num_tris, num_meshes, instance_o2w,
                                   instance_mesh, num_instances, mats, K, &s))
    return r;
  if (int r = drain_all(c)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  int users = 0;
  if (!c->scenes.empty() && c->scenes[0]) {
    users = c->scenes[0]->users;
    c->scenes[0]->mem.release();
    c->scenes[0]->tex_mem.release();
    delete c->scenes[0];
  }
  s->users = users;
  put_slot(c->scenes, s, 0);
  c->jobs_dirty = true;
  return 0;
}

int dspt_upload_scene_gpu_bvh(dspt_ctx* c, const float* vertices, int64_t V, const uint32_t* tri_idx, int64_t T,
                              const float* mats, int32_t K, int32_t max_leaf, int32_t method, double* build_ms) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  SceneRes* s = nullptr;
  if (int r = make_scene_gpu_bvh(c, vertices, V, tri_idx, T, mats, K, max_leaf, method, &s, build_ms)) return r;
  if (int r = drain_all(c)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  int users = 0;
  if (!c->scenes.empty() && c->scenes[0]) {
    users = c->scenes[0]->users;
    c->scenes[0]->mem.release();
    c->scenes[0]->tex_mem.release();
    delete c->scenes[0];
  }
  s->users = users;
  put_slot(c->scenes, s, 0);
  c->jobs_dirty = true;
  return 0;
}

int dspt_upload_ibl(dspt_ctx* c, const float* rgba, int32_t w, int32_t h, double* total_sum) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  IblRes* b = nullptr;
  if (int r = make_ibl(c, rgba, w, h, &b)) return r;
  if (int r = drain_all(c)) return r;
  HIPCHK(hipStreamSynchronize(c->stream));
  int users = 0;
  if (!c->ibls.empty() && c->ibls[0]) {
    users = c->ibls[0]->users;
    c->ibls[0]->mem.release();
    delete c->ibls[0];
  }
  b->users = users;
  put_slot(c->ibls, b, 0);
  if (total_sum) *total_sum = b->total;
  c->jobs_dirty = true;
  return 0;
}

int dspt_read_ibl_tables(dspt_ctx* c, float* pdf, float* cdf, float* ypdf, float* ycdf) {
  if (!c || c->ibls.empty() || !c->ibls[0]) return fail(DSPT_ERR_STATE, "no IBL uploaded");
  IblRes* b = c->ibls[0];
  size_t n = (size_t)b->w * b->h;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (pdf) HIPCHK(hipMemcpy(pdf, b->pdf, n * 4, hipMemcpyDeviceToHost));
  if (cdf) HIPCHK(hipMemcpy(cdf, b->cdf, n * 4, hipMemcpyDeviceToHost));
  if (ypdf) HIPCHK(hipMemcpy(ypdf, b->ypdf, (size_t)b->h * 4, hipMemcpyDeviceToHost));
  if (ycdf) HIPCHK(hipMemcpy(ycdf, b->ycdf, (size_t)b->h * 4, hipMemcpyDeviceToHost));
  return 0;
}

int dspt_resize(dspt_ctx* c, int32_t w, int32_t h) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (w <= 0 || h <= 0 || (int64_t)w * h > (1LL << 28)) return fail(DSPT_ERR_ARG, "invalid resolution");
  HIPCHK(hipSetDevice(c->device));
  c->W = w;
  c->H = h;
  if (c->legacy_job >= 0 && job_ok(c, c->legacy_job))
    if (int r = destroy_job(c, c->legacy_job)) return r;
  c->legacy_job = -1;
  return 0;
}

int dspt_set_frame(dspt_ctx* c, const float* block) {
  if (!c || !block) return fail(DSPT_ERR_ARG, "NULL argument");
  static_assert(sizeof(FrameParams) == DSPT_FRAME_FLOATS * 4, "frame block layout");
  FrameParams F;
  std::memcpy(&F, block, sizeof(FrameParams));
  if (!(F.max_bounces >= 0.0f) || F.max_bounces > 1024.0f) return fail(DSPT_ERR_ARG, "max_bounces out of range");
  c->F = F;
  c->have_frame = true;
  return 0;
}

int dspt_set_record_samples(dspt_ctx* c, int enable) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  c->record = enable != 0;
  return 0;
}

namespace {

// Frames first .. first+num_frames-1 of the implicit job over the region
// [x0, x0+w) x [y0, y0+h) (w == 0: the whole image, which restarts the
// accumulation at frame 1).
int render_impl(dspt_ctx* c, int32_t first_frame, int32_t num_frames, int32_t x0, int32_t y0, int32_t w, int32_t h) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (c->scenes.empty() || !c->scenes[0]) return fail(DSPT_ERR_STATE, "render before dspt_upload_scene");
  if (!c->have_frame) return fail(DSPT_ERR_STATE, "render before dspt_set_frame");
  if (c->W <= 0) return fail(DSPT_ERR_STATE, "render before dspt_resize");
  if (first_frame < 1 || num_frames < 0) return fail(DSPT_ERR_ARG, "frames are 1-based and num_frames >= 0");
  if ((int64_t)first_frame + num_frames >= (1LL << 30)) return fail(DSPT_ERR_ARG, "frame index out of range");
  if (w != 0 && (x0 < 0 || y0 < 0 || w < 1 || h < 1 || (int64_t)x0 + w > c->W || (int64_t)y0 + h > c->H))
    return fail(DSPT_ERR_ARG, "region outside the image");
  HIPCHK(hipSetDevice(c->device));
  int ibl = (!c->ibls.empty() && c->ibls[0]) ? 0 : -1;
  if (int r = check_job_frame(c, c->F, 0, ibl)) return r;
  // the implicit job is rebuilt when the recording request changes, or when
  // a call that restarts the accumulation (first_frame 1) could keep more
  // frames in flight than its ring holds (the ring was sized by an earlier,
  // shorter call); a continuing call keeps the job and its running mean
  if (c->legacy_job >= 0 && job_ok(c, c->legacy_job)) {
    JobRes* j = c->jobs[c->legacy_job];
    bool rec_fit = !c->record || (j->record && j->rec_count >= num_frames);
    bool ring_fit = first_frame > 1 || j->ring_frames >= ring_frames_for(c, j->npix(), num_frames);
    if (!rec_fit || !ring_fit || j->record != c->record || j->ibl != ibl) {
      if (int r = destroy_job(c, c->legacy_job)) return r;
      c->legacy_job = -1;
    }
  }
  if (c->legacy_job < 0 || !job_ok(c, c->legacy_job)) {
    int id = -1;
    if (int r = create_job(c, 0, ibl, c->F, c->W, c->H, num_frames, c->record, false, -1, &id)) return r;
    c->legacy_job = id;
  }
  if (num_frames == 0) return 0;
  const int id = c->legacy_job;
  JobRes* j = c->jobs[id];
  j->F = c->F;
  j->ibl = ibl;
  const bool whole = w == 0 || (x0 == 0 && y0 == 0 && w == c->W && h == c->H);
  j->rx0 = whole ? 0 : x0;
  j->ry0 = whole ? 0 : y0;
  j->rw = whole ? 0 : w;
  j->rh = whole ? 0 : h;
  j->rec_first = first_frame;
  j->frame_begin = first_frame;
  j->frame_end = first_frame + num_frames;
  j->next_sched = first_frame;
  j->acc_known = first_frame;
  j->cancel = false;
  j->epoch = c->chunk_no;
  c->jobs_dirty = true;
  k_fill_i32<<<1, 64, 0, c->stream>>>(c->d_prog + id, 1, first_frame);
  if (first_frame == 1 && w == 0)
    k_fill_f4<<<grid_for(j->npix(), 256), 256, 0, c->stream>>>(j->accum, j->npix(), make_float4(0, 0, 0, 0));
  HIPCHK(hipGetLastError());
  if (int r = run_until_complete(c, {id})) return r;
  c->last_frames = num_frames;
  return 0;
}

}  // namespace

int dspt_render(dspt_ctx* c, int32_t first_frame, int32_t num_frames) {
  return render_impl(c, first_frame, num_frames, 0, 0, 0, 0);
}

int dspt_render_region(dspt_ctx* c, int32_t first_frame, int32_t num_frames, int32_t x0, int32_t y0, int32_t width,
                       int32_t height) {
  if (width <= 0 || height <= 0) return fail(DSPT_ERR_ARG, "region width and height must be positive");
  return render_impl(c, first_frame, num_frames, x0, y0, width, height);
}

int dspt_clear_accum(dspt_ctx* c) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (c->legacy_job < 0 || !job_ok(c, c->legacy_job)) return 0;  // nothing accumulated yet
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (int r = settle(c)) return r;
  JobRes* j = c->jobs[c->legacy_job];
  k_fill_f4<<<grid_for(j->npix(), 256), 256, 0, c->stream>>>(j->accum, j->npix(), make_float4(0, 0, 0, 0));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_synchronize(dspt_ctx* c) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  HIPCHK(hipStreamSynchronize(c->stream));
  return settle(c);
}

int dspt_read_accum(dspt_ctx* c, float* rgba) {
  if (!c || !rgba) return fail(DSPT_ERR_ARG, "NULL argument");
  int id;
  if (int r = legacy_job(c, &id)) return r;
  return dspt_job_read_accum(c, id, rgba);
}

int dspt_read_samples(dspt_ctx* c, float* rgba, uint32_t* sigs, int32_t* prim) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  int id;
  if (int r = legacy_job(c, &id)) return r;
  JobRes* j = c->jobs[id];
  if (!j->record) return fail(DSPT_ERR_STATE, "sample recording was not enabled for the last render");
  size_t n = (size_t)c->last_frames * j->npix();
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rgba) HIPCHK(hipMemcpy(rgba, j->rec_L, n * 16, hipMemcpyDeviceToHost));
  if (sigs) HIPCHK(hipMemcpy(sigs, j->rec_sig, n * 4, hipMemcpyDeviceToHost));
  if (prim) HIPCHK(hipMemcpy(prim, j->rec_prim, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------------ jobs
int dspt_job_create(dspt_ctx* c, int32_t scene_id, int32_t ibl_id, const float* block, int32_t W, int32_t H,
                    int32_t num_frames, int32_t record_samples, int32_t* job_id) {
  if (!c || !block || !job_id) return fail(DSPT_ERR_ARG, "NULL argument");
  FrameParams F;
  std::memcpy(&F, block, sizeof(FrameParams));
  HIPCHK(hipSetDevice(c->device));
  int id = -1;
  if (int r = create_job(c, scene_id, ibl_id, F, W, H, num_frames, record_samples != 0, true, -1, &id)) return r;
  *job_id = id;
  return 0;
}

int dspt_job_destroy(dspt_ctx* c, int32_t id) {
  if (!c || !job_ok(c, id)) return fail(DSPT_ERR_ARG, "unknown job id");
  if (id == c->legacy_job) c->legacy_job = -1;
  return destroy_job(c, id);
}

int dspt_job_status(dspt_ctx* c, int32_t id, int32_t* frames_done, int32_t* frames_total) {
  if (!c || !job_ok(c, id)) return fail(DSPT_ERR_ARG, "unknown job id");
  JobRes* j = c->jobs[id];
  if (frames_done) *frames_done = j->acc_known - j->frame_begin;
  if (frames_total) *frames_total = j->frame_end - j->frame_begin;
  return 0;
}

int dspt_job_read_accum(dspt_ctx* c, int32_t id, float* rgba) {
  if (!c || !rgba || !job_ok(c, id)) return fail(DSPT_ERR_ARG, "unknown job id or NULL output");
  JobRes* j = c->jobs[id];
  HIPCHK(hipMemcpyAsync(rgba, j->accum, (size_t)j->npix() * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_job_postprocess(dspt_ctx* c, int32_t id, int32_t mode, float exposure, int32_t gamma, int32_t fxaa, int32_t ow,
                         int32_t oh, uint8_t* rgba8, float* rgba32) {
  if (!c || !job_ok(c, id)) return fail(DSPT_ERR_ARG, "unknown job id");
  if (ow <= 0 || oh <= 0 || (int64_t)ow * oh > (1LL << 28)) return fail(DSPT_ERR_ARG, "invalid canvas size");
  if (!(exposure == exposure)) return fail(DSPT_ERR_ARG, "exposure is NaN");
  if (!rgba8 && !rgba32) return 0;
  JobRes* j = c->jobs[id];
  HIPCHK(hipSetDevice(c->device));
  const int64_t n = (int64_t)ow * oh;
  const size_t need = (size_t)n * ((rgba8 ? 4 : 0) + (rgba32 ? 16 : 0));
  if (need > c->post_cap) {
    if (c->post_buf) HIPCHK(hipFree(c->post_buf));
    c->post_buf = nullptr;
    c->post_cap = 0;
    if (hipMalloc(&c->post_buf, need) != hipSuccess) return fail(DSPT_ERR_MEMORY, "display buffer");
    c->post_cap = need;
  }
  float4* d32 = rgba32 ? reinterpret_cast<float4*>(c->post_buf) : nullptr;
  uint32_t* d8 = rgba8 ? reinterpret_cast<uint32_t*>(c->post_buf + (rgba32 ? (size_t)n * 16 : 0)) : nullptr;
  post::Source S{j->accum, j->W, j->H, post::Params{mode, gamma ? 1 : 0, exposure}};
  // fxaa.frag's u_inv_res: 1 / renderRes in JS doubles, rounded by uniform2f
  const float ix = (float)(1.0 / (double)j->W), iy = (float)(1.0 / (double)j->H);
  timed(c, DSPT_K_POST, [&] {
    k_postprocess<<<grid_for(n, 256), 256, 0, c->stream>>>(S, fxaa ? 1 : 0, ix, iy, ow, oh, d8, d32);
  });
  HIPCHK(hipGetLastError());
  if (rgba32) HIPCHK(hipMemcpyAsync(rgba32, d32, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
  if (rgba8) HIPCHK(hipMemcpyAsync(rgba8, d8, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  drain_events(c);
  return 0;
}

int dspt_postprocess(dspt_ctx* c, int32_t mode, float exposure, int32_t gamma, int32_t fxaa, int32_t ow, int32_t oh,
                     uint8_t* rgba8, float* rgba32) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  int id;
  if (int r = legacy_job(c, &id)) return r;
  return dspt_job_postprocess(c, id, mode, exposure, gamma, fxaa, ow, oh, rgba8, rgba32);
}

int dspt_job_read_samples(dspt_ctx* c, int32_t id, float* rgba, uint32_t* sigs, int32_t* prim) {
  if (!c || !job_ok(c, id)) return fail(DSPT_ERR_ARG, "unknown job id");
  JobRes* j = c->jobs[id];
  if (!j->record) return fail(DSPT_ERR_STATE, "job was created without sample recording");
  size_t n = (size_t)j->rec_count * j->npix();
  HIPCHK(hipStreamSynchronize(c->stream));
  if (rgba) HIPCHK(hipMemcpy(rgba, j->rec_L, n * 16, hipMemcpyDeviceToHost));
  if (sigs) HIPCHK(hipMemcpy(sigs, j->rec_sig, n * 4, hipMemcpyDeviceToHost));
  if (prim) HIPCHK(hipMemcpy(prim, j->rec_prim, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int dspt_jobs_step(dspt_ctx* c, const int32_t* job_ids, int32_t n, int32_t* active) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  HIPCHK(hipSetDevice(c->device));
  std::vector<int> ids = listed_jobs(c, job_ids, n);
  if (int r = run_chunk(c, ids, true)) return r;
  if (active) {
    int k = 0;
    for (int id : ids)
      if (!c->jobs[id]->complete()) k++;
    *active = k;
  }
  return 0;
}

int dspt_jobs_run(dspt_ctx* c, const int32_t* job_ids, int32_t n) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  HIPCHK(hipSetDevice(c->device));
  return run_until_complete(c, listed_jobs(c, job_ids, n));
}

int dspt_set_memory_limit(dspt_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) return fail(DSPT_ERR_ARG, "invalid limit");
  c->lim.cap = bytes;
  return 0;
}

int dspt_set_memory_budget(dspt_ctx* c, int64_t bytes) {
  if (!c || bytes < 0) return fail(DSPT_ERR_ARG, "invalid budget");
  c->budget = bytes;
  return 0;
}

int dspt_set_launch_config(dspt_ctx* c, int64_t max_paths, int32_t trace_blocks, int32_t shade_blocks) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (max_paths > (1LL << 30)) return fail(DSPT_ERR_ARG, "max_paths above 2^30");
  const bool repool = (max_paths > 0 && std::max<int64_t>(65536, max_paths) != c->max_paths) ||
                      (trace_blocks > 0 && trace_blocks != c->trace_blocks_max);
  if (repool) {
    if (int r = drain_all(c)) return r;
    if (int r = settle(c)) return r;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (max_paths > 0) c->max_paths = std::max<int64_t>(65536, max_paths);
    if (trace_blocks > 0) c->trace_blocks_max = trace_blocks;
    c->cap = 0;  // reallocated (with the spill area) on the next chunk
    c->have_prev = false;
  }
  if (shade_blocks > 0) c->shade_blocks_max = shade_blocks;
  return 0;
}

int dspt_get_launch_config(dspt_ctx* c, int64_t* max_paths, int32_t* trace_blocks, int32_t* shade_blocks) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (max_paths) *max_paths = c->max_paths;
  if (trace_blocks) *trace_blocks = c->trace_blocks_max;
  if (shade_blocks) *shade_blocks = c->shade_blocks_max;
  return 0;
}

int dspt_set_schedule(dspt_ctx* c, int32_t mode, int32_t pass_frames, int32_t chunk_iters, float ring_factor) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (mode != DSPT_SCHED_STREAM && mode != DSPT_SCHED_PASS) return fail(DSPT_ERR_ARG, "unknown schedule mode");
  if (chunk_iters < 0 || chunk_iters > 64) return fail(DSPT_ERR_ARG, "chunk_iters must be in 1..64 (0 keeps)");
  if (!(ring_factor >= 0.0f) || ring_factor > 1024.0f) return fail(DSPT_ERR_ARG, "ring_factor must be in [1, 1024] (0 keeps)");
  c->mode = mode;
  c->pass_frames = std::max(0, pass_frames);
  if (chunk_iters > 0) c->chunk_iters = chunk_iters;
  if (ring_factor > 0.0f) c->ring_factor = std::max(1.0f, ring_factor);
  return 0;
}

int dspt_set_policy(dspt_ctx* c, int32_t policy) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (policy != DSPT_POLICY_RR && policy != DSPT_POLICY_FIFO && policy != DSPT_POLICY_SRPT)
    return fail(DSPT_ERR_ARG, "unknown scheduling policy");
  c->policy = policy;
  return 0;
}

int dspt_get_schedule(dspt_ctx* c, int32_t* mode, int32_t* pass_frames, int32_t* chunk_iters, float* ring_factor) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (mode) *mode = c->mode;
  if (pass_frames) *pass_frames = c->pass_frames;
  if (chunk_iters) *chunk_iters = c->chunk_iters;
  if (ring_factor) *ring_factor = c->ring_factor;
  return 0;
}

// ------------------------------------------------------------------ counters
int dspt_get_counters(dspt_ctx* c, int64_t* out) {
  if (!c || !out) return fail(DSPT_ERR_ARG, "NULL argument");
  unsigned long long h[kNumCounters];
  HIPCHK(hipStreamSynchronize(c->stream));
  if (int r = read_counters(c, h)) return r;
  for (int i = 0; i < DSPT_CNT_COUNT; i++) out[i] = (int64_t)h[i];
  out[DSPT_CNT_NODE_VISITS] = (int64_t)(h[4] + h[10]);
  out[DSPT_CNT_TRI_TESTS] = (int64_t)(h[5] + h[11]);
  out[DSPT_CNT_RAYS_PRIMARY] = (int64_t)h[12];  // camera rays issued by the refill
  return 0;
}

int dspt_reset_counters(dspt_ctx* c) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  drain_events(c);
  std::memset(c->kstat, 0, sizeof(c->kstat));
  HIPCHK(hipMemsetAsync(c->cnt, 0, (size_t)kCntSlots * kNumCounters * sizeof(unsigned long long), c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_set_ray_order(dspt_ctx* c, int mode) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  if (mode != DSPT_ORDER_NONE && mode != DSPT_ORDER_BUCKETS && mode != DSPT_ORDER_SHADE)
    return fail(DSPT_ERR_ARG, "unknown ray order mode");
  c->ray_order = mode;  // takes effect with the next chunk (the graph key includes it)
  return 0;
}

int dspt_set_bounce_graphs(dspt_ctx* c, int enable) {
  if (!c) return fail(DSPT_ERR_ARG, "ctx is NULL");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->use_graphs = enable != 0;
  drop_graphs(c);
  c->graph_key = {};
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
  if (int r = read_counters(c, h)) return r;
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

// ------------------------------------------------------------------ traversal probes
static int trace_common(dspt_ctx* c, const float* rays, int64_t n, int any, int32_t* tri, float* t, float* u, float* v,
                        int32_t* inst) {
  if (!c || !rays || !tri) return fail(DSPT_ERR_ARG, "NULL argument");
  if (c->scenes.empty() || !c->scenes[0]) return fail(DSPT_ERR_STATE, "trace before dspt_upload_scene");
  if (n <= 0) return 0;
  HIPCHK(hipSetDevice(c->device));
  JobRes tmp;
  tmp.scene = 0;
  SceneView S = view_of(c, &tmp);
  ScopedAllocs A;  // freed on every return
  A.lim = &c->lim;
  float4* d_rays = nullptr;
  int32_t *d_tri = nullptr, *d_inst = nullptr;
  float *d_t = nullptr, *d_u = nullptr, *d_v = nullptr;
  int* d_spill = nullptr;
  int blocks = (int)std::min<int64_t>(2048, grid_for(n, kTraceBlock));
  int r = 0;
  if ((r = A.get(&d_rays, (size_t)n * 2)) || (r = A.get(&d_tri, (size_t)n))) return r;
  if (inst && (r = A.get(&d_inst, (size_t)n))) return r;
  if (t && ((r = A.get(&d_t, (size_t)n)) || (r = A.get(&d_u, (size_t)n)) || (r = A.get(&d_v, (size_t)n)))) return r;
  if ((r = A.get(&d_spill, (size_t)blocks * kTraceBlock * (kStackSize - kLdsStack)))) return r;
  HIPCHK(hipMemcpyAsync(d_rays, rays, n * 32, hipMemcpyHostToDevice, c->stream));
  k_trace_rays<<<blocks, kTraceBlock, 0, c->stream>>>(S, d_rays, n, any, d_tri, d_t, d_u, d_v, d_inst, d_spill, c->cnt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(tri, d_tri, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (inst) HIPCHK(hipMemcpyAsync(inst, d_inst, n * 4, hipMemcpyDeviceToHost, c->stream));
  if (t) {
    HIPCHK(hipMemcpyAsync(t, d_t, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(u, d_u, n * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(v, d_v, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int dspt_trace_closest(dspt_ctx* c, const float* rays, int64_t n, int32_t* tri, float* t, float* u, float* v) {
  if (!t || !u || !v) return fail(DSPT_ERR_ARG, "NULL output");
  return trace_common(c, rays, n, 0, tri, t, u, v, nullptr);
}

int dspt_trace_closest_instanced(dspt_ctx* c, const float* rays, int64_t n, int32_t* inst, int32_t* tri, float* t, float* u,
                                 float* v) {
  if (!inst || !t || !u || !v) return fail(DSPT_ERR_ARG, "NULL output");
  return trace_common(c, rays, n, 0, tri, t, u, v, inst);
}

int dspt_trace_closest_noalloc(dspt_ctx* c, const float* d_rays, int64_t n, int32_t* d_tri, float* d_t, float* d_u,
                               float* d_v, int32_t blocks) {
  if (!c || !d_rays || !d_tri || !d_t || !d_u || !d_v) return fail(DSPT_ERR_ARG, "NULL argument");
  if (c->scenes.empty() || !c->scenes[0]) return fail(DSPT_ERR_STATE, "trace before dspt_upload_scene");
  if (n < 0) return fail(DSPT_ERR_ARG, "negative ray count");
  if (blocks <= 0) blocks = kNoAllocBlocks;
  if (blocks > kNoAllocBlocks)
    return fail(DSPT_ERR_ARG, "blocks " + std::to_string(blocks) + " exceeds the no-alloc scratch (" +
                                  std::to_string(kNoAllocBlocks) + " blocks of " + std::to_string(kTraceBlock) +
                                  " lanes, " + std::to_string(kNoAllocScratchInts * 4 >> 20) + " MiB)");
  if (!c->scratch) return fail(DSPT_ERR_STATE, "no-alloc scratch address unknown");
  if (n == 0) return 0;
  HIPCHK(hipSetDevice(c->device));
  JobRes tmp;
  tmp.scene = 0;
  SceneView S = view_of(c, &tmp);
  blocks = (int)std::min<int64_t>(blocks, grid_for(n, kTraceBlock));
  // the spill is written before it is read (stack discipline): nothing to clear
  k_trace_rays<<<blocks, kTraceBlock, 0, c->stream>>>(S, (const float4*)d_rays, n, 0, d_tri, d_t, d_u, d_v, nullptr,
                                                      c->scratch, c->cnt);
  HIPCHK(hipGetLastError());
  return 0;
}

int dspt_trace_occluded(dspt_ctx* c, const float* rays, int64_t n, int32_t* occluded) {
  if (!occluded) return fail(DSPT_ERR_ARG, "NULL output");
  std::vector<int32_t> tri(std::max<int64_t>(n, 1));
  int r = trace_common(c, rays, n, 1, tri.data(), nullptr, nullptr, nullptr, nullptr);
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

int dspt_selftest_pin_math(int device, int32_t fn, const float* x, const float* y, float* out, int64_t n) {
  if (!x || !out || (fn == 4 || fn == 5) && !y) return fail(DSPT_ERR_ARG, "NULL argument");
  if (fn < 0 || fn > 8) return fail(DSPT_ERR_ARG, "fn must be 0..8");
  if (n <= 0) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(DSPT_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(DSPT_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  ScopedAllocs A;
  float *dx = nullptr, *dy = nullptr, *dout = nullptr;
  int r;
  if ((r = A.get(&dx, (size_t)n)) || (r = A.get(&dout, (size_t)n))) return r;
  if (y && (r = A.get(&dy, (size_t)n))) return r;
  HIPCHK(hipMemcpy(dx, x, n * 4, hipMemcpyHostToDevice));
  if (y) HIPCHK(hipMemcpy(dy, y, n * 4, hipMemcpyHostToDevice));
  k_pin_math<<<grid_for(n, 256), 256>>>(fn, dx, dy, dout, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
  return 0;
}

int dspt_selftest_copy(int device, int64_t bytes, int32_t reps) {
  if (bytes <= 0 || bytes % 16 || reps <= 0) return fail(DSPT_ERR_ARG, "bytes must be a positive multiple of 16, reps > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(DSPT_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(DSPT_ERR_ARG, "device index out of range");
  HIPCHK(hipSetDevice(device));
  ScopedAllocs A;
  float4 *src = nullptr, *dst = nullptr;
  const int64_t n = bytes / 16;
  int r;
  if ((r = A.get(&src, (size_t)n)) || (r = A.get(&dst, (size_t)n))) return r;
  HIPCHK(hipMemset(src, 0, (size_t)bytes));
  for (int i = 0; i < reps; i++) k_copy_calib<<<4096, 256>>>(src, dst, n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return 0;
}

int dspt_memory_usage(dspt_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return fail(DSPT_ERR_ARG, "NULL argument");
  *bytes = memory_in_use(c);
  return 0;
}

}  // extern "C"
