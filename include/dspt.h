/*
 * dspt.h — C ABI of the MI355X path-tracing engine (gfx950, HIP).
 *
 * The engine replaces the GPU render loop behind the reference's
 * PathtracingRenderer (packages/lib/renderer.ts:116-1109, "RT" below) and
 * keeps its data formats: the 20-float vertex stream, u32 triangle indices,
 * the TinyBVH "bvh_gpu" node array (16 floats per node) with its i32 index
 * array, 68-float materials, an RGBA32F equirect IBL and the 52-float
 * uniform block of packages/lib/shader/renderer.frag:10-31.
 *
 * Conventions: every call returns 0 on success and a negative code on
 * failure; dspt_last_error() then gives a thread-local message (mirrors the
 * int-status + last-error channel of packages/lib/bvh/tinybvh_builder.ts:86-97).
 * Buffers are caller-owned host memory unless stated; the engine copies what
 * it keeps into device memory (HBM) it owns. Images are RGBA f32, row 0 =
 * bottom row (GL convention), pixel (x, y) at index y*W + x.
 */
#ifndef DSPT_H
#define DSPT_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dspt_ctx dspt_ctx;

enum {
  DSPT_OK = 0,
  DSPT_ERR_ARG = -1,       /* invalid argument / malformed input */
  DSPT_ERR_DEVICE = -2,    /* HIP runtime error or no GPU */
  DSPT_ERR_STATE = -3,     /* call order (e.g. render before upload) */
  DSPT_ERR_MEMORY = -4,    /* device allocation or admission budget exceeded */
  DSPT_ERR_UNSUPPORTED = -5
};

enum { DSPT_PROFILE_LEAN = 0, DSPT_PROFILE_FULL = 1 };

/* Number of floats in the frame parameter block (std140 uniform block of
 * renderer.frag:10-31, packed in the key order of RT:376-438). */
#define DSPT_FRAME_FLOATS 52

/* Counter slots returned by dspt_get_counters(). */
enum {
  DSPT_CNT_PATHS = 0,        /* pixel samples completed */
  DSPT_CNT_RAYS_PRIMARY = 1,
  DSPT_CNT_RAYS_EXTENSION = 2,
  DSPT_CNT_RAYS_SHADOW = 3,
  DSPT_CNT_NODE_VISITS = 4,  /* closest-hit + any-hit traversals */
  DSPT_CNT_TRI_TESTS = 5,
  DSPT_CNT_SHADE_HITS = 6,
  DSPT_CNT_ENV_MISSES = 7,
  DSPT_CNT_NEE_SAMPLES = 8,
  DSPT_CNT_COUNT = 16
};

const char* dspt_last_error(void);
const char* dspt_version(void);

/* Replaces `new PathtracingRenderer({materialProfile})` (RT:309-325). */
int dspt_create(int profile, int device, dspt_ctx** out);
int dspt_destroy(dspt_ctx* ctx);

/* Replaces RT.setScene (RT:914-958) after the BVH build of RT.initBvh
 * (RT:872-912): vertex stream f32[20*V] (packages/lib/scene_data.ts:21),
 * triangle indices u32[3*T], BVH nodes f32[16*N] and BVH indices i32[M]
 * (packages/lib/bvh/tinybvh_builder.ts:1-5), materials f32[68*K]
 * (packages/lib/material.ts:2-481). Materials that reference textures
 * need the scene's textures (dspt_scene_set_textures) before rendering;
 * replacing the scene drops them. */
int dspt_upload_scene(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                      int64_t num_triangles, const float* bvh_nodes, int64_t num_nodes, const int32_t* bvh_indices,
                      int64_t num_bvh_indices, const float* materials, int32_t num_materials);

/* Replaces RT.setIBL (RT:840-865): uploads the RGBA32F equirect map and
 * builds the importance-sampling pdf/cdf tables on the device, bit-identical
 * to precomputeIBLImportanceSamplingData (RT:777-838). If total_sum is not
 * NULL it receives the marginal sum (u_ibl_pdf_total_sum). */
int dspt_upload_ibl(dspt_ctx* ctx, const float* rgba, int32_t width, int32_t height, double* total_sum);
/* Reads the device CDF tables back (pdf, cdf f32[w*h]; ypdf, ycdf f32[h]). */
int dspt_read_ibl_tables(dspt_ctx* ctx, float* pdf, float* cdf, float* ypdf, float* ycdf);

/* Replaces RT.resize (RT:365-372): render resolution; clears accumulation. */
int dspt_resize(dspt_ctx* ctx, int32_t width, int32_t height);

/* Replaces the uniform upload of RT.updatePathracingUniforms (RT:376-438).
 * u_frame_count in the block is ignored: dspt_render() sets it per frame. */
int dspt_set_frame(dspt_ctx* ctx, const float block[DSPT_FRAME_FLOATS]);

/* Replaces RT.render(camera, num_samples, ...) + the renderFrame loop
 * (RT:512-714): renders frames first_frame .. first_frame+num_frames-1
 * (u_frame_count is 1-based) into the accumulation buffer with the running
 * mean of renderer.frag:153-156. first_frame == 1 resets the accumulation
 * (RT.resetAccumulation, RT:327-336). Asynchronous on the context stream;
 * dspt_synchronize() or dspt_read_accum() waits. */
int dspt_render(dspt_ctx* ctx, int32_t first_frame, int32_t num_frames);
int dspt_synchronize(dspt_ctx* ctx);

/* One scissored tile pass of RT.renderFrame (RT:632-643): frames first_frame
 * .. first_frame+num_frames-1 of the pixels [x0, x0+width) x [y0, y0+height)
 * only (GL window coordinates: row 0 is the bottom row, as gl.scissor), each
 * folded into those pixels' running mean with u_frame_count = its frame
 * (renderer.frag:153-156); every other pixel keeps its value. A tile pass
 * never clears, first_frame 1 included: as in the reference, the clear
 * belongs to the reset (dspt_clear_accum). Overlapping calls for one frame
 * fold that frame again into the overlap, as overlapping scissor rectangles
 * do in the reference. */
int dspt_render_region(dspt_ctx* ctx, int32_t first_frame, int32_t num_frames, int32_t x0, int32_t y0, int32_t width,
                       int32_t height);
/* Clears the accumulation to zero: the clear of both render targets when
 * RT.renderFrame starts after RT.resetAccumulation (RT:327-336, 604-621). */
int dspt_clear_accum(dspt_ctx* ctx);

/* Accumulation readback: f32[4*W*H]. No counterpart in the reference lib
 * (the app reads the canvas, packages/app/demo_viewer.ts:246-251). */
int dspt_read_accum(dspt_ctx* ctx, float* rgba);

/* Per-sample radiance of the last dspt_render() call, for parity checks:
 * f32[4 * num_frames * W * H] ordered [frame][y][x][rgba]. Only valid when
 * dspt_set_record_samples(ctx, 1) was called before rendering. sigs receives
 * the per-sample path signature (branch/hit-id hash), prim the primary hit
 * triangle id (-1 on miss); either may be NULL. */
int dspt_set_record_samples(dspt_ctx* ctx, int enable);
int dspt_read_samples(dspt_ctx* ctx, float* rgba, uint32_t* sigs, int32_t* prim);

/* Work counters accumulated since the last reset (int64[DSPT_CNT_COUNT]). */
int dspt_get_counters(dspt_ctx* ctx, int64_t* out);
int dspt_reset_counters(dspt_ctx* ctx);

/* Per-kernel timing with HIP events recorded on the engine stream around
 * every launch of the bounce loop. Kernel classes: */
enum {
  DSPT_K_RAYGEN = 0,       /* refill of freed slots with camera rays (plan + regen) */
  DSPT_K_TRACE_EXT = 1,    /* closest-hit traversal of camera + extension rays */
  DSPT_K_SHADE = 2,
  DSPT_K_TRACE_SHADOW = 3, /* any-hit traversal of NEE shadow rays */
  DSPT_K_ACCUM = 4,        /* running mean of complete frames (accum_ready + advance) */
  DSPT_K_POST = 5,         /* display chain (dspt_postprocess / dspt_job_postprocess) */
  DSPT_K_ORDER = 6,        /* ray order of the extension queue (keys, bucket scan, scatter) */
  DSPT_K_COUNT = 8
};
/* Stats per class, double[DSPT_K_COUNT * 4]: {launches, total_ms,
 * node_visits, tri_tests} (visits/tests only for the two traversal classes).
 * Reset together with dspt_reset_counters(). */
int dspt_set_kernel_timing(dspt_ctx* ctx, int enable);
int dspt_get_kernel_stats(dspt_ctx* ctx, double* out);

/* The bounce loop of a chunk (per bounce: plan + regen, closest-hit trace,
 * shade, shadow trace) is captured once as a HIP graph per pool layout and
 * replayed; the work sizes live in device counters, so every replay is the
 * same launch sequence. On by default (the environment variable
 * DSPT_BOUNCE_GRAPHS=0 turns it off at dspt_create); kernel timing launches
 * eagerly. Results are identical either way. No counterpart in the
 * reference, whose loop is one WebGL draw per tile and frame
 * (renderer.ts:595-714). */
int dspt_set_bounce_graphs(dspt_ctx* ctx, int enable);

/* Ray order of the path queue:
 * DSPT_ORDER_BUCKETS sorts the continuing paths into 8192 buckets by (Morton
 * code of the origin's cell in a 16x16x16 grid over the scene's root box,
 * sign of dir.y) before each closest-hit trace; camera rays keep pixel
 * order, and shading follows the same order.
 * DSPT_ORDER_SHADE (default) does that and then, between the closest-hit
 * trace and shading, sorts the whole queue by the material of each hit
 * (misses together), so a wave shades one material's lobes.
 * DSPT_ORDER_NONE keeps the compaction order.
 * Results are identical in every mode (every path is computed from its own
 * state); only speed changes. DSPT_RAY_ORDER=<mode> in the environment sets
 * the mode at dspt_create; the batched renderer's autotuner tries the modes.
 * No counterpart in the reference (one path per fragment). */
enum { DSPT_ORDER_NONE = 0, DSPT_ORDER_BUCKETS = 1, DSPT_ORDER_SHADE = 2 };
int dspt_set_ray_order(dspt_ctx* ctx, int mode);

/* Display chain of renderer.ts:665-700 on the accumulation (the implicit
 * job of the single-image API, or job `job_id`): the tonemap pass
 * (SH/tonemap.frag; tonemap_mode = index into the renderer's
 * ["None", "Reinhard", "Cineon", "AcesFilm"] (renderer.ts:192): 0 exposure
 * only, 1 Reinhard, 2 Cineon, 3 ACES; 4 is the shader's Uncharted2; any other
 * value applies no operator, as indexOf() == -1 does), gamma 1/2.2 when
 * gamma != 0, the FXAA pass (SH/fxaa.frag) when fxaa != 0, and the display
 * pass (SH/display.frag) resampling to an out_w x out_h canvas (the
 * pixelRatio scaling; pass the render size for none). All buffers behave as
 * the reference's RGBA32F NEAREST + CLAMP_TO_EDGE textures
 * (gl_utils.ts:22-33). Outputs, either may be NULL: rgba8 u8[4*out_w*out_h],
 * the canvas value floor(clamp(c, 0, 1) * 255 + 0.5); rgba32
 * f32[4*out_w*out_h], the display buffer. Row 0 = bottom (glReadPixels
 * order). Synchronous. */
int dspt_postprocess(dspt_ctx* ctx, int32_t tonemap_mode, float exposure, int32_t gamma, int32_t fxaa, int32_t out_w,
                     int32_t out_h, uint8_t* rgba8, float* rgba32);
int dspt_job_postprocess(dspt_ctx* ctx, int32_t job_id, int32_t tonemap_mode, float exposure, int32_t gamma, int32_t fxaa,
                         int32_t out_w, int32_t out_h, uint8_t* rgba8, float* rgba32);

/* Traversal entry points on caller rays {org.xyz, tfar, dir.xyz, pad}
 * (f32[8*n]): closest hit (SH/bvh.glsl:214-216) -> tri id (-1 = miss), t,
 * barycentric u, v; and occlusion (SH/bvh.glsl:225-229) -> 1 if any
 * triangle lies in [1e-8, tfar). Synchronous. */
int dspt_trace_closest(dspt_ctx* ctx, const float* rays, int64_t n, int32_t* tri, float* t, float* u, float* v);
int dspt_trace_occluded(dspt_ctx* ctx, const float* rays, int64_t n, int32_t* occluded);

/* Allocation-free closest-hit query. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * results as dspt_trace_closest, but every buffer is caller-owned DEVICE
 * memory: d_rays f32[8n] ({org.xyz, tfar, dir.xyz, pad}), outputs i32[n] /
 * f32[n]. The call enqueues one kernel on the context's stream and returns;
 * the inputs must be ready before the call and dspt_synchronize() waits for
 * the outputs. It allocates nothing: the traversal-stack spill lives in a
 * fixed per-device scratch array of 24 MiB (512 blocks of 256 lanes), held
 * until the process ends. blocks = grid size (<= 0: 512); more than 512 is
 * rejected with DSPT_ERR_ARG. Calls that overlap on one GPU (from several
 * contexts or threads) share that scratch and race. */
int dspt_trace_closest_noalloc(dspt_ctx* ctx, const float* d_rays, int64_t n, int32_t* d_tri, float* d_t, float* d_u,
                               float* d_v, int32_t blocks);

/* Engine-side BVH builder emitting the TinyBVH bvh_gpu layout consumed by
 * SH/bvh.glsl:1-6 (replaces the WASM call of RT:880-882 /
 * tinybvh_builder.ts:142-215). positions f32[3*V], indices u32[3*T].
 * Two calls: with nodes == NULL it returns the sizes in *num_nodes and
 * *num_bvh_indices; then fill caller buffers of those sizes. */
int dspt_build_bvh(const float* positions, int64_t num_vertices, const uint32_t* indices, int64_t num_triangles,
                   int32_t max_leaf, float* nodes, int64_t* num_nodes, int32_t* bvh_indices, int64_t* num_bvh_indices);

/* GPU BVH build (engine extension; the reference builds on the host,
 * renderer.ts:872-912 -> tinybvh_builder.ts:142-215), emitted in the same
 * bvh_gpu layout, for admitting scenes while serving. Both methods start
 * from 63-bit Morton codes of the triangle centroids:
 *   DSPT_BVH_LBVH  Karras' linear BVH (fastest build, lower tree quality)
 *   DSPT_BVH_PLOC  parallel locally-ordered clustering (Meister & Bittner):
 *                  nearest-neighbour merges by surface area, trees close to
 *                  a SAH build
 * and collapse subtrees of at most max_leaf triangles into leaves. Inputs
 * are the scene buffers themselves (vertex stream f32[20*V], u32[3*T]).
 * dspt_scene_create_gpu_bvh / dspt_upload_scene_gpu_bvh: a resident scene
 * (or scene 0) whose BVH is built on the device; build_ms (may be NULL)
 * receives the build time measured with HIP events.
 * dspt_build_bvh_gpu: the same build returned to the host: nodes f32[16 *
 * (2T-1)] capacity, *num_nodes = nodes written, bvh_indices i32[T]. */
enum { DSPT_BVH_LBVH = 0, DSPT_BVH_PLOC = 1 };
int dspt_scene_create_gpu_bvh(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                              int64_t num_triangles, const float* materials, int32_t num_materials, int32_t max_leaf,
                              int32_t method, int32_t* scene_id, double* build_ms);
int dspt_upload_scene_gpu_bvh(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                              int64_t num_triangles, const float* materials, int32_t num_materials, int32_t max_leaf,
                              int32_t method, double* build_ms);
int dspt_build_bvh_gpu(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                       int64_t num_triangles, int32_t max_leaf, int32_t method, float* nodes, int64_t* num_nodes,
                       int32_t* bvh_indices, double* build_ms);

/* Device memory currently held by the context, bytes. */
int dspt_memory_usage(dspt_ctx* ctx, int64_t* bytes);

/* ---- Resident scenes / IBLs and render jobs (batched and served modes).
 * No counterpart in the reference, which renders one scene at a time; these
 * generalise setScene/setIBL/render to many scenes kept resident in HBM. */
int dspt_scene_create(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                      int64_t num_triangles, const float* bvh_nodes, int64_t num_nodes, const int32_t* bvh_indices,
                      int64_t num_bvh_indices, const float* materials, int32_t num_materials, int32_t* scene_id);
int dspt_scene_destroy(dspt_ctx* ctx, int32_t scene_id);

/* Two-level (instanced) scene. No counterpart in the reference, which
 * flattens instances into world space (packages/lib/scene_loader.ts:178-222,
 * 288-305); this keeps one copy per mesh. vertices f32[20*V] in object space,
 * tri_indices u32[3*T]; mesh m owns triangles [mesh_first_tri[m],
 * + mesh_num_tris[m]); instance i places mesh instance_mesh[i] with the
 * object->world affine instance_o2w[12*i .. +12] (rows of a 3x4 matrix).
 * The engine builds a bottom-level BVH per mesh and a top-level BVH over the
 * instances. Hits report the mesh triangle id and the instance. */
int dspt_scene_create_instanced(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                                int64_t num_triangles, const int64_t* mesh_first_tri, const int64_t* mesh_num_tris,
                                int32_t num_meshes, const float* instance_o2w, const int32_t* instance_mesh,
                                int32_t num_instances, const float* materials, int32_t num_materials, int32_t* scene_id);
/* The same as the single-image scene (replaces scene 0). */
int dspt_upload_scene_instanced(dspt_ctx* ctx, const float* vertices, int64_t num_vertices, const uint32_t* tri_indices,
                                int64_t num_triangles, const int64_t* mesh_first_tri, const int64_t* mesh_num_tris,
                                int32_t num_meshes, const float* instance_o2w, const int32_t* instance_mesh,
                                int32_t num_instances, const float* materials, int32_t num_materials);
/* Closest hit on scene 0 reporting the instance too (-1 on a one-level scene). */
int dspt_trace_closest_instanced(dspt_ctx* ctx, const float* rays, int64_t n, int32_t* inst, int32_t* tri, float* t, float* u,
                                 float* v);
int dspt_ibl_create(dspt_ctx* ctx, const float* rgba, int32_t width, int32_t height, double* total_sum, int32_t* ibl_id);
int dspt_ibl_destroy(dspt_ctx* ctx, int32_t ibl_id);

/* Material textures of a scene (scene 0 is the one dspt_upload_scene
 * replaces), replacing earlier ones. Replaces the reference's texture-array
 * and TexInfo upload (packages/lib/scene_data_webgl2.ts:233-363):
 *   texinfos    f32[8 * num_texinfos], TexInfo of packages/lib/material.ts:483-523
 *               (offset.xy, texture array index, layer, scale.xy, uv set, pad);
 *   array_dims  i32[3 * num_arrays]: width, height, layers of each RGBA8 array;
 *   texels      the arrays' RGBA8 texels concatenated in array order, each
 *               array layer by layer, row 0 first (num_texel_bytes in total).
 * Sampling is GL LINEAR + REPEAT without mipmaps, as the reference's arrays.
 * Rendering a scene whose materials reference a texture id >= num_texinfos
 * fails with DSPT_ERR_STATE. */
int dspt_scene_set_textures(dspt_ctx* ctx, int32_t scene_id, const float* texinfos, int32_t num_texinfos,
                            const int32_t* array_dims, int32_t num_arrays, const uint8_t* texels, int64_t num_texel_bytes);

/* A job renders frames 1..num_frames of (scene, IBL or -1, frame block, W x H)
 * into its own accumulation buffer. Admission: fails with DSPT_ERR_MEMORY when
 * the job's buffers would exceed the memory budget (dspt_set_memory_budget)
 * or, budget or not, the device's free HBM (hipMemGetInfo) less a 256 MiB
 * reserve. */
int dspt_job_create(dspt_ctx* ctx, int32_t scene_id, int32_t ibl_id, const float block[DSPT_FRAME_FLOATS], int32_t width,
                    int32_t height, int32_t num_frames, int32_t record_samples, int32_t* job_id);
int dspt_job_destroy(dspt_ctx* ctx, int32_t job_id);
int dspt_job_status(dspt_ctx* ctx, int32_t job_id, int32_t* frames_done, int32_t* frames_total);
int dspt_job_read_accum(dspt_ctx* ctx, int32_t job_id, float* rgba);
int dspt_job_read_samples(dspt_ctx* ctx, int32_t job_id, float* rgba, uint32_t* sigs, int32_t* prim);

/* One chunk of the path pool (a few bounces; see dspt_set_schedule) over the
 * listed jobs (all live jobs when job_ids is NULL): schedules their pending
 * frames, runs the bounces and folds every complete frame into its job's
 * running mean. Progress (dspt_job_status) is that of the previous chunk;
 * *active receives how many listed jobs are not finished yet. A job's
 * result does not depend on which other jobs shared the pool. */
int dspt_jobs_step(dspt_ctx* ctx, const int32_t* job_ids, int32_t n, int32_t* active);
/* Runs chunks until the listed jobs have folded all their frames. */
int dspt_jobs_run(dspt_ctx* ctx, const int32_t* job_ids, int32_t n);
int dspt_set_memory_budget(dspt_ctx* ctx, int64_t bytes);
/* Hard limit on the device memory the context allocates for scenes, IBLs,
 * textures, jobs, the path pool and per-call temporaries (0 = none, the
 * default). A call that would pass it fails with DSPT_ERR_MEMORY and frees
 * what it had allocated, so dspt_memory_usage() is unchanged. */
int dspt_set_memory_limit(dspt_ctx* ctx, int64_t bytes);

/* Launch configuration: path-pool size (paths in flight; default 16 Mi,
 * about 5 GB of path state; larger pools amortise each launch's drain over
 * more rays), traversal and shading grid caps in blocks of 256 threads
 * (defaults 2048 and 16384; <= 0 keeps a value). Chosen
 * by the autotuner in the host layer. A new pool size takes effect once the
 * work in flight drains. */
int dspt_set_launch_config(dspt_ctx* ctx, int64_t max_paths, int32_t trace_blocks, int32_t shade_blocks);
int dspt_get_launch_config(dspt_ctx* ctx, int64_t* max_paths, int32_t* trace_blocks, int32_t* shade_blocks);

/* Scheduling of the path pool.
 *   DSPT_SCHED_STREAM  frames of the listed jobs are queued round-robin ahead
 *                      of the pool; a slot freed by a finished path takes the
 *                      next sample on the next bounce (default).
 *   DSPT_SCHED_PASS    frame-synchronous passes: a new pass (up to a pool of
 *                      frames, FIFO over the listed jobs, at most pass_frames
 *                      frames of a job when > 0) starts only when the pool is
 *                      idle; pass_frames = 1 is the reference's one frame per
 *                      draw loop (RT:595-714).
 * chunk_iters: bounces per chunk (0 keeps); ring_factor: per-job ring of
 * unfolded frames, in pools (ring_factor * max_paths / pixels + 2 frames;
 * applies to jobs created afterwards; 0 keeps). */
enum { DSPT_SCHED_STREAM = 0, DSPT_SCHED_PASS = 1 };
int dspt_set_schedule(dspt_ctx* ctx, int32_t mode, int32_t pass_frames, int32_t chunk_iters, float ring_factor);
int dspt_get_schedule(dspt_ctx* ctx, int32_t* mode, int32_t* pass_frames, int32_t* chunk_iters, float* ring_factor);

/* Which listed job's frames the STREAM schedule queues next:
 *   DSPT_POLICY_RR    one frame per job in turn (fair progress; default);
 *   DSPT_POLICY_FIFO  the listed order, whole runs of frames per job, the
 *                     next job starting as soon as the pool has room (few
 *                     scenes share the caches at a time);
 *   DSPT_POLICY_SRPT  like FIFO, least remaining work (pixels x frames) first. */
enum { DSPT_POLICY_RR = 0, DSPT_POLICY_FIFO = 1, DSPT_POLICY_SRPT = 2 };
int dspt_set_policy(dspt_ctx* ctx, int32_t policy);

/* Engine self-test (no counterpart in the reference): evaluates the pinned
 * float transcendentals of pin_math.hpp on `device` over n arguments, the
 * same code the kernels and the CPU oracle's float build run. fn: 0 sin,
 * 1 cos, 2 exp, 3 log, 4 pow(x, y), 5 atan2(x, y) (x = the y-coordinate),
 * 6 acos, 7 / 8 the sine / cosine of psincos. y is read for fn 4 and 5. */
int dspt_selftest_pin_math(int device, int32_t fn, const float* x, const float* y, float* out, int64_t n);

/* Measurement aid (no counterpart in the reference): `reps` launches of a
 * plain device copy of `bytes` bytes (k_copy_calib), whose memory counters
 * (FETCH_SIZE, WRITE_SIZE) per dispatch calibrate the counters against bytes
 * actually read and written. */
int dspt_selftest_copy(int device, int64_t bytes, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif /* DSPT_H */
