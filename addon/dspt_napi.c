/*
 * N-API addon exposing the C ABI of include/dspt.h to Node, so the
 * reference's TypeScript PathtracingRenderer surface (packages/lib/renderer.ts)
 * can drive the MI355X engine. Typed arrays are passed by pointer (no copy
 * on the JS side); the engine copies what it keeps into HBM. render() runs
 * on a libuv worker (napi_async_work) and resolves a Promise, so the event
 * loop is not blocked (the reference renders from rAF/setTimeout callbacks,
 * renderer.ts:527-565). Errors become thrown JS Errors carrying
 * dspt_last_error(), as the reference throws Error objects
 * (packages/lib/bvh/tinybvh_builder.ts:48-97).
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/dspt.h"

#define CALL(env, expr)                                   \
  do {                                                    \
    if ((expr) != napi_ok) {                              \
      napi_throw_error((env), NULL, "N-API call failed: " #expr); \
      return NULL;                                        \
    }                                                     \
  } while (0)

static napi_value throw_dspt(napi_env env, int code) {
  char msg[1024];
  snprintf(msg, sizeof msg, "dspt error %d: %s", code, dspt_last_error());
  napi_throw_error(env, NULL, msg);
  return NULL;
}

/* An engine handle: the context plus a flag set while an async call (render,
 * jobsStep, jobsRun) owns it; the context is not re-entrant, so other calls
 * on the handle throw until that call settles. Both sides of the flag run on
 * the JS thread. */
typedef struct {
  dspt_ctx* c;
  int busy;
} Handle;

static Handle* get_handle(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected an engine handle");
    return NULL;
  }
  return (Handle*)p;
}

static dspt_ctx* get_ctx(napi_env env, napi_value v) {
  Handle* h = get_handle(env, v);
  if (!h) return NULL;
  if (h->busy) {
    napi_throw_error(env, NULL, "engine busy: an asynchronous call on this handle has not settled");
    return NULL;
  }
  return h->c;
}

static int get_typed(napi_env env, napi_value v, napi_typedarray_type want, void** data, size_t* len) {
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, NULL, "unexpected typed array type");
    return -1;
  }
  return 0;
}

static int64_t get_i64(napi_env env, napi_value v) {
  int64_t x = 0;
  napi_get_value_int64(env, v, &x);
  return x;
}

static void finalize_ctx(napi_env env, void* data, void* hint) {
  (void)env; (void)hint;
  Handle* h = (Handle*)data;
  if (h) {
    dspt_destroy(h->c);
    free(h);
  }
}

/* create(profile, device) -> handle */
static napi_value js_create(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int profile = argc > 0 ? (int)get_i64(env, argv[0]) : 0;
  int device = argc > 1 ? (int)get_i64(env, argv[1]) : 0;
  dspt_ctx* c = NULL;
  int rc = dspt_create(profile, device, &c);
  if (rc) return throw_dspt(env, rc);
  Handle* hd = (Handle*)calloc(1, sizeof(Handle));
  hd->c = c;
  napi_value h;
  CALL(env, napi_create_external(env, hd, finalize_ctx, NULL, &h));
  return h;
}

/* uploadScene(h, verts f32, tris u32, nodes f32, bvhIdx i32, mats f32) */
static napi_value js_upload_scene(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 6) { napi_throw_error(env, NULL, "uploadScene expects 6 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void *v, *t, *n, *b, *m;
  size_t lv, lt, ln, lb, lm;
  if (get_typed(env, argv[1], napi_float32_array, &v, &lv) || get_typed(env, argv[2], napi_uint32_array, &t, &lt) ||
      get_typed(env, argv[3], napi_float32_array, &n, &ln) || get_typed(env, argv[4], napi_int32_array, &b, &lb) ||
      get_typed(env, argv[5], napi_float32_array, &m, &lm))
    return NULL;
  int rc = dspt_upload_scene(c, (const float*)v, (int64_t)(lv / 20), (const uint32_t*)t, (int64_t)(lt / 3), (const float*)n,
                             (int64_t)(ln / 16), (const int32_t*)b, (int64_t)lb, (const float*)m, (int32_t)(lm / 68));
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

/* uploadIbl(h, rgba f32, w, h) -> totalSum */
static napi_value js_upload_ibl(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[1], napi_float32_array, &d, &l)) return NULL;
  int w = (int)get_i64(env, argv[2]), h = (int)get_i64(env, argv[3]);
  if ((size_t)w * h * 4 != l) { napi_throw_range_error(env, NULL, "IBL data length != 4*w*h"); return NULL; }
  double tot = 0;
  int rc = dspt_upload_ibl(c, (const float*)d, w, h, &tot);
  if (rc) return throw_dspt(env, rc);
  napi_value r;
  CALL(env, napi_create_double(env, tot, &r));
  return r;
}

static napi_value js_resize(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  int rc = dspt_resize(c, (int32_t)get_i64(env, argv[1]), (int32_t)get_i64(env, argv[2]));
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

/* setFrame(h, Float32Array(52)) */
static napi_value js_set_frame(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[1], napi_float32_array, &d, &l)) return NULL;
  if (l != DSPT_FRAME_FLOATS) { napi_throw_range_error(env, NULL, "frame block must hold 52 floats"); return NULL; }
  int rc = dspt_set_frame(c, (const float*)d);
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

enum { OP_RENDER = 0, OP_STEP = 1, OP_RUN = 2, OP_REGION = 3 };

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref handle_ref;
  Handle* h;
  int op, first, count, rc, active;
  int x0, y0, rw, rh; /* OP_REGION: the tile */
  int32_t* ids;
  int nids;
  char err[512];
} AsyncCall;

static void async_execute(napi_env env, void* data) {
  (void)env;
  AsyncCall* j = (AsyncCall*)data;
  dspt_ctx* c = j->h->c;
  if (j->op == OP_RENDER) {
    j->rc = dspt_render(c, j->first, j->count);
    if (!j->rc) j->rc = dspt_synchronize(c);
  } else if (j->op == OP_REGION) {
    j->rc = dspt_render_region(c, j->first, j->count, j->x0, j->y0, j->rw, j->rh);
    if (!j->rc) j->rc = dspt_synchronize(c);
  } else if (j->op == OP_STEP) {
    j->rc = dspt_jobs_step(c, j->ids, j->nids, &j->active);
  } else {
    j->rc = dspt_jobs_run(c, j->ids, j->nids);
    if (!j->rc) j->rc = dspt_synchronize(c);
  }
  if (j->rc) snprintf(j->err, sizeof j->err, "dspt error %d: %s", j->rc, dspt_last_error());
}

static void async_complete(napi_env env, napi_status status, void* data) {
  AsyncCall* j = (AsyncCall*)data;
  j->h->busy = 0;
  napi_value v;
  if (status == napi_ok && j->rc == 0) {
    if (j->op == OP_STEP) napi_create_int32(env, j->active, &v);
    else napi_get_undefined(env, &v);
    napi_resolve_deferred(env, j->deferred, v);
  } else {
    napi_value msg;
    napi_create_string_utf8(env, j->rc ? j->err : "call cancelled", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &v);
    napi_reject_deferred(env, j->deferred, v);
  }
  napi_delete_reference(env, j->handle_ref);
  napi_delete_async_work(env, j->work);
  free(j->ids);
  free(j);
}

static napi_value queue_call(napi_env env, napi_value handle, AsyncCall* j, const char* what) {
  napi_value promise, name;
  j->h->busy = 1;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok ||
      napi_create_reference(env, handle, 1, &j->handle_ref) != napi_ok ||
      napi_create_string_utf8(env, what, NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, NULL, name, async_execute, async_complete, j, &j->work) != napi_ok ||
      napi_queue_async_work(env, j->work) != napi_ok) {
    j->h->busy = 0;
    napi_throw_error(env, NULL, "could not queue the asynchronous call");
    return NULL;
  }
  return promise;
}

/* render(h, firstFrame, numFrames) -> Promise<void> */
static napi_value js_render(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (!get_ctx(env, argv[0])) return NULL;
  AsyncCall* j = (AsyncCall*)calloc(1, sizeof(AsyncCall));
  j->h = get_handle(env, argv[0]);
  j->op = OP_RENDER;
  j->first = (int)get_i64(env, argv[1]);
  j->count = (int)get_i64(env, argv[2]);
  return queue_call(env, argv[0], j, "dspt_render");
}

/* renderRegion(h, firstFrame, numFrames, x0, y0, width, height) -> Promise<void>:
 * one scissored tile pass (rows from the bottom) */
static napi_value js_render_region(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value argv[7];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 7) { napi_throw_type_error(env, NULL, "renderRegion(h, first, count, x0, y0, width, height)"); return NULL; }
  if (!get_ctx(env, argv[0])) return NULL;
  AsyncCall* j = (AsyncCall*)calloc(1, sizeof(AsyncCall));
  j->h = get_handle(env, argv[0]);
  j->op = OP_REGION;
  j->first = (int)get_i64(env, argv[1]);
  j->count = (int)get_i64(env, argv[2]);
  j->x0 = (int)get_i64(env, argv[3]);
  j->y0 = (int)get_i64(env, argv[4]);
  j->rw = (int)get_i64(env, argv[5]);
  j->rh = (int)get_i64(env, argv[6]);
  return queue_call(env, argv[0], j, "dspt_render_region");
}

/* jobsStep(h, Int32Array ids | null) -> Promise<number of unfinished jobs>;
 * jobsRun(h, Int32Array ids | null) -> Promise<void> */
static napi_value jobs_call(napi_env env, napi_callback_info info, int op) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (!get_ctx(env, argv[0])) return NULL;
  AsyncCall* j = (AsyncCall*)calloc(1, sizeof(AsyncCall));
  j->h = get_handle(env, argv[0]);
  j->op = op;
  napi_valuetype t = napi_null;
  if (argc > 1) napi_typeof(env, argv[1], &t);
  if (argc > 1 && t != napi_null && t != napi_undefined) {
    void* d;
    size_t l;
    if (get_typed(env, argv[1], napi_int32_array, &d, &l)) { free(j); return NULL; }
    j->ids = (int32_t*)malloc(l * sizeof(int32_t) + 1);
    memcpy(j->ids, d, l * sizeof(int32_t));
    j->nids = (int)l;
  }
  return queue_call(env, argv[0], j, op == OP_STEP ? "dspt_jobs_step" : "dspt_jobs_run");
}
static napi_value js_jobs_step(napi_env env, napi_callback_info info) { return jobs_call(env, info, OP_STEP); }
static napi_value js_jobs_run(napi_env env, napi_callback_info info) { return jobs_call(env, info, OP_RUN); }

/* sceneCreate(h, verts f32, tris u32, nodes f32, bvhIdx i32, mats f32) -> scene id */
static napi_value js_scene_create(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 6) { napi_throw_error(env, NULL, "sceneCreate expects 6 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void *v, *t, *n, *b, *m;
  size_t lv, lt, ln, lb, lm;
  if (get_typed(env, argv[1], napi_float32_array, &v, &lv) || get_typed(env, argv[2], napi_uint32_array, &t, &lt) ||
      get_typed(env, argv[3], napi_float32_array, &n, &ln) || get_typed(env, argv[4], napi_int32_array, &b, &lb) ||
      get_typed(env, argv[5], napi_float32_array, &m, &lm))
    return NULL;
  int32_t id = -1;
  int rc = dspt_scene_create(c, (const float*)v, (int64_t)(lv / 20), (const uint32_t*)t, (int64_t)(lt / 3), (const float*)n,
                             (int64_t)(ln / 16), (const int32_t*)b, (int64_t)lb, (const float*)m, (int32_t)(lm / 68), &id);
  if (rc) return throw_dspt(env, rc);
  napi_value r;
  CALL(env, napi_create_int32(env, id, &r));
  return r;
}

/* sceneCreateInstanced(h, verts f32, tris u32, meshFirst i32, meshCount i32, o2w f32[12I], instMesh i32[I], mats f32)
 * -> scene id (two-level scene: object-space meshes placed by instances) */
static napi_value js_scene_create_instanced(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value argv[8];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 8) { napi_throw_error(env, NULL, "sceneCreateInstanced expects 8 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void *v, *t, *mf, *mc, *x, *im, *m;
  size_t lv, lt, lmf, lmc, lx, lim, lm;
  if (get_typed(env, argv[1], napi_float32_array, &v, &lv) || get_typed(env, argv[2], napi_uint32_array, &t, &lt) ||
      get_typed(env, argv[3], napi_int32_array, &mf, &lmf) || get_typed(env, argv[4], napi_int32_array, &mc, &lmc) ||
      get_typed(env, argv[5], napi_float32_array, &x, &lx) || get_typed(env, argv[6], napi_int32_array, &im, &lim) ||
      get_typed(env, argv[7], napi_float32_array, &m, &lm))
    return NULL;
  if (lmf != lmc || lx != 12 * lim) {
    napi_throw_range_error(env, NULL, "mesh ranges or instance arrays have mismatched lengths");
    return NULL;
  }
  int64_t* first = (int64_t*)malloc((lmf + 1) * sizeof(int64_t));
  int64_t* count = (int64_t*)malloc((lmf + 1) * sizeof(int64_t));
  if (!first || !count) {
    free(first);
    free(count);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  for (size_t i = 0; i < lmf; i++) {
    first[i] = ((const int32_t*)mf)[i];
    count[i] = ((const int32_t*)mc)[i];
  }
  int32_t id = -1;
  int rc = dspt_scene_create_instanced(c, (const float*)v, (int64_t)(lv / 20), (const uint32_t*)t, (int64_t)(lt / 3), first,
                                       count, (int32_t)lmf, (const float*)x, (const int32_t*)im, (int32_t)lim,
                                       (const float*)m, (int32_t)(lm / 68), &id);
  free(first);
  free(count);
  if (rc) return throw_dspt(env, rc);
  napi_value r;
  CALL(env, napi_create_int32(env, id, &r));
  return r;
}

/* sceneCreateGpuBvh(h, verts f32, tris u32, mats f32, maxLeaf, method) -> {id, buildMs}: the BVH is built on the GPU */
static napi_value js_scene_create_gpu_bvh(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 6) { napi_throw_error(env, NULL, "sceneCreateGpuBvh expects 6 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void *v, *t, *m;
  size_t lv, lt, lm;
  if (get_typed(env, argv[1], napi_float32_array, &v, &lv) || get_typed(env, argv[2], napi_uint32_array, &t, &lt) ||
      get_typed(env, argv[3], napi_float32_array, &m, &lm))
    return NULL;
  int32_t id = -1;
  double ms = 0.0;
  int rc = dspt_scene_create_gpu_bvh(c, (const float*)v, (int64_t)(lv / 20), (const uint32_t*)t, (int64_t)(lt / 3),
                                     (const float*)m, (int32_t)(lm / 68), (int32_t)get_i64(env, argv[4]),
                                     (int32_t)get_i64(env, argv[5]), &id, &ms);
  if (rc) return throw_dspt(env, rc);
  napi_value r, vid, vms;
  CALL(env, napi_create_object(env, &r));
  CALL(env, napi_create_int32(env, id, &vid));
  CALL(env, napi_create_double(env, ms, &vms));
  CALL(env, napi_set_named_property(env, r, "id", vid));
  CALL(env, napi_set_named_property(env, r, "buildMs", vms));
  return r;
}

/* sceneSetTextures(h, sceneId, texinfos f32[8X], dims i32[3A], texels u8) */
static napi_value js_scene_set_textures(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 5) { napi_throw_error(env, NULL, "sceneSetTextures expects 5 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void *ti, *dm, *tx;
  size_t lti, ldm, ltx;
  if (get_typed(env, argv[2], napi_float32_array, &ti, &lti) || get_typed(env, argv[3], napi_int32_array, &dm, &ldm) ||
      get_typed(env, argv[4], napi_uint8_array, &tx, &ltx))
    return NULL;
  int rc = dspt_scene_set_textures(c, (int32_t)get_i64(env, argv[1]), (const float*)ti, (int32_t)(lti / 8),
                                   (const int32_t*)dm, (int32_t)(ldm / 3), (const uint8_t*)tx, (int64_t)ltx);
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

/* iblCreate(h, rgba f32, w, h) -> {id, totalSum} */
static napi_value js_ibl_create(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[1], napi_float32_array, &d, &l)) return NULL;
  int w = (int)get_i64(env, argv[2]), hh = (int)get_i64(env, argv[3]);
  if ((size_t)w * hh * 4 != l) { napi_throw_range_error(env, NULL, "IBL data length != 4*w*h"); return NULL; }
  double tot = 0;
  int32_t id = -1;
  int rc = dspt_ibl_create(c, (const float*)d, w, hh, &tot, &id);
  if (rc) return throw_dspt(env, rc);
  napi_value obj, vid, vt;
  CALL(env, napi_create_object(env, &obj));
  CALL(env, napi_create_int32(env, id, &vid));
  CALL(env, napi_create_double(env, tot, &vt));
  CALL(env, napi_set_named_property(env, obj, "id", vid));
  CALL(env, napi_set_named_property(env, obj, "totalSum", vt));
  return obj;
}

/* jobCreate(h, scene, ibl, Float32Array(52) block, w, h, frames) -> job id */
static napi_value js_job_create(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value argv[7];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 7) { napi_throw_error(env, NULL, "jobCreate expects 7 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[3], napi_float32_array, &d, &l)) return NULL;
  if (l != DSPT_FRAME_FLOATS) { napi_throw_range_error(env, NULL, "frame block must hold 52 floats"); return NULL; }
  int32_t id = -1;
  int rc = dspt_job_create(c, (int32_t)get_i64(env, argv[1]), (int32_t)get_i64(env, argv[2]), (const float*)d,
                           (int32_t)get_i64(env, argv[4]), (int32_t)get_i64(env, argv[5]), (int32_t)get_i64(env, argv[6]), 0, &id);
  if (rc) return throw_dspt(env, rc);
  napi_value r;
  CALL(env, napi_create_int32(env, id, &r));
  return r;
}

/* jobStatus(h, id) -> [framesDone, framesTotal] */
static napi_value js_job_status(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  int32_t d = 0, t = 0;
  int rc = dspt_job_status(c, (int32_t)get_i64(env, argv[1]), &d, &t);
  if (rc) return throw_dspt(env, rc);
  napi_value arr, vd, vt;
  CALL(env, napi_create_array_with_length(env, 2, &arr));
  CALL(env, napi_create_int32(env, d, &vd));
  CALL(env, napi_create_int32(env, t, &vt));
  CALL(env, napi_set_element(env, arr, 0, vd));
  CALL(env, napi_set_element(env, arr, 1, vt));
  return arr;
}

/* jobReadAccum(h, id, Float32Array out) */
static napi_value js_job_read_accum(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[2], napi_float32_array, &d, &l)) return NULL;
  int rc = dspt_job_read_accum(c, (int32_t)get_i64(env, argv[1]), (float*)d);
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

/* small setters: jobDestroy(h, id), sceneDestroy(h, id), iblDestroy(h, id),
 * setMemoryBudget(h, bytes), setLaunchConfig(h, maxPaths, traceBlocks,
 * shadeBlocks), setSchedule(h, mode, passFrames, chunkIters, ringFactor),
 * setPolicy(h, policy) */
enum { S_JOB_DESTROY, S_SCENE_DESTROY, S_IBL_DESTROY, S_BUDGET, S_LAUNCH, S_SCHEDULE, S_POLICY, S_CLEAR_ACCUM };
static napi_value setter(napi_env env, napi_callback_info info, int what) {
  size_t argc = 5;
  napi_value argv[5];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  int64_t a[4] = {0, 0, 0, 0};
  double rf = 0.0;
  for (size_t i = 1; i < argc && i < 5; i++) a[i - 1] = get_i64(env, argv[i]);
  if (what == S_SCHEDULE && argc > 4) napi_get_value_double(env, argv[4], &rf);
  int rc = 0;
  switch (what) {
    case S_JOB_DESTROY: rc = dspt_job_destroy(c, (int32_t)a[0]); break;
    case S_SCENE_DESTROY: rc = dspt_scene_destroy(c, (int32_t)a[0]); break;
    case S_IBL_DESTROY: rc = dspt_ibl_destroy(c, (int32_t)a[0]); break;
    case S_BUDGET: rc = dspt_set_memory_budget(c, a[0]); break;
    case S_LAUNCH: rc = dspt_set_launch_config(c, a[0], (int32_t)a[1], (int32_t)a[2]); break;
    case S_SCHEDULE: rc = dspt_set_schedule(c, (int32_t)a[0], (int32_t)a[1], (int32_t)a[2], (float)rf); break;
    case S_POLICY: rc = dspt_set_policy(c, (int32_t)a[0]); break;
    case S_CLEAR_ACCUM: rc = dspt_clear_accum(c); break;
  }
  if (rc) return throw_dspt(env, rc);
  return NULL;
}
static napi_value js_job_destroy(napi_env e, napi_callback_info i) { return setter(e, i, S_JOB_DESTROY); }
static napi_value js_clear_accum(napi_env e, napi_callback_info i) { return setter(e, i, S_CLEAR_ACCUM); }
static napi_value js_scene_destroy(napi_env e, napi_callback_info i) { return setter(e, i, S_SCENE_DESTROY); }
static napi_value js_ibl_destroy(napi_env e, napi_callback_info i) { return setter(e, i, S_IBL_DESTROY); }
static napi_value js_set_budget(napi_env e, napi_callback_info i) { return setter(e, i, S_BUDGET); }
static napi_value js_set_launch(napi_env e, napi_callback_info i) { return setter(e, i, S_LAUNCH); }
static napi_value js_set_schedule(napi_env e, napi_callback_info i) { return setter(e, i, S_SCHEDULE); }
static napi_value js_set_policy(napi_env e, napi_callback_info i) { return setter(e, i, S_POLICY); }

/* memoryUsage(h) -> bytes */
static napi_value js_memory_usage(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  int64_t b = 0;
  int rc = dspt_memory_usage(c, &b);
  if (rc) return throw_dspt(env, rc);
  napi_value r;
  CALL(env, napi_create_double(env, (double)b, &r));
  return r;
}

/* readAccum(h, Float32Array out) */
static napi_value js_read_accum(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  void* d;
  size_t l;
  if (get_typed(env, argv[1], napi_float32_array, &d, &l)) return NULL;
  int rc = dspt_read_accum(c, (float*)d);
  if (rc) return throw_dspt(env, rc);
  return NULL;
}

/* postprocess(h, jobId (-1: the single-image accumulation), tonemapMode, exposure, gamma, fxaa, outW, outH)
 * -> Uint8Array(4*outW*outH), the display chain's RGBA8 canvas, row 0 = bottom */
static napi_value js_postprocess(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value argv[8];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  if (argc < 8) { napi_throw_error(env, NULL, "postprocess expects 8 arguments"); return NULL; }
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  const int32_t job = (int32_t)get_i64(env, argv[1]), mode = (int32_t)get_i64(env, argv[2]);
  double exposure = 1.0;
  CALL(env, napi_get_value_double(env, argv[3], &exposure));
  bool gamma = false, fxaa = false;
  CALL(env, napi_get_value_bool(env, argv[4], &gamma));
  CALL(env, napi_get_value_bool(env, argv[5], &fxaa));
  const int64_t ow = get_i64(env, argv[6]), oh = get_i64(env, argv[7]);
  if (ow <= 0 || oh <= 0 || ow * oh > (1LL << 28)) { napi_throw_range_error(env, NULL, "invalid canvas size"); return NULL; }
  void* data = NULL;
  napi_value ab, out;
  CALL(env, napi_create_arraybuffer(env, (size_t)(4 * ow * oh), &data, &ab));
  int rc = job < 0 ? dspt_postprocess(c, mode, (float)exposure, gamma, fxaa, (int32_t)ow, (int32_t)oh, (uint8_t*)data, NULL)
                   : dspt_job_postprocess(c, job, mode, (float)exposure, gamma, fxaa, (int32_t)ow, (int32_t)oh, (uint8_t*)data,
                                          NULL);
  if (rc) return throw_dspt(env, rc);
  CALL(env, napi_create_typedarray(env, napi_uint8_array, (size_t)(4 * ow * oh), ab, 0, &out));
  return out;
}

/* counters(h) -> Float64Array(16) */
static napi_value js_counters(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  int64_t cnt[DSPT_CNT_COUNT];
  int rc = dspt_get_counters(c, cnt);
  if (rc) return throw_dspt(env, rc);
  napi_value ab, arr;
  void* p;
  CALL(env, napi_create_arraybuffer(env, DSPT_CNT_COUNT * sizeof(double), &p, &ab));
  for (int i = 0; i < DSPT_CNT_COUNT; i++) ((double*)p)[i] = (double)cnt[i];
  CALL(env, napi_create_typedarray(env, napi_float64_array, DSPT_CNT_COUNT, ab, 0, &arr));
  return arr;
}

/* buildBvh(positions f32, indices u32, maxLeaf) -> {nodeData, triangleIndices} */
static napi_value js_build_bvh(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *p, *ix;
  size_t lp, li;
  if (get_typed(env, argv[0], napi_float32_array, &p, &lp) || get_typed(env, argv[1], napi_uint32_array, &ix, &li)) return NULL;
  int max_leaf = argc > 2 ? (int)get_i64(env, argv[2]) : 4;
  int64_t nn = 0, nm = 0;
  int rc = dspt_build_bvh((const float*)p, (int64_t)(lp / 3), (const uint32_t*)ix, (int64_t)(li / 3), max_leaf, NULL, &nn,
                          NULL, &nm);
  if (rc) return throw_dspt(env, rc);
  napi_value abn, abi, an, ai, obj;
  void *dn, *di;
  CALL(env, napi_create_arraybuffer(env, (size_t)nn * 64, &dn, &abn));
  CALL(env, napi_create_arraybuffer(env, (size_t)nm * 4, &di, &abi));
  rc = dspt_build_bvh((const float*)p, (int64_t)(lp / 3), (const uint32_t*)ix, (int64_t)(li / 3), max_leaf, (float*)dn, &nn,
                      (int32_t*)di, &nm);
  if (rc) return throw_dspt(env, rc);
  CALL(env, napi_create_typedarray(env, napi_float32_array, (size_t)nn * 16, abn, 0, &an));
  CALL(env, napi_create_typedarray(env, napi_int32_array, (size_t)nm, abi, 0, &ai));
  CALL(env, napi_create_object(env, &obj));
  CALL(env, napi_set_named_property(env, obj, "nodeData", an));
  CALL(env, napi_set_named_property(env, obj, "triangleIndices", ai));
  return obj;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value s;
  CALL(env, napi_create_string_utf8(env, dspt_version(), NAPI_AUTO_LENGTH, &s));
  return s;
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      {"create", NULL, js_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"uploadScene", NULL, js_upload_scene, NULL, NULL, NULL, napi_enumerable, NULL},
      {"uploadIbl", NULL, js_upload_ibl, NULL, NULL, NULL, napi_enumerable, NULL},
      {"resize", NULL, js_resize, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setFrame", NULL, js_set_frame, NULL, NULL, NULL, napi_enumerable, NULL},
      {"render", NULL, js_render, NULL, NULL, NULL, napi_enumerable, NULL},
      {"renderRegion", NULL, js_render_region, NULL, NULL, NULL, napi_enumerable, NULL},
      {"clearAccum", NULL, js_clear_accum, NULL, NULL, NULL, napi_enumerable, NULL},
      {"readAccum", NULL, js_read_accum, NULL, NULL, NULL, napi_enumerable, NULL},
      {"postprocess", NULL, js_postprocess, NULL, NULL, NULL, napi_enumerable, NULL},
      {"counters", NULL, js_counters, NULL, NULL, NULL, napi_enumerable, NULL},
      {"buildBvh", NULL, js_build_bvh, NULL, NULL, NULL, napi_enumerable, NULL},
      {"version", NULL, js_version, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sceneCreate", NULL, js_scene_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sceneDestroy", NULL, js_scene_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sceneSetTextures", NULL, js_scene_set_textures, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sceneCreateInstanced", NULL, js_scene_create_instanced, NULL, NULL, NULL, napi_enumerable, NULL},
      {"sceneCreateGpuBvh", NULL, js_scene_create_gpu_bvh, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iblCreate", NULL, js_ibl_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"iblDestroy", NULL, js_ibl_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobCreate", NULL, js_job_create, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobDestroy", NULL, js_job_destroy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobStatus", NULL, js_job_status, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobReadAccum", NULL, js_job_read_accum, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobsStep", NULL, js_jobs_step, NULL, NULL, NULL, napi_enumerable, NULL},
      {"jobsRun", NULL, js_jobs_run, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setMemoryBudget", NULL, js_set_budget, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setLaunchConfig", NULL, js_set_launch, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setSchedule", NULL, js_set_schedule, NULL, NULL, NULL, napi_enumerable, NULL},
      {"setPolicy", NULL, js_set_policy, NULL, NULL, NULL, napi_enumerable, NULL},
      {"memoryUsage", NULL, js_memory_usage, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE_INIT() { return init(env, exports); }
