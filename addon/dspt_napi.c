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

static dspt_ctx* get_ctx(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected an engine handle");
    return NULL;
  }
  return (dspt_ctx*)p;
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
  if (data) dspt_destroy((dspt_ctx*)data);
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
  napi_value h;
  CALL(env, napi_create_external(env, c, finalize_ctx, NULL, &h));
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

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref handle_ref;
  dspt_ctx* c;
  int first, count, rc;
  char err[512];
} RenderJob;

static void render_execute(napi_env env, void* data) {
  (void)env;
  RenderJob* j = (RenderJob*)data;
  j->rc = dspt_render(j->c, j->first, j->count);
  if (!j->rc) j->rc = dspt_synchronize(j->c);
  if (j->rc) snprintf(j->err, sizeof j->err, "dspt error %d: %s", j->rc, dspt_last_error());
}

static void render_complete(napi_env env, napi_status status, void* data) {
  RenderJob* j = (RenderJob*)data;
  napi_value v;
  if (status == napi_ok && j->rc == 0) {
    napi_get_undefined(env, &v);
    napi_resolve_deferred(env, j->deferred, v);
  } else {
    napi_value msg;
    napi_create_string_utf8(env, j->rc ? j->err : "render cancelled", NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, NULL, msg, &v);
    napi_reject_deferred(env, j->deferred, v);
  }
  napi_delete_reference(env, j->handle_ref);
  napi_delete_async_work(env, j->work);
  free(j);
}

/* render(h, firstFrame, numFrames) -> Promise<void> */
static napi_value js_render(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CALL(env, napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  dspt_ctx* c = get_ctx(env, argv[0]);
  if (!c) return NULL;
  RenderJob* j = (RenderJob*)calloc(1, sizeof(RenderJob));
  j->c = c;
  j->first = (int)get_i64(env, argv[1]);
  j->count = (int)get_i64(env, argv[2]);
  napi_value promise, name;
  CALL(env, napi_create_promise(env, &j->deferred, &promise));
  CALL(env, napi_create_reference(env, argv[0], 1, &j->handle_ref));
  CALL(env, napi_create_string_utf8(env, "dspt_render", NAPI_AUTO_LENGTH, &name));
  CALL(env, napi_create_async_work(env, NULL, name, render_execute, render_complete, j, &j->work));
  CALL(env, napi_queue_async_work(env, j->work));
  return promise;
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
      {"readAccum", NULL, js_read_accum, NULL, NULL, NULL, napi_enumerable, NULL},
      {"counters", NULL, js_counters, NULL, NULL, NULL, napi_enumerable, NULL},
      {"buildBvh", NULL, js_build_bvh, NULL, NULL, NULL, napi_enumerable, NULL},
      {"version", NULL, js_version, NULL, NULL, NULL, napi_enumerable, NULL},
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE_INIT() { return init(env, exports); }
