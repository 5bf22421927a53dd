/*
 * N-API addon over libgsr (include/gsr.h): the binding a JavaScript trainer
 * loads in place of WebGPU (trainer/gpu.js:704-712).  One function per C entry
 * point; device pointers travel as BigInt, uniforms (camera 80 B, step 32 B) as
 * ArrayBuffers or typed arrays laid out as cameraUniform / the Step struct
 * (gpu.js:696-701, 866-869).  A non-zero status throws an Error carrying
 * gsr_last_error() and a numeric `code` (as the reference throws from
 * createRenderer, gpu.js:705).
 *
 * Threading: calls run on the JavaScript thread and queue work on the context's
 * stream.  read() (the reference's read(), gpu.js:716-726) and binSortAsync()
 * wait for the device on a worker thread (napi_async_work) and settle a Promise,
 * so the event loop stays free.  While one is pending its context is busy: any
 * other call on that context (or on a batch of it) throws, so only one thread
 * ever uses a context.  download() and binSort() are the blocking forms.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gsr.h"

#define CHECK_NAPI(call)                                             \
  do {                                                               \
    if ((call) != napi_ok) {                                         \
      napi_throw_error(env, NULL, "N-API call failed: " #call);      \
      return NULL;                                                   \
    }                                                                \
  } while (0)

static napi_value throw_gsr(napi_env env, gsr_ctx* ctx, int rc) {
  char msg[600];
  snprintf(msg, sizeof(msg), "gsr error %d: %s", rc, gsr_last_error(ctx));
  napi_value err, m, code;
  napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &m);
  napi_create_error(env, NULL, m, &err);
  napi_create_int32(env, rc, &code);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
  return NULL;
}

/* A context as JavaScript holds it: busy while an asynchronous call owns it. */
typedef struct {
  gsr_ctx* ctx;
  int busy;
  int batches;   /* live batches made on this context */
  int released;  /* JavaScript dropped the context: destroy it with its last batch */
} addon_ctx;

/* Destroys the context once JavaScript has dropped it and no batch uses it
 * (at teardown Node may finalize the context before its batches). */
static void ctx_release(addon_ctx* a) {
  if (a->released && a->batches == 0) {
    gsr_destroy(a->ctx);
    free(a);
  }
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) {
    napi_throw_type_error(env, NULL, "too few arguments");
    return 0;
  }
  return 1;
}

static addon_ctx* get_actx(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a gsr context");
    return NULL;
  }
  addon_ctx* a = (addon_ctx*)p;
  if (a->busy) {
    napi_throw_error(env, NULL, "gsr context busy: await its pending read() or binSortAsync() first");
    return NULL;
  }
  return a;
}

static gsr_ctx* get_ctx(napi_env env, napi_value v) {
  addon_ctx* a = get_actx(env, v);
  return a ? a->ctx : NULL;
}

/* Device pointer: BigInt, or null/undefined for NULL. */
static int get_ptr(napi_env env, napi_value v, void** out) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_null || t == napi_undefined) {
    *out = NULL;
    return 1;
  }
  uint64_t x = 0;
  bool lossless = true;
  if (napi_get_value_bigint_uint64(env, v, &x, &lossless) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a device pointer (BigInt)");
    return 0;
  }
  *out = (void*)(uintptr_t)x;
  return 1;
}

static int get_u32(napi_env env, napi_value v, uint32_t* out) {
  if (napi_get_value_uint32(env, v, out) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected an unsigned integer");
    return 0;
  }
  return 1;
}

static int get_f32(napi_env env, napi_value v, float* out) {
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a number");
    return 0;
  }
  *out = (float)d;
  return 1;
}

static int get_u64(napi_env env, napi_value v, uint64_t* out) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_bigint) {
    bool lossless;
    return napi_get_value_bigint_uint64(env, v, out, &lossless) == napi_ok;
  }
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok || d < 0) {
    napi_throw_type_error(env, NULL, "expected a size");
    return 0;
  }
  *out = (uint64_t)d;
  return 1;
}

/* Bytes of an ArrayBuffer or typed array. */
static int get_bytes(napi_env env, napi_value v, void** data, size_t* len) {
  bool is_ab = false, is_ta = false;
  napi_is_arraybuffer(env, v, &is_ab);
  if (is_ab) return napi_get_arraybuffer_info(env, v, data, len) == napi_ok;
  napi_is_typedarray(env, v, &is_ta);
  if (is_ta) {
    napi_typedarray_type type;
    size_t length, offset;
    napi_value ab;
    if (napi_get_typedarray_info(env, v, &type, &length, data, &ab, &offset) != napi_ok) return 0;
    size_t el = 1;
    switch (type) {
      case napi_int16_array: case napi_uint16_array: el = 2; break;
      case napi_int32_array: case napi_uint32_array: case napi_float32_array: el = 4; break;
      case napi_float64_array: case napi_bigint64_array: case napi_biguint64_array: el = 8; break;
      default: el = 1;
    }
    *len = length * el;
    return 1;
  }
  napi_throw_type_error(env, NULL, "expected an ArrayBuffer or typed array");
  return 0;
}

static int get_struct(napi_env env, napi_value v, void* dst, size_t size) {
  void* data;
  size_t len;
  if (!get_bytes(env, v, &data, &len)) return 0;
  if (len < size) {
    napi_throw_range_error(env, NULL, "uniform buffer too small");
    return 0;
  }
  memcpy(dst, data, size);
  return 1;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static void ctx_finalize(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  addon_ctx* a = (addon_ctx*)data;
  a->released = 1;
  ctx_release(a);
}

/* create(device) -> context */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  uint32_t dev;
  if (!get_u32(env, argv[0], &dev)) return NULL;
  gsr_ctx* ctx = NULL;
  int rc = gsr_create((int)dev, &ctx);
  if (rc) return throw_gsr(env, NULL, rc);
  addon_ctx* a = (addon_ctx*)calloc(1, sizeof(addon_ctx));
  if (!a) {
    gsr_destroy(ctx);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  a->ctx = ctx;
  napi_value ext;
  if (napi_create_external(env, a, ctx_finalize, NULL, &ext) != napi_ok) {
    gsr_destroy(ctx);
    free(a);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return ext;
}

/* alloc(ctx, bytes) -> BigInt pointer (zero-filled) */
static napi_value js_alloc(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  uint64_t bytes;
  if (!ctx || !get_u64(env, argv[1], &bytes)) return NULL;
  void* p = NULL;
  int rc = gsr_alloc(ctx, (size_t)bytes, &p);
  if (rc) return throw_gsr(env, ctx, rc);
  napi_value out;
  CHECK_NAPI(napi_create_bigint_uint64(env, (uint64_t)(uintptr_t)p, &out));
  return out;
}

static napi_value js_free(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  void* p;
  if (!ctx || !get_ptr(env, argv[1], &p)) return NULL;
  int rc = gsr_free(ctx, p);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* upload(ctx, ptr, typedArray) */
static napi_value js_upload(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  void *p, *data;
  size_t len;
  if (!ctx || !get_ptr(env, argv[1], &p) || !get_bytes(env, argv[2], &data, &len)) return NULL;
  int rc = gsr_upload(ctx, p, data, len);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* download(ctx, ptr, typedArray): fills the array (waits for the device) */
static napi_value js_download(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  void *p, *data;
  size_t len;
  if (!ctx || !get_ptr(env, argv[1], &p) || !get_bytes(env, argv[2], &data, &len)) return NULL;
  int rc = gsr_download(ctx, data, p, len);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

static napi_value js_clear(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  void* p;
  uint64_t bytes;
  if (!ctx || !get_ptr(env, argv[1], &p) || !get_u64(env, argv[2], &bytes)) return NULL;
  int rc = gsr_clear(ctx, p, (size_t)bytes);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* copy(ctx, dst, src, bytes) — copyBufferToBuffer gpu.js:719 */
static napi_value js_copy(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  void *dst, *src;
  uint64_t bytes;
  if (!ctx || !get_ptr(env, argv[1], &dst) || !get_ptr(env, argv[2], &src) || !get_u64(env, argv[3], &bytes))
    return NULL;
  int rc = gsr_copy(ctx, dst, src, (size_t)bytes);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  if (!ctx) return NULL;
  int rc = gsr_synchronize(ctx);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* project(ctx, cam, gs, sh, shDegree, n, splats) — PREPROCESS gpu.js:64-88 */
static napi_value js_project(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *gs, *sh, *splats;
  uint32_t deg, n;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &gs) ||
      !get_ptr(env, argv[3], &sh) || !get_u32(env, argv[4], &deg) || !get_u32(env, argv[5], &n) ||
      !get_ptr(env, argv[6], &splats))
    return NULL;
  int rc = gsr_project(ctx, &cam, (const float*)gs, (const float*)sh, (int)deg, n, (float*)splats);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* binSort(ctx, cam, splats, n, order, capacity, ranges) -> entries (Number).
 * Too small an order buffer throws with code GSR_E_CAPACITY and `entries`. */
static napi_value js_bin_sort(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *splats, *order, *ranges;
  uint32_t n;
  uint64_t cap;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &splats) ||
      !get_u32(env, argv[3], &n) || !get_ptr(env, argv[4], &order) || !get_u64(env, argv[5], &cap) ||
      !get_ptr(env, argv[6], &ranges))
    return NULL;
  uint64_t entries = 0;
  int rc = gsr_bin_sort(ctx, &cam, (const float*)splats, n, (uint32_t*)order, cap, (uint32_t*)ranges, &entries);
  if (rc == GSR_E_CAPACITY) {
    char msg[600];
    snprintf(msg, sizeof(msg), "gsr error %d: %s", rc, gsr_last_error(ctx));
    napi_value err, m, code, e;
    napi_create_string_utf8(env, msg, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &err);
    napi_create_int32(env, rc, &code);
    napi_set_named_property(env, err, "code", code);
    napi_create_double(env, (double)entries, &e);
    napi_set_named_property(env, err, "entries", e);
    napi_throw(env, err);
    return NULL;
  }
  if (rc) return throw_gsr(env, ctx, rc);
  napi_value out;
  CHECK_NAPI(napi_create_double(env, (double)entries, &out));
  return out;
}

/* forward(ctx, cam, splats, order, ranges, image, state|null) — FORWARD gpu.js:105-137 */
static napi_value js_forward(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *splats, *order, *ranges, *image, *state;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &splats) ||
      !get_ptr(env, argv[3], &order) || !get_ptr(env, argv[4], &ranges) || !get_ptr(env, argv[5], &image) ||
      !get_ptr(env, argv[6], &state))
    return NULL;
  int rc = gsr_forward(ctx, &cam, (const float*)splats, (const uint32_t*)order, (const uint32_t*)ranges,
                       (float*)image, (float*)state);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* lossGrad(ctx, cam, image, target, lambda, out) — SSIM_COEF + LOSS_GRAD gpu.js:336-409 */
static napi_value js_loss_grad(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *image, *target, *out;
  float lambda;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &image) ||
      !get_ptr(env, argv[3], &target) || !get_f32(env, argv[4], &lambda) || !get_ptr(env, argv[5], &out))
    return NULL;
  int rc = gsr_loss_grad(ctx, &cam, (const float*)image, (const float*)target, lambda, (float*)out);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* backward(ctx, cam, splats, order, ranges, image, lossGrad, grads) — BACKWARD_SG gpu.js:204-293 */
static napi_value js_backward(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  if (!get_args(env, info, 8, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *splats, *order, *ranges, *image, *lg, *grads;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &splats) ||
      !get_ptr(env, argv[3], &order) || !get_ptr(env, argv[4], &ranges) || !get_ptr(env, argv[5], &image) ||
      !get_ptr(env, argv[6], &lg) || !get_ptr(env, argv[7], &grads))
    return NULL;
  int rc = gsr_backward(ctx, &cam, (const float*)splats, (const uint32_t*)order, (const uint32_t*)ranges,
                        (const float*)image, (const float*)lg, (float*)grads);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* shape(ctx, cam, gs, sh, shDegree, n, grads, shape, shGrad) — SHAPE gpu.js:297-328 */
static napi_value js_shape(napi_env env, napi_callback_info info) {
  napi_value argv[9];
  if (!get_args(env, info, 9, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *gs, *sh, *grads, *shape, *shg;
  uint32_t deg, n;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &gs) ||
      !get_ptr(env, argv[3], &sh) || !get_u32(env, argv[4], &deg) || !get_u32(env, argv[5], &n) ||
      !get_ptr(env, argv[6], &grads) || !get_ptr(env, argv[7], &shape) || !get_ptr(env, argv[8], &shg))
    return NULL;
  int rc = gsr_shape(ctx, &cam, (const float*)gs, (const float*)sh, (int)deg, n, (const float*)grads,
                     (float*)shape, (float*)shg);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* adam(ctx, step, gs, grads, shape, moments) — ADAM gpu.js:411-443 */
static napi_value js_adam(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  if (!get_args(env, info, 6, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_step st;
  void *gs, *grads, *shape, *mom;
  if (!ctx || !get_struct(env, argv[1], &st, sizeof st) || !get_ptr(env, argv[2], &gs) ||
      !get_ptr(env, argv[3], &grads) || !get_ptr(env, argv[4], &shape) || !get_ptr(env, argv[5], &mom))
    return NULL;
  int rc = gsr_adam(ctx, &st, (float*)gs, (const float*)grads, (const float*)shape, (float*)mom);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* stats(ctx, step, grads, splats, stats) — STATS gpu.js:483-497 */
static napi_value js_stats(napi_env env, napi_callback_info info) {
  napi_value argv[5];
  if (!get_args(env, info, 5, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_step st;
  void *grads, *splats, *stats;
  if (!ctx || !get_struct(env, argv[1], &st, sizeof st) || !get_ptr(env, argv[2], &grads) ||
      !get_ptr(env, argv[3], &splats) || !get_ptr(env, argv[4], &stats))
    return NULL;
  int rc = gsr_stats(ctx, &st, (const float*)grads, (const float*)splats, (float*)stats);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* ---- asynchronous readbacks (read(), gpu.js:716-726) ---------------------- */

enum { OP_READ = 0, OP_BIN_SORT = 1 };

typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref ctx_ref; /* keeps the context alive while the worker uses it */
  addon_ctx* actx;
  int kind;
  /* read */
  const void* src;
  size_t bytes;
  void* host;
  /* binSortAsync */
  gsr_camera cam;
  const float* splats;
  uint32_t n;
  uint32_t* order;
  uint64_t cap;
  uint32_t* ranges;
  uint64_t entries;
  int rc;
  char msg[600];
} async_op;

/* Worker thread: the context is busy, so this thread is its only user. */
static void op_execute(napi_env env, void* data) {
  (void)env;
  async_op* op = (async_op*)data;
  gsr_ctx* ctx = op->actx->ctx;
  if (op->kind == OP_READ)
    op->rc = op->bytes ? gsr_download(ctx, op->host, op->src, op->bytes) : GSR_OK;
  else
    op->rc = gsr_bin_sort(ctx, &op->cam, op->splats, op->n, op->order, op->cap, op->ranges, &op->entries);
  if (op->rc) snprintf(op->msg, sizeof(op->msg), "gsr error %d: %s", op->rc, gsr_last_error(ctx));
}

static void free_host(napi_env env, void* data, void* hint) {
  (void)env;
  (void)hint;
  free(data);
}

/* JavaScript thread: settles the Promise and releases the context. */
static void op_complete(napi_env env, napi_status status, void* data) {
  async_op* op = (async_op*)data;
  op->actx->busy = 0;
  napi_value result = NULL;
  if (status != napi_ok || op->rc) {
    napi_value m, code;
    if (status != napi_ok) snprintf(op->msg, sizeof(op->msg), "asynchronous gsr call was not completed");
    napi_create_string_utf8(env, op->msg, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, NULL, m, &result);
    napi_create_int32(env, status != napi_ok ? GSR_E_INVALID : op->rc, &code);
    napi_set_named_property(env, result, "code", code);
    if (op->kind == OP_BIN_SORT && op->rc == GSR_E_CAPACITY) {
      napi_value e;
      napi_create_double(env, (double)op->entries, &e);
      napi_set_named_property(env, result, "entries", e);
    }
    napi_reject_deferred(env, op->deferred, result);
    free(op->host);
  } else if (op->kind == OP_READ) {
    /* the ArrayBuffer takes the host block (freed with it) */
    if (op->bytes) {
      if (napi_create_external_arraybuffer(env, op->host, op->bytes, free_host, NULL, &result) != napi_ok) {
        free(op->host);
        napi_create_arraybuffer(env, 0, NULL, &result);
      }
    } else {
      napi_create_arraybuffer(env, 0, NULL, &result);
    }
    napi_resolve_deferred(env, op->deferred, result);
  } else {
    napi_create_double(env, (double)op->entries, &result);
    napi_resolve_deferred(env, op->deferred, result);
  }
  napi_delete_async_work(env, op->work);
  napi_delete_reference(env, op->ctx_ref);
  free(op);
}

/* Queues op on a worker thread and returns its Promise; the context is busy
 * until the Promise settles. */
static napi_value start_op(napi_env env, napi_value ctx_value, async_op* op, const char* name) {
  napi_value promise, rname;
  if (napi_create_promise(env, &op->deferred, &promise) != napi_ok ||
      napi_create_reference(env, ctx_value, 1, &op->ctx_ref) != napi_ok ||
      napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &rname) != napi_ok ||
      napi_create_async_work(env, NULL, rname, op_execute, op_complete, op, &op->work) != napi_ok) {
    free(op->host);
    free(op);
    napi_throw_error(env, NULL, "N-API call failed while starting an asynchronous gsr call");
    return NULL;
  }
  op->actx->busy = 1;
  if (napi_queue_async_work(env, op->work) != napi_ok) {
    op->actx->busy = 0;
    napi_delete_async_work(env, op->work);
    napi_delete_reference(env, op->ctx_ref);
    free(op->host);
    free(op);
    napi_throw_error(env, NULL, "N-API call failed: napi_queue_async_work");
    return NULL;
  }
  return promise;
}

/* read(ctx, ptr, bytes) -> Promise<ArrayBuffer>: a copy of device memory after
 * all work queued so far (read(), gpu.js:716-726). */
static napi_value js_read(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  addon_ctx* a = get_actx(env, argv[0]);
  void* p;
  uint64_t bytes;
  if (!a || !get_ptr(env, argv[1], &p) || !get_u64(env, argv[2], &bytes)) return NULL;
  async_op* op = (async_op*)calloc(1, sizeof(async_op));
  if (!op) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  op->actx = a;
  op->kind = OP_READ;
  op->src = p;
  op->bytes = (size_t)bytes;
  if (bytes && !(op->host = malloc((size_t)bytes))) {
    free(op);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  return start_op(env, argv[0], op, "gsr.read");
}

/* binSortAsync(ctx, cam, splats, n, order, capacity, ranges) -> Promise<entries>;
 * too small an order buffer rejects with code GSR_E_CAPACITY and `entries`. */
static napi_value js_bin_sort_async(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  addon_ctx* a = get_actx(env, argv[0]);
  async_op* op = (async_op*)calloc(1, sizeof(async_op));
  if (!op) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  void *splats, *order, *ranges;
  if (!a || !get_struct(env, argv[1], &op->cam, sizeof op->cam) || !get_ptr(env, argv[2], &splats) ||
      !get_u32(env, argv[3], &op->n) || !get_ptr(env, argv[4], &order) || !get_u64(env, argv[5], &op->cap) ||
      !get_ptr(env, argv[6], &ranges)) {
    free(op);
    return NULL;
  }
  op->actx = a;
  op->kind = OP_BIN_SORT;
  op->splats = (const float*)splats;
  op->order = (uint32_t*)order;
  op->ranges = (uint32_t*)ranges;
  return start_op(env, argv[0], op, "gsr.binSort");
}

/* setCulling(ctx, enabled) — verification knob of forward / backward */
static napi_value js_set_culling(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  bool on;
  if (!ctx) return NULL;
  if (napi_get_value_bool(env, argv[1], &on) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a boolean");
    return NULL;
  }
  int rc = gsr_set_culling(ctx, on ? 1 : 0);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* lossGradWindow(ctx, cam, image, target, lambda, window, out) — 7 (box) or 11 (Gaussian) */
static napi_value js_loss_grad_window(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  if (!get_args(env, info, 7, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_camera cam;
  void *image, *target, *out;
  float lambda;
  uint32_t window;
  if (!ctx || !get_struct(env, argv[1], &cam, sizeof cam) || !get_ptr(env, argv[2], &image) ||
      !get_ptr(env, argv[3], &target) || !get_f32(env, argv[4], &lambda) || !get_u32(env, argv[5], &window) ||
      !get_ptr(env, argv[6], &out))
    return NULL;
  int rc = gsr_loss_grad_window(ctx, &cam, (const float*)image, (const float*)target, lambda, (int)window,
                                (float*)out);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* adamSh(ctx, step, lrShRest, gs, sh, grads, shape, shGrad, moments) — ADAM over SH3 */
static napi_value js_adam_sh(napi_env env, napi_callback_info info) {
  napi_value argv[9];
  if (!get_args(env, info, 9, argv)) return NULL;
  gsr_ctx* ctx = get_ctx(env, argv[0]);
  gsr_step st;
  float lr_rest;
  void *gs, *sh, *grads, *shape, *shg, *mom;
  if (!ctx || !get_struct(env, argv[1], &st, sizeof st) || !get_f32(env, argv[2], &lr_rest) ||
      !get_ptr(env, argv[3], &gs) || !get_ptr(env, argv[4], &sh) || !get_ptr(env, argv[5], &grads) ||
      !get_ptr(env, argv[6], &shape) || !get_ptr(env, argv[7], &shg) || !get_ptr(env, argv[8], &mom))
    return NULL;
  int rc = gsr_adam_sh(ctx, &st, lr_rest, (float*)gs, (float*)sh, (const float*)grads, (const float*)shape,
                       (const float*)shg, (float*)mom);
  if (rc) return throw_gsr(env, ctx, rc);
  return undefined(env);
}

/* ---- batched views --------------------------------------------------------- */

typedef struct {
  gsr_batch* batch;
  addon_ctx* actx;
  napi_ref ctx_ref; /* the context outlives its batches */
  uint32_t views;
} addon_batch;

static void batch_finalize(napi_env env, void* data, void* hint) {
  (void)hint;
  addon_batch* b = (addon_batch*)data;
  addon_ctx* a = b->actx;
  if (b->batch) {
    gsr_batch_destroy(b->batch);
    a->batches--;
  }
  napi_delete_reference(env, b->ctx_ref);
  free(b);
  ctx_release(a);
}

static addon_batch* get_batch(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a gsr batch");
    return NULL;
  }
  addon_batch* b = (addon_batch*)p;
  if (!b->batch) {
    napi_throw_error(env, NULL, "gsr batch destroyed");
    return NULL;
  }
  if (b->actx->busy) {
    napi_throw_error(env, NULL, "gsr context busy: await its pending read() or binSortAsync() first");
    return NULL;
  }
  return b;
}

static napi_value throw_batch(napi_env env, addon_batch* b, int rc) { return throw_gsr(env, b->actx->ctx, rc); }

/* Cameras: one ArrayBuffer / typed array of views * 80 bytes (cameraUniform each). */
static gsr_camera* get_cams(napi_env env, napi_value v, uint32_t* count) {
  void* data;
  size_t len;
  if (!get_bytes(env, v, &data, &len)) return NULL;
  if (len == 0 || len % sizeof(gsr_camera)) {
    napi_throw_range_error(env, NULL, "cameras: a multiple of 80 bytes expected");
    return NULL;
  }
  *count = (uint32_t)(len / sizeof(gsr_camera));
  return (gsr_camera*)data;
}

/* batchCreate(ctx, cams, n, shDegree) -> batch */
static napi_value js_batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  addon_ctx* a = get_actx(env, argv[0]);
  uint32_t views = 0, n, deg;
  gsr_camera* cams = a ? get_cams(env, argv[1], &views) : NULL;
  if (!cams || !get_u32(env, argv[2], &n) || !get_u32(env, argv[3], &deg)) return NULL;
  gsr_batch* batch = NULL;
  int rc = gsr_batch_create(a->ctx, views, cams, n, (int)deg, &batch);
  if (rc) return throw_gsr(env, a->ctx, rc);
  addon_batch* b = (addon_batch*)calloc(1, sizeof(addon_batch));
  napi_value ext;
  if (!b || napi_create_reference(env, argv[0], 1, &b->ctx_ref) != napi_ok) {
    gsr_batch_destroy(batch);
    free(b);
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  b->batch = batch;
  b->actx = a;
  b->views = views;
  a->batches++;
  if (napi_create_external(env, b, batch_finalize, NULL, &ext) != napi_ok) {
    napi_delete_reference(env, b->ctx_ref);
    gsr_batch_destroy(batch);
    a->batches--;
    free(b);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  return ext;
}

/* batchDestroy(batch): releases the device buffers now (else at collection) */
static napi_value js_batch_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  if (!b) return NULL;
  gsr_batch_destroy(b->batch);
  b->batch = NULL;
  b->actx->batches--;
  return undefined(env);
}

static napi_value js_batch_set_cameras(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  uint32_t views = 0;
  gsr_camera* cams = b ? get_cams(env, argv[1], &views) : NULL;
  if (!cams) return NULL;
  if (views != b->views) {
    napi_throw_range_error(env, NULL, "camera count differs from the batch's");
    return NULL;
  }
  int rc = gsr_batch_set_cameras(b->batch, cams);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

/* batchStep(batch, gs, sh, targets, lambda, grads, shape, shGrad, stats, images, useGraph) */
static napi_value js_batch_step(napi_env env, napi_callback_info info) {
  napi_value argv[11];
  if (!get_args(env, info, 11, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  void *gs, *sh, *tg, *grads, *shape, *shg, *stats, *images;
  float lambda;
  bool graph;
  if (!b || !get_ptr(env, argv[1], &gs) || !get_ptr(env, argv[2], &sh) || !get_ptr(env, argv[3], &tg) ||
      !get_f32(env, argv[4], &lambda) || !get_ptr(env, argv[5], &grads) || !get_ptr(env, argv[6], &shape) ||
      !get_ptr(env, argv[7], &shg) || !get_ptr(env, argv[8], &stats) || !get_ptr(env, argv[9], &images))
    return NULL;
  if (napi_get_value_bool(env, argv[10], &graph) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a boolean");
    return NULL;
  }
  int rc = gsr_batch_step_stats(b->batch, (const float*)gs, (const float*)sh, (const float*)tg, lambda,
                                (float*)grads, (float*)shape, (float*)shg, (float*)stats, (float*)images,
                                graph ? 1 : 0);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

/* batchEntries(batch) -> Float64Array(views): tile-list entries per view of the last step */
static napi_value js_batch_entries(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  if (!b) return NULL;
  uint64_t* e = (uint64_t*)calloc(b->views, sizeof(uint64_t));
  if (!e) {
    napi_throw_error(env, NULL, "out of host memory");
    return NULL;
  }
  int rc = gsr_batch_entries(b->batch, e);
  if (rc) {
    free(e);
    return throw_batch(env, b, rc);
  }
  napi_value ab, arr;
  double* d;
  if (napi_create_arraybuffer(env, b->views * sizeof(double), (void**)&d, &ab) != napi_ok) {
    free(e);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_arraybuffer");
    return NULL;
  }
  for (uint32_t v = 0; v < b->views; ++v) d[v] = (double)e[v];
  free(e);
  CHECK_NAPI(napi_create_typedarray(env, napi_float64_array, b->views, ab, 0, &arr));
  return arr;
}

/* batchAutotune(batch, gs, sh, targets, lambda, grads, shape, shGrad) -> tuning (Uint32Array(8)) */
static napi_value js_batch_tuning_array(napi_env env, addon_batch* b) {
  uint32_t t[8];
  int rc = gsr_batch_tuning(b->batch, t);
  if (!rc) rc = gsr_batch_passes(b->batch, &t[6], &t[7]);
  if (rc) return throw_batch(env, b, rc);
  napi_value ab, arr;
  uint32_t* d;
  CHECK_NAPI(napi_create_arraybuffer(env, sizeof t, (void**)&d, &ab));
  memcpy(d, t, sizeof t);
  CHECK_NAPI(napi_create_typedarray(env, napi_uint32_array, 8, ab, 0, &arr));
  return arr;
}

static napi_value js_batch_autotune(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  if (!get_args(env, info, 8, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  void *gs, *sh, *tg, *grads, *shape, *shg;
  float lambda;
  if (!b || !get_ptr(env, argv[1], &gs) || !get_ptr(env, argv[2], &sh) || !get_ptr(env, argv[3], &tg) ||
      !get_f32(env, argv[4], &lambda) || !get_ptr(env, argv[5], &grads) || !get_ptr(env, argv[6], &shape) ||
      !get_ptr(env, argv[7], &shg))
    return NULL;
  int rc = gsr_batch_autotune(b->batch, (const float*)gs, (const float*)sh, (const float*)tg, lambda,
                              (float*)grads, (float*)shape, (float*)shg, NULL, NULL);
  if (rc) return throw_batch(env, b, rc);
  return js_batch_tuning_array(env, b);
}

/* batchTuning(batch) -> Uint32Array(8): tile batch, chunk, split first/later/backward, emit band,
 * depth-slice passes and growth */
static napi_value js_batch_tuning(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  if (!b) return NULL;
  return js_batch_tuning_array(env, b);
}

/* batchSetTuning(batch, tuning[6 or 8]): the knobs batchTuning returned (the depth slices optional) */
static napi_value js_batch_set_tuning(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  void* data;
  size_t len;
  if (!b || !get_bytes(env, argv[1], &data, &len)) return NULL;
  if (len < 6 * sizeof(uint32_t)) {
    napi_throw_range_error(env, NULL, "tuning: 6 unsigned integers expected");
    return NULL;
  }
  const uint32_t* t = (const uint32_t*)data;
  int rc = len >= 8 * sizeof(uint32_t) ? gsr_batch_set_passes(b->batch, t[6], t[7]) : GSR_OK;
  if (!rc) rc = gsr_batch_configure(b->batch, t[0], t[1]);
  if (!rc) rc = gsr_batch_set_split(b->batch, t[2], t[3], t[4]);
  if (!rc) rc = gsr_batch_set_emit_band(b->batch, t[5]);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

/* batchSetPasses(batch, passes, growth), batchSetSsimWindow(batch, window), batchSetCulling(batch, on) */
static napi_value js_batch_set_passes(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  uint32_t passes, growth;
  if (!b || !get_u32(env, argv[1], &passes) || !get_u32(env, argv[2], &growth)) return NULL;
  int rc = gsr_batch_set_passes(b->batch, passes, growth);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

static napi_value js_batch_set_ssim_window(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  uint32_t window;
  if (!b || !get_u32(env, argv[1], &window)) return NULL;
  int rc = gsr_batch_set_ssim_window(b->batch, (int)window);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

static napi_value js_batch_set_culling(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  addon_batch* b = get_batch(env, argv[0]);
  bool on;
  if (!b) return NULL;
  if (napi_get_value_bool(env, argv[1], &on) != napi_ok) {
    napi_throw_type_error(env, NULL, "expected a boolean");
    return NULL;
  }
  int rc = gsr_batch_set_culling(b->batch, on ? 1 : 0);
  if (rc) return throw_batch(env, b, rc);
  return undefined(env);
}

static napi_value init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback fn;
  } fns[] = {
      {"create", js_create},     {"alloc", js_alloc},       {"free", js_free},
      {"upload", js_upload},     {"download", js_download}, {"clear", js_clear},
      {"copy", js_copy},
      {"synchronize", js_sync},  {"project", js_project},   {"binSort", js_bin_sort},
      {"forward", js_forward},   {"lossGrad", js_loss_grad}, {"backward", js_backward},
      {"shape", js_shape},       {"adam", js_adam},         {"stats", js_stats},
      {"read", js_read},         {"binSortAsync", js_bin_sort_async},
      {"setCulling", js_set_culling}, {"lossGradWindow", js_loss_grad_window}, {"adamSh", js_adam_sh},
      {"batchCreate", js_batch_create}, {"batchDestroy", js_batch_destroy},
      {"batchSetCameras", js_batch_set_cameras}, {"batchStep", js_batch_step},
      {"batchEntries", js_batch_entries}, {"batchAutotune", js_batch_autotune},
      {"batchTuning", js_batch_tuning}, {"batchSetTuning", js_batch_set_tuning},
      {"batchSetPasses", js_batch_set_passes}, {"batchSetSsimWindow", js_batch_set_ssim_window},
      {"batchSetCulling", js_batch_set_culling},
  };
  for (size_t i = 0; i < sizeof(fns) / sizeof(fns[0]); ++i) {
    napi_value f;
    if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f) != napi_ok ||
        napi_set_named_property(env, exports, fns[i].name, f) != napi_ok)
      return NULL;
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
