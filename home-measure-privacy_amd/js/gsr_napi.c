/*
 * N-API addon over libgsr (include/gsr.h): the binding a JavaScript trainer
 * loads in place of WebGPU (trainer/gpu.js:704-712).  One function per C entry
 * point; device pointers travel as BigInt, uniforms (camera 80 B, step 32 B) as
 * ArrayBuffers or typed arrays laid out as cameraUniform / the Step struct
 * (gpu.js:696-701, 866-869).  A non-zero status throws an Error carrying
 * gsr_last_error() and a numeric `code` (as the reference throws from
 * createRenderer, gpu.js:705).  Calls run on the JavaScript thread; downloads
 * wait for the device.
 */
#define NAPI_VERSION 8
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
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

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  if (argc < want) {
    napi_throw_type_error(env, NULL, "too few arguments");
    return 0;
  }
  return 1;
}

static gsr_ctx* get_ctx(napi_env env, napi_value v) {
  void* p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, NULL, "expected a gsr context");
    return NULL;
  }
  return (gsr_ctx*)p;
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
  gsr_destroy((gsr_ctx*)data);
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
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, ctx, ctx_finalize, NULL, &ext));
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
