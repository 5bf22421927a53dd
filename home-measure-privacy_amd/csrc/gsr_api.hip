// C ABI of libgsr (include/gsr.h): context, buffers, the single-view operators
// that mirror the pipelines of trainer/gpu.js (createRenderer, gpu.js:704-980)
// and the batched multi-view step.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "gsr_internal.h"

using namespace gsr;

struct gsr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int num_cus = 0;
  // Buffers reused between calls, grown by 1.25x (as the reference's scratch
  // pool, gpu.js:764-772).
  std::map<std::string, std::pair<void*, size_t>> pool;
  // Single-view work list, rebuilt when the resolution changes.
  uint32_t work_w = 0, work_h = 0;
  // Pinned staging for small device->host reads.
  uint64_t* pinned = nullptr;
  // Verification knob of gsr_forward / gsr_backward: no hit-box culling.
  int no_cull = 0;
  // Device address of the static scratch on this context's device (null when
  // the lookup failed; scratch_error says why).
  void* scratch = nullptr;
  const char* scratch_error = "";
};

namespace gsr {
hipError_t ctx_malloc(gsr_ctx* ctx, void** p, size_t bytes);
}  // namespace gsr

// Scratch of gsr_bin_sort_noalloc: a file-scope device array, so that call
// allocates nothing.  Every GPU the library runs on holds its own copy until
// the process ends; calls on one GPU share it (and race if they overlap).
constexpr size_t kScratchBytes = size_t(48) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

thread_local std::string g_create_error;

// The scratch's device address per device, looked up once per device with that
// device current (host code must not use the array's name as an address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

void lookup_scratch(gsr_ctx* ctx) {
  if (ctx->device >= kMaxDevices) {
    ctx->scratch_error = "device index above the scratch table";
    return;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_done[ctx->device]) {
    void* p = nullptr;
    g_scratch_err[ctx->device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[ctx->device] = g_scratch_err[ctx->device] == hipSuccess ? p : nullptr;
    g_scratch_done[ctx->device] = true;
  }
  ctx->scratch = g_scratch_ptr[ctx->device];
  if (!ctx->scratch) ctx->scratch_error = hipGetErrorString(g_scratch_err[ctx->device]);
}

int fail(gsr_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  else g_create_error = buf;
  return code;
}

#define GSR_HIP(ctx, call)                                                                        \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail((ctx), GSR_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define GSR_LAUNCHED(ctx)                                                                              \
  do {                                                                                                 \
    hipError_t e_ = hipGetLastError();                                                                 \
    if (e_ != hipSuccess) return fail((ctx), GSR_E_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

// A pooled device buffer of at least `bytes` (contents undefined).
int pooled(gsr_ctx* ctx, const char* name, size_t bytes, void** out) {
  auto it = ctx->pool.find(name);
  if (it != ctx->pool.end() && it->second.second >= bytes) {
    *out = it->second.first;
    return GSR_OK;
  }
  if (it != ctx->pool.end()) {
    GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GSR_HIP(ctx, hipFree(it->second.first));
    ctx->pool.erase(it);
  }
  size_t sz = std::max<size_t>(256, ((bytes + bytes / 4) + 255) / 256 * 256);
  void* p = nullptr;
  if (gsr::ctx_malloc(ctx, &p, sz) != hipSuccess)
    return fail(ctx, GSR_E_NOMEM, "out of device memory (%zu bytes for %s)", sz, name);
  ctx->pool[name] = {p, sz};
  *out = p;
  return GSR_OK;
}

template <class T>
int pooled_t(gsr_ctx* ctx, const char* name, size_t count, T** out) {
  void* p = nullptr;
  int rc = pooled(ctx, name, count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return rc;
}

ViewDesc view_of(const gsr_camera& c, uint32_t tile_base, uint64_t pixel_base, uint64_t splat_base) {
  ViewDesc v{};
  v.width = uint32_t(c.width);
  v.height = uint32_t(c.height);
  v.tiles_x = (v.width + kTile - 1) / kTile;
  v.tiles_y = (v.height + kTile - 1) / kTile;
  v.tile_base = tile_base;
  v.num_tiles = v.tiles_x * v.tiles_y;
  v.pixel_base = pixel_base;
  v.splat_base = splat_base;
  return v;
}

int check_camera(gsr_ctx* ctx, const gsr_camera* cam) {
  if (!cam) return fail(ctx, GSR_E_INVALID, "camera is NULL");
  if (!(cam->width >= 1.0f && cam->height >= 1.0f && cam->width <= 65536.0f && cam->height <= 65536.0f) ||
      cam->width != float(uint32_t(cam->width)) || cam->height != float(uint32_t(cam->height)))
    return fail(ctx, GSR_E_INVALID, "camera size %gx%g is not a positive integer size", cam->width, cam->height);
  return GSR_OK;
}

// Uploads the single-view descriptor and camera; rebuilds the work list when the
// resolution changed.
int single_view(gsr_ctx* ctx, const gsr_camera* cam, ViewDesc** dview, gsr_camera** dcam, TileWork** work,
                ViewDesc* host) {
  int rc = check_camera(ctx, cam);
  if (rc) return rc;
  *host = view_of(*cam, 0, 0, 0);
  if ((rc = pooled_t(ctx, "view", 1, dview))) return rc;
  if ((rc = pooled_t(ctx, "cam", 1, dcam))) return rc;
  GSR_HIP(ctx, hipMemcpyAsync(*dview, host, sizeof(ViewDesc), hipMemcpyHostToDevice, ctx->stream));
  GSR_HIP(ctx, hipMemcpyAsync(*dcam, cam, sizeof(gsr_camera), hipMemcpyHostToDevice, ctx->stream));
  if (work) {
    if ((rc = pooled_t(ctx, "work", host->num_tiles, work))) return rc;
    if (ctx->work_w != host->width || ctx->work_h != host->height) {
      launch_fill_work(ctx->stream, *dview, 1, host->num_tiles, *work);
      GSR_LAUNCHED(ctx);
      ctx->work_w = host->width;
      ctx->work_h = host->height;
    }
  }
  return GSR_OK;
}

uint32_t pick_chunk(uint64_t items) {
  // Aim for about 1024 binning blocks; chunks between 256 and 4096 Gaussians.
  uint64_t c = 256;
  while (c < 4096 && c * 1024 < items) c <<= 1;
  return uint32_t(c);
}

}  // namespace

extern "C" {

int gsr_create(int device, gsr_ctx** out) {
  if (!out) return fail(nullptr, GSR_E_INVALID, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(nullptr, GSR_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0 || device >= count) return fail(nullptr, GSR_E_INVALID, "device %d out of range [0, %d)", device, count);
  if ((e = hipSetDevice(device)) != hipSuccess) return fail(nullptr, GSR_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  gsr_ctx* ctx = new gsr_ctx();
  ctx->device = device;
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
    delete ctx;
    return fail(nullptr, GSR_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  if (ctx->num_cus <= 0) ctx->num_cus = 256;
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&ctx->pinned), 4096, hipHostMallocDefault)) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return fail(nullptr, GSR_E_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  lookup_scratch(ctx);  // device `device` is current
  *out = ctx;
  return GSR_OK;
}

void gsr_destroy(gsr_ctx* ctx) {
  if (!ctx) return;
  // Teardown: nothing to report to a caller that is destroying the context.
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->pool) (void)hipFree(kv.second.first);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* gsr_last_error(const gsr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int gsr_alloc(gsr_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(ctx, GSR_E_INVALID, "NULL argument");
  const size_t sz = (std::max<size_t>(16, bytes) + 15) / 16 * 16;
  void* p = nullptr;
  if (gsr::ctx_malloc(ctx, &p, sz) != hipSuccess) return fail(ctx, GSR_E_NOMEM, "out of device memory (%zu bytes)", sz);
  launch_zero(ctx->stream, p, sz);
  GSR_LAUNCHED(ctx);
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *dptr = p;
  return GSR_OK;
}

int gsr_free(gsr_ctx* ctx, void* dptr) {
  if (!ctx) return GSR_E_INVALID;
  if (!dptr) return GSR_OK;
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  GSR_HIP(ctx, hipFree(dptr));
  return GSR_OK;
}

int gsr_upload(gsr_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (bytes == 0) return GSR_OK;
  GSR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GSR_OK;
}

int gsr_download(gsr_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (bytes == 0) return GSR_OK;
  GSR_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GSR_OK;
}

int gsr_clear(gsr_ctx* ctx, void* dptr, size_t bytes) {
  if (!ctx || (!dptr && bytes)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (bytes % 4 || reinterpret_cast<uintptr_t>(dptr) % 4)
    return fail(ctx, GSR_E_INVALID, "clear of %zu bytes at %p is not 4-byte aligned", bytes, dptr);
  launch_zero(ctx->stream, dptr, bytes);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_copy(gsr_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || ((!dst || !src) && bytes)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (bytes % 4 || reinterpret_cast<uintptr_t>(dst) % 4 || reinterpret_cast<uintptr_t>(src) % 4)
    return fail(ctx, GSR_E_INVALID, "copy of %zu bytes is not 4-byte aligned", bytes);
  const uintptr_t d = reinterpret_cast<uintptr_t>(dst), s = reinterpret_cast<uintptr_t>(src);
  if (bytes && d < s + bytes && s < d + bytes) return fail(ctx, GSR_E_INVALID, "copy regions overlap");
  launch_copy(ctx->stream, dst, src, bytes);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_set_culling(gsr_ctx* ctx, int enabled) {
  if (!ctx) return GSR_E_INVALID;
  ctx->no_cull = enabled ? 0 : 1;
  return GSR_OK;
}

int gsr_synchronize(gsr_ctx* ctx) {
  if (!ctx) return GSR_E_INVALID;
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GSR_OK;
}

int gsr_project(gsr_ctx* ctx, const gsr_camera* cam, const float* gs, const float* sh, int sh_degree, uint32_t n,
                float* splats) {
  if (!ctx) return GSR_E_INVALID;
  if (n && (!gs || !splats)) return fail(ctx, GSR_E_INVALID, "gs / splats is NULL");
  if (sh_degree < 0 || sh_degree > 3) return fail(ctx, GSR_E_INVALID, "sh_degree %d not in [0, 3]", sh_degree);
  if (sh_degree > 0 && !sh) return fail(ctx, GSR_E_INVALID, "sh is NULL with sh_degree %d", sh_degree);
  ViewDesc* dview;
  gsr_camera* dcam;
  ViewDesc host;
  int rc = single_view(ctx, cam, &dview, &dcam, nullptr, &host);
  if (rc) return rc;
  launch_project(ctx->stream, dcam, dview, 1, gs, sh, sh_degree, n, splats, nullptr, nullptr, nullptr);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

}  // extern "C"

namespace {

// Device scratch of one gsr_bin_sort call.
struct BinSortScratch {
  uint32_t *keys, *keys_alt, *vals, *vals_alt, *hist, *piece, *matrix, *tile_tot, *overflow;
  uint2 *rects, *rect_sorted;
  uint64_t *ebase, *entries, *run;
};

// Element counts of each scratch region for n Gaussians binned on `host`.
struct BinSortSizes {
  size_t nn, hist, piece, matrix, tiles;
  uint32_t chunk;
};

BinSortSizes bin_sort_sizes(const ViewDesc& host, uint32_t n) {
  BinSortSizes z;
  z.nn = std::max<uint32_t>(n, 1);
  z.chunk = pick_chunk(n);
  z.hist = std::max<size_t>(1, radix_hist_elems(n, 1));
  z.piece = std::max<size_t>(1, scan_piece_elems(z.hist, 1));
  z.matrix = std::max<size_t>(1, bin_matrix_elems(&host, 1, n, z.chunk));
  z.tiles = std::max<uint32_t>(host.num_tiles, 1);
  return z;
}

// tileLists (gpu.js:450-481) over the scratch `w`: prep, depth sort, counting
// binning, the entry count back to the host, then the order when it fits.
int bin_sort_run(gsr_ctx* ctx, const ViewDesc& host, const ViewDesc* dview, const float* splats, uint32_t n,
                 uint32_t* order, uint64_t order_capacity, uint32_t* ranges, uint64_t* num_entries,
                 const BinSortScratch& w, uint32_t chunk) {
  launch_zero(ctx->stream, w.overflow, sizeof(uint32_t));
  const bool packed = rect_packed(std::max(host.tiles_x, host.tiles_y));
  launch_prep(ctx->stream, dview, 1, splats, n, w.keys, w.rects, packed);
  SortScratch sc{w.keys_alt, w.vals, w.vals_alt, w.hist, w.piece};
  launch_depth_sort(ctx->stream, w.keys, n, 1, sc, w.rects, w.rect_sorted, packed);
  BinScratch bs{w.matrix, w.tile_tot, w.rect_sorted, w.run, packed};
  BinPass pass{0, n, nullptr, ranges, w.ebase, w.entries, true};
  launch_bin_count(ctx->stream, dview, &host, 1, n, chunk, bs, pass);
  GSR_LAUNCHED(ctx);
  GSR_HIP(ctx, hipMemcpyAsync(ctx->pinned, w.entries, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  GSR_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const uint64_t E = ctx->pinned[0];
  *num_entries = E;
  if (E > order_capacity)
    return fail(ctx, GSR_E_CAPACITY, "order holds %llu entries, %llu needed", (unsigned long long)order_capacity,
                (unsigned long long)E);
  launch_bin_emit(ctx->stream, dview, &host, 1, n, chunk, w.vals, bs, pass, order, order_capacity, w.overflow);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int check_bin_sort_args(gsr_ctx* ctx, const float* splats, uint32_t n, uint32_t* order, uint64_t order_capacity,
                        uint32_t* ranges, uint64_t* num_entries) {
  if (!ranges || !num_entries || (n && !splats)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (order_capacity && !order)
    return fail(ctx, GSR_E_INVALID, "order is NULL with capacity %llu", (unsigned long long)order_capacity);
  return GSR_OK;
}

}  // namespace

extern "C" {

int gsr_bin_sort(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, uint32_t n, uint32_t* order,
                 uint64_t order_capacity, uint32_t* ranges, uint64_t* num_entries) {
  if (!ctx) return GSR_E_INVALID;
  int rc = check_bin_sort_args(ctx, splats, n, order, order_capacity, ranges, num_entries);
  if (rc) return rc;
  ViewDesc* dview;
  gsr_camera* dcam;
  ViewDesc host;
  if ((rc = single_view(ctx, cam, &dview, &dcam, nullptr, &host))) return rc;
  const BinSortSizes z = bin_sort_sizes(host, n);
  BinSortScratch w;
  if ((rc = pooled_t(ctx, "keys", z.nn, &w.keys)) || (rc = pooled_t(ctx, "keys_alt", z.nn, &w.keys_alt)) ||
      (rc = pooled_t(ctx, "vals", z.nn, &w.vals)) || (rc = pooled_t(ctx, "vals_alt", z.nn, &w.vals_alt)) ||
      (rc = pooled_t(ctx, "rects", z.nn, &w.rects)) || (rc = pooled_t(ctx, "rect_sorted", z.nn, &w.rect_sorted)) ||
      (rc = pooled_t(ctx, "hist", z.hist, &w.hist)) || (rc = pooled_t(ctx, "piece", z.piece, &w.piece)) ||
      (rc = pooled_t(ctx, "matrix", z.matrix, &w.matrix)) || (rc = pooled_t(ctx, "tile_tot", z.tiles, &w.tile_tot)) ||
      (rc = pooled_t(ctx, "entries", 1, &w.entries)) || (rc = pooled_t(ctx, "ebase", 1, &w.ebase)) ||
      (rc = pooled_t(ctx, "run", 1, &w.run)) || (rc = pooled_t(ctx, "overflow", 1, &w.overflow)))
    return rc;
  return bin_sort_run(ctx, host, dview, splats, n, order, order_capacity, ranges, num_entries, w, z.chunk);
}

int gsr_bin_sort_noalloc(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, uint32_t n, uint32_t* order,
                         uint64_t order_capacity, uint32_t* ranges, uint64_t* num_entries) {
  if (!ctx) return GSR_E_INVALID;
  int rc = check_bin_sort_args(ctx, splats, n, order, order_capacity, ranges, num_entries);
  if (rc) return rc;
  if ((rc = check_camera(ctx, cam))) return rc;
  if (!ctx->scratch) return fail(ctx, GSR_E_HIP, "the static scratch is unavailable: %s", ctx->scratch_error);
  const ViewDesc host = view_of(*cam, 0, 0, 0);
  const BinSortSizes z = bin_sort_sizes(host, n);
  // Regions of the device's static scratch, 256-byte aligned, in this order:
  // offsets first, so the total is checked before any address is formed.
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t off = at;
    at += (bytes + 255) / 256 * 256;
    return off;
  };
  const size_t o_view = take(sizeof(ViewDesc)), o_keys = take(z.nn * 4), o_keys_alt = take(z.nn * 4),
               o_vals = take(z.nn * 4), o_vals_alt = take(z.nn * 4), o_rects = take(z.nn * 8),
               o_rect_sorted = take(z.nn * 8), o_hist = take(z.hist * 4), o_piece = take(z.piece * 4),
               o_matrix = take(z.matrix * 4), o_tile_tot = take(z.tiles * 4), o_entries = take(8),
               o_ebase = take(8), o_run = take(8), o_overflow = take(4);
  if (at > kScratchBytes)
    return fail(ctx, GSR_E_CAPACITY,
                "gsr_bin_sort_noalloc: %u Gaussians on a %ux%u view need %zu bytes of scratch, above its %zu",
                n, host.width, host.height, at, kScratchBytes);
  char* base = static_cast<char*>(ctx->scratch);
  ViewDesc* dview = reinterpret_cast<ViewDesc*>(base + o_view);
  BinSortScratch w;
  w.keys = reinterpret_cast<uint32_t*>(base + o_keys);
  w.keys_alt = reinterpret_cast<uint32_t*>(base + o_keys_alt);
  w.vals = reinterpret_cast<uint32_t*>(base + o_vals);
  w.vals_alt = reinterpret_cast<uint32_t*>(base + o_vals_alt);
  w.rects = reinterpret_cast<uint2*>(base + o_rects);
  w.rect_sorted = reinterpret_cast<uint2*>(base + o_rect_sorted);
  w.hist = reinterpret_cast<uint32_t*>(base + o_hist);
  w.piece = reinterpret_cast<uint32_t*>(base + o_piece);
  w.matrix = reinterpret_cast<uint32_t*>(base + o_matrix);
  w.tile_tot = reinterpret_cast<uint32_t*>(base + o_tile_tot);
  w.entries = reinterpret_cast<uint64_t*>(base + o_entries);
  w.ebase = reinterpret_cast<uint64_t*>(base + o_ebase);
  w.run = reinterpret_cast<uint64_t*>(base + o_run);
  w.overflow = reinterpret_cast<uint32_t*>(base + o_overflow);
  GSR_HIP(ctx, hipMemcpyAsync(dview, &host, sizeof(ViewDesc), hipMemcpyHostToDevice, ctx->stream));
  return bin_sort_run(ctx, host, dview, splats, n, order, order_capacity, ranges, num_entries, w, z.chunk);
}

static int raster_common(gsr_ctx* ctx, const gsr_camera* cam, RasterArgs& a, ViewDesc& host) {
  ViewDesc* dview;
  gsr_camera* dcam;
  TileWork* work;
  int rc = single_view(ctx, cam, &dview, &dcam, &work, &host);
  if (rc) return rc;
  a.views = dview;
  a.work = work;
  a.num_work = host.num_tiles;
  a.grads_view_stride = 0;
  a.queue = nullptr;
  a.tile_batch = 1;
  a.order_capacity = ~uint64_t(0);  // the caller's order holds every entry its ranges name
  a.num_passes = 1;
  a.ebase[0] = nullptr;
  a.pass = 0;
  a.tstate = nullptr;
  a.done = nullptr;
  a.no_cull = ctx->no_cull;
  return GSR_OK;
}

int gsr_forward(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, const uint32_t* order,
                const uint32_t* ranges, float* image, float* state) {
  if (!ctx) return GSR_E_INVALID;
  if (!splats || !order || !ranges || !image) return fail(ctx, GSR_E_INVALID, "NULL argument");
  RasterArgs a{};
  ViewDesc host;
  int rc = raster_common(ctx, cam, a, host);
  if (rc) return rc;
  a.splats = splats;
  a.order = order;
  a.ranges[0] = ranges;
  a.image = image;
  a.state = state;
  launch_forward(ctx->stream, a, 0, 0);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_loss_grad(gsr_ctx* ctx, const gsr_camera* cam, const float* image, const float* target, float lambda,
                  float* loss_grad) {
  return gsr_loss_grad_window(ctx, cam, image, target, lambda, GSR_SSIM_BOX7, loss_grad);
}

int gsr_loss_grad_window(gsr_ctx* ctx, const gsr_camera* cam, const float* image, const float* target,
                         float lambda, int window, float* loss_grad) {
  if (!ctx) return GSR_E_INVALID;
  if (!image || !target || !loss_grad) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (window != GSR_SSIM_BOX7 && window != GSR_SSIM_GAUSS11)
    return fail(ctx, GSR_E_INVALID, "unknown SSIM window %d", window);
  if (!(lambda >= 0.0f && lambda <= 1.0f)) return fail(ctx, GSR_E_INVALID, "lambda %g not in [0, 1]", lambda);
  ViewDesc* dview;
  gsr_camera* dcam;
  ViewDesc host;
  int rc = single_view(ctx, cam, &dview, &dcam, nullptr, &host);
  if (rc) return rc;
  float* coef = nullptr;
  if (lambda > 0.0f && (rc = pooled_t(ctx, "ssim_coef", size_t(host.width) * host.height * 9, &coef))) return rc;
  launch_loss_grad(ctx->stream, dview, &host, 1, image, target, lambda, coef, loss_grad, window);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_backward(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, const uint32_t* order,
                 const uint32_t* ranges, const float* image, const float* loss_grad, float* grads) {
  if (!ctx) return GSR_E_INVALID;
  if (!splats || !order || !ranges || !image || !loss_grad || !grads)
    return fail(ctx, GSR_E_INVALID, "NULL argument");
  RasterArgs a{};
  ViewDesc host;
  int rc = raster_common(ctx, cam, a, host);
  if (rc) return rc;
  a.splats = splats;
  a.order = order;
  a.ranges[0] = ranges;
  a.image = const_cast<float*>(image);
  a.dl = loss_grad;
  a.grads = grads;
  launch_backward(ctx->stream, a, 0, 0);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_shape(gsr_ctx* ctx, const gsr_camera* cam, const float* gs, const float* sh, int sh_degree, uint32_t n,
              const float* grads, float* shape, float* sh_grad) {
  if (!ctx) return GSR_E_INVALID;
  if (n && (!gs || !grads || !shape)) return fail(ctx, GSR_E_INVALID, "NULL argument");
  if (sh_degree < 0 || sh_degree > 3) return fail(ctx, GSR_E_INVALID, "sh_degree %d not in [0, 3]", sh_degree);
  if (sh_degree > 0 && (!sh || !sh_grad)) return fail(ctx, GSR_E_INVALID, "sh / sh_grad is NULL");
  ViewDesc* dview;
  gsr_camera* dcam;
  ViewDesc host;
  int rc = single_view(ctx, cam, &dview, &dcam, nullptr, &host);
  if (rc) return rc;
  launch_shape(ctx->stream, dcam, 1, gs, sh, sh_degree, n, const_cast<float*>(grads), 0, shape, sh_grad, nullptr, 0);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_adam(gsr_ctx* ctx, const gsr_step* step, float* gs, const float* grads, const float* shape,
             float* moments) {
  if (!ctx) return GSR_E_INVALID;
  if (!step || (step->count && (!gs || !grads || !shape || !moments))) return fail(ctx, GSR_E_INVALID, "NULL argument");
  launch_adam(ctx->stream, *step, gs, grads, shape, moments);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_adam_sh(gsr_ctx* ctx, const gsr_step* step, float lr_sh_rest, float* gs, float* sh, const float* grads,
                const float* shape, const float* sh_grad, float* moments) {
  if (!ctx) return GSR_E_INVALID;
  if (!step || (step->count && (!gs || !sh || !grads || !shape || !sh_grad || !moments)))
    return fail(ctx, GSR_E_INVALID, "NULL argument");
  for (const void* p : {static_cast<const void*>(gs), static_cast<const void*>(sh), static_cast<const void*>(sh_grad),
                        static_cast<const void*>(moments)})
    if (reinterpret_cast<uintptr_t>(p) % 16)
      return fail(ctx, GSR_E_INVALID, "gs, sh, sh_grad and moments must be 16-byte aligned");
  launch_adam_sh(ctx->stream, *step, lr_sh_rest, gs, sh, grads, shape, sh_grad, moments);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

int gsr_stats(gsr_ctx* ctx, const gsr_step* step, const float* grads, const float* splats, float* stats) {
  if (!ctx) return GSR_E_INVALID;
  if (!step || (step->count && (!grads || !splats || !stats))) return fail(ctx, GSR_E_INVALID, "NULL argument");
  launch_stats(ctx->stream, *step, grads, splats, stats);
  GSR_LAUNCHED(ctx);
  return GSR_OK;
}

}  // extern "C"

namespace gsr {
hipStream_t ctx_stream(gsr_ctx* ctx) { return ctx->stream; }
int ctx_num_cus(gsr_ctx* ctx) { return ctx->num_cus; }
int ctx_fail(gsr_ctx* ctx, int code, const char* msg) { return fail(ctx, code, "%s", msg); }
// Device memory on the context's device, whichever device the calling thread
// has current (the current device is per thread).
hipError_t ctx_malloc(gsr_ctx* ctx, void** p, size_t bytes) {
  const hipError_t e = hipSetDevice(ctx->device);
  return e != hipSuccess ? e : hipMalloc(p, bytes);
}
}  // namespace gsr
