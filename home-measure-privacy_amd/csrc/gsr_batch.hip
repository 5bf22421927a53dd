// Batched views: one shared Gaussian set rendered and differentiated from many
// cameras of any resolution per step (the reference renders one view per step,
// trainer/gpu.js:915-933).
//
// Step = project (all views, one launch) -> per-view depth radix sort (one
// segmented launch per pass) -> tile binning (count, scan, emit) -> persistent
// forward over a work queue of every view's tiles -> loss gradient -> persistent
// backward -> chain to the 3D parameters summed over views.  The sequence is
// captured once into a HIP graph and replayed; every size it launches with is
// fixed by the cameras except the entry count, which is bounded by the order
// capacity: an overflowing step is detected, storage grows and the step reruns.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "gsr_internal.h"

using namespace gsr;

struct gsr_ctx;  // defined in gsr_api.hip
extern "C" const char* gsr_last_error(const gsr_ctx* ctx);

namespace gsr {
// Accessors into the context (gsr_api.hip).
hipStream_t ctx_stream(gsr_ctx* ctx);
int ctx_num_cus(gsr_ctx* ctx);
int ctx_fail(gsr_ctx* ctx, int code, const char* msg);
hipError_t ctx_malloc(gsr_ctx* ctx, void** p, size_t bytes);
}  // namespace gsr

constexpr uint32_t kLiveBase = 32;
constexpr uint32_t kQueueBase = 64;
// Per-view gradient records hold the backward's nine pixel moments, packed 9
// floats (36 bytes) apart: the chain reads every record of every view, whole
// cache lines at a time, so padding would be fetched with them.
constexpr uint32_t kRecord = 9;
constexpr uint32_t kFlagWords = kQueueBase + 9 * (gsr::kMaxPasses + 1) * gsr::kQueueStride;

struct gsr_batch {
  gsr_ctx* ctx = nullptr;
  uint32_t V = 0, n = 0;
  int deg = 0;
  std::vector<gsr_camera> cams;
  std::vector<ViewDesc> views;
  uint32_t total_tiles = 0, max_tiles = 0;
  uint64_t total_pixels = 0;
  uint32_t chunk = 0, tile_batch = 1;
  uint32_t grid_fwd = 0, grid_bwd = 0;
  // device
  gsr_camera* d_cams = nullptr;
  ViewDesc* d_views = nullptr;
  TileWork* d_work = nullptr;
  float* splats = nullptr;
  uint32_t *keys = nullptr, *keys_alt = nullptr, *vals = nullptr, *vals_alt = nullptr;
  uint2* rects = nullptr;
  uint2* rect_sorted = nullptr;
  uint32_t *hist = nullptr, *piece = nullptr;
  // The range of the depth keys' top 16 bits (three-pass depth sort): [V][2],
  // then the chunks' [V][chunks][2].
  uint32_t* key_range = nullptr;
  // The depth sort takes four passes from the first step whose view had keys too
  // wide for the three-pass sort (its order was not used: the step ran again).
  bool wide_keys = false;
  uint32_t* matrix = nullptr;
  uint32_t* tile_tot = nullptr;
  // Depth-slice passes: pass p bins sorted positions [pass_end[p-1], pass_end[p]).
  uint32_t num_passes = 1;
  uint32_t growth = 1;
  uint32_t pass_end[kMaxPasses] = {};
  uint64_t* entries = nullptr;  // [kMaxPasses][V]
  uint64_t* ebase = nullptr;    // [kMaxPasses][V]
  uint64_t* run = nullptr;      // slots used so far
  uint32_t* ranges = nullptr;   // [kMaxPasses][total_tiles][2]
  float* tstate = nullptr;      // per pixel transmittance between forward passes
  uint8_t* done = nullptr;      // per tile: saturated
  uint8_t* band_open = nullptr; // per tile, 4 bytes: band still open after the last forward pass
  // Later slices bin only the Gaussians that reach an open tile (launch_slice_filter):
  // their rectangles and indices go to the depth sort's scratch (rects, vals_alt),
  // unused once the sort is done.
  uint16_t* open_sat = nullptr;    // [total_tiles]: summed-area table of the open tiles
  uint64_t* slice_mask = nullptr;  // [V][ceil(n / 64)]
  uint32_t* slice_piece = nullptr; // [V][slice_filter_pieces(n)]
  uint32_t* slice_kept = nullptr;  // [kMaxPasses][V]: Gaussians kept for each later pass
  int ssim_window = GSR_SSIM_BOX7;
  uint32_t* walk = nullptr;     // [kMaxPasses][total_tiles]: entries the forward examined per pass
  uint32_t* prev_walk = nullptr;  // [total_tiles]: the previous step's first-slice walk (a cost estimate)
  TileWork* heavy = nullptr;    // [2 total_tiles]: half-tile items of the split tiles of a launch
  TileWork* live_work = nullptr;  // [total_tiles]: the whole-tile items of a launch (live list)
  uint32_t* live_counts = nullptr;  // [ceil(total_tiles / 256)]: live list scratch
  uint8_t* split = nullptr;     // [total_tiles]: tile split in the current launch
  // Tiles are blended as two half-tile items when their cost exceeds these (0: never):
  // the forward of the first depth slice (most tiles saturate early, so list
  // length says little about the walk), the forward of later slices (unsaturated
  // tiles walk their whole segment) and the backward (entries the forward walked).
  uint32_t split_first = 0, split_later = 0, split_backward = 1024;
  // Tiles per band of the emit kernel: 128 (both workloads measured emit faster
  // there than at 256 since its ranking rounds went branch-light).
  uint32_t emit_band = 128;
  int no_cull = 0;           // verification knob: no hit-box binning or culling
  uint64_t* valid_mask = nullptr;  // [V][ceil(n / 64)]: projected validly (STATS)
  uint32_t* order = nullptr;
  uint64_t order_cap = 0;
  // [0] overflow (bit 0: entry storage, bit 1: a view's depth keys too wide for the
  // three-pass sort), [16 + p] split-item count of forward pass p, [16 + kMaxPasses]
  // the backward's; [kLiveBase + p] live-list length of forward pass p,
  // [kLiveBase + kMaxPasses] the backward's; from kQueueBase, 9 work counters
  // per forward pass, then the backward's (a shared one and one per die part),
  // kQueueStride words apart.
  uint32_t* flags = nullptr;
  float* image = nullptr;
  float* dl = nullptr;
  float* coef = nullptr;
  float* vgrads = nullptr;
  uint64_t* pinned = nullptr;  // host: entries[kMaxPasses][V], overflow
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> graph_key;
  float graph_lambda = -1.0f;
};

namespace {

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

template <class T>
int dalloc(gsr_batch* b, T** p, size_t count) {
  dfree(*p);
  *p = nullptr;
  size_t bytes = std::max<size_t>(256, count * sizeof(T));
  if (ctx_malloc(b->ctx, reinterpret_cast<void**>(p), bytes) != hipSuccess) {
    *p = nullptr;
    return ctx_fail(b->ctx, GSR_E_NOMEM, "out of device memory for a batch buffer");
  }
  launch_zero(ctx_stream(b->ctx), *p, bytes / 4 * 4);
  return GSR_OK;
}

void drop_graph(gsr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

int layout(gsr_batch* b) {
  b->views.resize(b->V);
  uint32_t tb = 0;
  uint64_t pb = 0;
  b->max_tiles = 0;
  for (uint32_t v = 0; v < b->V; ++v) {
    const gsr_camera& c = b->cams[v];
    ViewDesc& d = b->views[v];
    d = ViewDesc{};
    d.width = uint32_t(c.width);
    d.height = uint32_t(c.height);
    d.tiles_x = (d.width + kTile - 1) / kTile;
    d.tiles_y = (d.height + kTile - 1) / kTile;
    d.num_tiles = d.tiles_x * d.tiles_y;
    d.tile_base = tb;
    d.pixel_base = pb;
    d.splat_base = uint64_t(v) * b->n;
    tb += d.num_tiles;
    pb += uint64_t(d.width) * d.height;
    b->max_tiles = std::max(b->max_tiles, d.num_tiles);
  }
  b->total_tiles = tb;
  b->total_pixels = pb;
  return GSR_OK;
}

int alloc_all(gsr_batch* b) {
  const size_t vn = size_t(b->V) * std::max<uint32_t>(b->n, 1);
  int rc;
  hipStream_t s = ctx_stream(b->ctx);
  if ((rc = dalloc(b, &b->d_cams, b->V)) || (rc = dalloc(b, &b->d_views, b->V)) ||
      (rc = dalloc(b, &b->d_work, b->total_tiles)) || (rc = dalloc(b, &b->splats, vn * 12)) ||
      (rc = dalloc(b, &b->keys, vn)) || (rc = dalloc(b, &b->keys_alt, vn)) || (rc = dalloc(b, &b->vals, vn)) ||
      (rc = dalloc(b, &b->vals_alt, vn)) || (rc = dalloc(b, &b->rects, vn)) ||
      (rc = dalloc(b, &b->rect_sorted, vn)) ||
      (rc = dalloc(b, &b->hist, std::max<size_t>(1, radix_hist_elems(b->n, b->V, true)))) ||
      (rc = dalloc(b, &b->piece,
                   std::max<size_t>(1, scan_piece_elems(radix_hist_elems(b->n, 1, true), b->V)))) ||
      (rc = dalloc(b, &b->key_range, depth_sort_range_elems(b->n, b->V))) ||
      (rc = dalloc(b, &b->tile_tot, b->total_tiles)) || (rc = dalloc(b, &b->entries, kMaxPasses * b->V)) ||
      (rc = dalloc(b, &b->ebase, kMaxPasses * b->V)) || (rc = dalloc(b, &b->run, 1)) ||
      (rc = dalloc(b, &b->ranges, size_t(kMaxPasses) * b->total_tiles * 2)) ||
      (rc = dalloc(b, &b->tstate, b->total_pixels)) || (rc = dalloc(b, &b->done, b->total_tiles)) ||
      (rc = dalloc(b, &b->band_open, size_t(b->total_tiles) * 4)) ||
      (rc = dalloc(b, &b->open_sat, b->total_tiles)) ||
      (rc = dalloc(b, &b->slice_mask, size_t(b->V) * ((size_t(b->n) + 63) / 64 + 1))) ||
      (rc = dalloc(b, &b->slice_piece, size_t(b->V) * slice_filter_pieces(std::max<uint32_t>(b->n, 1)))) ||
      (rc = dalloc(b, &b->slice_kept, size_t(kMaxPasses) * b->V)) ||
      (rc = dalloc(b, &b->walk, size_t(kMaxPasses) * b->total_tiles)) ||
      (rc = dalloc(b, &b->prev_walk, b->total_tiles)) ||
      (rc = dalloc(b, &b->heavy, size_t(b->total_tiles) * 2)) || (rc = dalloc(b, &b->split, b->total_tiles)) ||
      (rc = dalloc(b, &b->live_work, b->total_tiles)) ||
      (rc = dalloc(b, &b->live_counts, (size_t(b->total_tiles) + 255) / 256 + 1)) ||
      (rc = dalloc(b, &b->flags, kFlagWords)) || (rc = dalloc(b, &b->image, b->total_pixels * 3)) ||
      (rc = dalloc(b, &b->dl, b->total_pixels * 3)) ||
      (rc = dalloc(b, &b->vgrads, size_t(b->V) * std::max<uint32_t>(b->n, 1) * kRecord)) ||
      (rc = dalloc(b, &b->valid_mask, size_t(b->V) * valid_mask_words(std::max<uint32_t>(b->n, 1)))))
    return rc;
  (void)s;
  return GSR_OK;
}

int set_chunk(gsr_batch* b, uint32_t chunk) {
  b->chunk = chunk;
  const size_t m = bin_matrix_elems(b->views.data(), b->V, b->n, chunk);
  int rc = dalloc(b, &b->matrix, std::max<size_t>(1, m));
  if (rc) return rc;
  drop_graph(b);
  return GSR_OK;
}

// passes == 1: full tile lists.  passes == k > 1: depth slices whose ends grow
// by `growth`, the first holding ceil(n / growth^(k-1)) Gaussians.
void set_passes(gsr_batch* b, uint32_t passes, uint32_t growth) {
  b->num_passes = passes;
  b->growth = growth;
  uint64_t div = 1;
  for (uint32_t p = 1; p < passes; ++p) div *= growth;
  for (uint32_t p = 0; p < passes; ++p) {
    b->pass_end[p] = p + 1 == passes ? b->n : uint32_t((uint64_t(b->n) + div - 1) / div);
    div /= growth;
  }
  drop_graph(b);
}

int upload_cams(gsr_batch* b) {
  hipStream_t s = ctx_stream(b->ctx);
  if (hipMemcpyAsync(b->d_cams, b->cams.data(), b->V * sizeof(gsr_camera), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_views, b->views.data(), b->V * sizeof(ViewDesc), hipMemcpyHostToDevice, s) != hipSuccess)
    return ctx_fail(b->ctx, GSR_E_HIP, "camera upload failed");
  launch_fill_work(s, b->d_views, b->V, b->total_tiles, b->d_work);
  if (hipStreamSynchronize(s) != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "camera upload failed");
  return GSR_OK;
}

// Stage boundaries and workload counters of a profiled step.
struct Prof {
  hipEvent_t ev[7];  // boundaries around project, sort, binning + forward, loss, backward, chain
  float binned_ms[3];             // bin count, bin emit, forward (summed over passes)
  uint32_t* last;                 // optional, [kMaxPasses][total_tiles]: furthest last hit - segment start
  uint32_t* touched;
};

// Issues one full step on the stream (capturable: kernels only).
void issue_step(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda, float* grads,
                float* shape, float* sh_grad, float* stats, float* images, Prof* prof = nullptr) {
  hipStream_t s = ctx_stream(b->ctx);
  float* image = images ? images : b->image;
  int stage = 0;
  auto mark = [&]() {
    if (prof) (void)hipEventRecord(prof->ev[stage++], s);  // a failed record fails the timing query
  };
  launch_zero(s, b->flags, kFlagWords * sizeof(uint32_t));
  // The first slice's walk of the previous step estimates this step's: the
  // forward splits its longest tiles by it (scheduling only).
  if (b->split_first) launch_copy(s, b->prev_walk, b->walk, size_t(b->total_tiles) * sizeof(uint32_t));
  launch_zero(s, b->walk, size_t(b->num_passes) * b->total_tiles * sizeof(uint32_t));
  if (prof && prof->last) launch_zero(s, prof->last, size_t(b->num_passes) * b->total_tiles * sizeof(uint32_t));
  mark();
  uint32_t axis = 0;
  for (const ViewDesc& vd : b->views) axis = std::max(axis, std::max(vd.tiles_x, vd.tiles_y));
  const bool packed = rect_packed(axis);
  launch_project(s, b->d_cams, b->d_views, b->V, gs, sh, b->deg, b->n, b->splats, b->keys, b->rects, nullptr,
                 packed, !b->no_cull, b->valid_mask);
  mark();
  SortScratch sc{b->keys_alt, b->vals, b->vals_alt, b->hist, b->piece};
  launch_depth_sort(s, b->keys, b->n, b->V, sc, b->rects, b->rect_sorted, packed,
                    b->wide_keys ? nullptr : b->key_range, b->flags);
  mark();
  RasterArgs a{};
  a.views = b->d_views;
  a.work = b->d_work;
  a.num_work = b->total_tiles;
  a.splats = b->splats;
  a.order = b->order;
  a.order_capacity = b->order_cap;
  a.num_passes = int(b->num_passes);
  for (uint32_t p = 0; p < b->num_passes; ++p) {
    a.ranges[p] = b->ranges + size_t(p) * b->total_tiles * 2;
    a.ebase[p] = b->ebase + size_t(p) * b->V;
  }
  a.image = image;
  a.state = nullptr;
  a.dl = b->dl;
  a.grads = b->vgrads;
  a.grads_view_stride = uint64_t(b->n) * kRecord;
  a.grads_record = kRecord;
  a.moments = 1;  // the shape launch below maps them to gradients
  a.tile_batch = b->tile_batch;
  a.tstate = b->num_passes > 1 ? b->tstate : nullptr;
  a.done = b->num_passes > 1 ? b->done : nullptr;
  a.band_open = b->num_passes > 1 ? b->band_open : nullptr;
  a.no_cull = b->no_cull;
  // Binning and forward alternate per depth slice: a slice is binned only for
  // the tiles the slices before it left unsaturated.  Profiled steps time all
  // binning as "bin count" + "bin emit" and all forward passes as "forward".
  BinScratch bs{b->matrix, b->tile_tot, b->rect_sorted, b->run, packed, b->emit_band};
  float t_count = 0.f, t_emit = 0.f, t_fwd = 0.f;
  hipEvent_t e0, e1, e2, e3;
  if (prof) {
    // Profiling only: a failed event call leaves the time at 0 and the step as it is.
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventCreate(&e2);
    (void)hipEventCreate(&e3);
  }
  for (uint32_t p = 0; p < b->num_passes; ++p) {
    BinPass pass{p ? b->pass_end[p - 1] : 0u, b->pass_end[p], p ? b->done : nullptr,
                 b->ranges + size_t(p) * b->total_tiles * 2, b->ebase + size_t(p) * b->V,
                 b->entries + size_t(p) * b->V, p == 0};
    if (prof) (void)hipEventRecord(e0, s);
    if (p > 0 && slice_filter_fits(b->max_tiles)) {
      uint32_t* kept = b->slice_kept + size_t(p) * b->V;
      launch_slice_filter(s, b->d_views, b->V, b->max_tiles, b->n, pass.pbeg, pass.pend, b->rect_sorted, b->vals,
                          packed, b->done, SliceFilter{b->open_sat, b->slice_mask, b->slice_piece}, b->rects,
                          b->vals_alt, kept);
      pass.rects = b->rects;
      pass.vals = b->vals_alt;
      pass.kept = kept;
    }
    launch_bin_count(s, b->d_views, b->views.data(), b->V, b->n, b->chunk, bs, pass);
    if (prof) (void)hipEventRecord(e1, s);
    launch_bin_emit(s, b->d_views, b->views.data(), b->V, b->n, b->chunk, b->vals, bs, pass, b->order, b->order_cap,
                    b->flags);
    if (prof) (void)hipEventRecord(e2, s);
    a.pass = int(p);
    a.queue = b->flags + kQueueBase + 9 * p * kQueueStride;
    const uint32_t split_cost = p ? b->split_later : b->split_first;
    a.split = split_cost ? b->split : nullptr;
    a.heavy = b->heavy;
    a.heavy_count = b->flags + 16 + p;
    if (split_cost) {
      SplitArgs sa{};
      sa.tiles = b->d_work;
      sa.total_tiles = b->total_tiles;
      sa.split_cost = split_cost;
      sa.ranges[0] = a.ranges[p];
      sa.num_passes = 1;
      if (p == 0) sa.walk = b->prev_walk;
      sa.split = b->split;
      sa.heavy = b->heavy;
      sa.heavy_count = b->flags + 16 + p;
      launch_split_list(s, sa);
    }
    // Later slices blend only the live tiles: most are saturated by then, and a
    // saturated or empty tile writes nothing a later stage reads.  (The first
    // slice blends every tile: an empty one still writes its image, T and
    // open bands.)
    a.work = b->d_work;
    a.work_count = nullptr;
    if (p > 0) {
      LiveArgs la{};
      la.tiles = b->d_work;
      la.total_tiles = b->total_tiles;
      la.ranges[0] = a.ranges[p];
      la.num_passes = 1;
      la.done = b->done;
      la.split = a.split;
      la.counts = b->live_counts;
      la.out = b->live_work;
      la.out_count = b->flags + kLiveBase + p;
      launch_live_list(s, la);
      a.work = b->live_work;
      a.work_count = la.out_count;
    }
    a.walk = b->walk + size_t(p) * b->total_tiles;
    a.last = prof && prof->last ? prof->last + size_t(p) * b->total_tiles : nullptr;
    launch_forward(s, a, 1, b->grid_fwd);
    if (a.done && p + 1 < b->num_passes) launch_done_update(s, b->total_tiles, b->band_open, b->done, p == 0);
    if (prof) {
      (void)hipEventRecord(e3, s);
      (void)hipEventSynchronize(e3);
      float x;
      if (hipEventElapsedTime(&x, e0, e1) != hipSuccess) x = 0.f;
      t_count += x;
      if (hipEventElapsedTime(&x, e1, e2) != hipSuccess) x = 0.f;
      t_emit += x;
      if (hipEventElapsedTime(&x, e2, e3) != hipSuccess) x = 0.f;
      t_fwd += x;
    }
  }
  if (prof) {
    prof->binned_ms[0] = t_count;
    prof->binned_ms[1] = t_emit;
    prof->binned_ms[2] = t_fwd;
    for (hipEvent_t e : {e0, e1, e2, e3}) (void)hipEventDestroy(e);
  }
  a.walk = nullptr;
  a.last = nullptr;
  mark();
  // Without the SSIM term the backward forms the L1 gradient itself.
  if (lambda > 0.0f)
    launch_loss_grad(s, b->d_views, b->views.data(), b->V, image, targets, lambda, b->coef, b->dl, b->ssim_window);
  a.l1_target = lambda > 0.0f ? nullptr : targets;
  mark();
  a.split = b->split_backward ? b->split : nullptr;
  a.heavy_count = b->flags + 16 + kMaxPasses;
  if (b->split_backward) {
    SplitArgs sa{};
    sa.tiles = b->d_work;
    sa.total_tiles = b->total_tiles;
    sa.split_cost = b->split_backward;
    for (uint32_t p = 0; p < b->num_passes; ++p) sa.ranges[p] = a.ranges[p];
    sa.num_passes = int(b->num_passes);
    sa.walk = b->walk;
    sa.split = b->split;
    sa.heavy = b->heavy;
    sa.heavy_count = b->flags + 16 + kMaxPasses;
    launch_split_list(s, sa);
  }
  // The backward blends the tiles with entries in some slice (it writes only
  // gradients, so an empty tile has nothing to do).
  {
    LiveArgs la{};
    la.tiles = b->d_work;
    la.total_tiles = b->total_tiles;
    for (uint32_t p = 0; p < b->num_passes; ++p) la.ranges[p] = a.ranges[p];
    la.num_passes = int(b->num_passes);
    la.split = a.split;
    la.counts = b->live_counts;
    la.out = b->live_work;
    la.out_count = b->flags + kLiveBase + kMaxPasses;
    launch_live_list(s, la);
    a.work = b->live_work;
    a.work_count = la.out_count;
  }
  a.queue = b->flags + kQueueBase + 9 * kMaxPasses * kQueueStride;
  launch_backward(s, a, 1, b->grid_bwd);
  mark();
  launch_shape(s, b->d_cams, b->V, gs, sh, b->deg, b->n, b->vgrads, uint64_t(b->n) * kRecord, shape, sh_grad, grads,
               1, prof && prof->last ? prof->touched : nullptr, kRecord, b->valid_mask, stats, b->flags, 1);
  mark();
}

int run_step(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda, float* grads,
             float* shape, float* sh_grad, float* stats, float* images, int use_graph) {
  hipStream_t s = ctx_stream(b->ctx);
  if (lambda > 0.0f && !b->coef) {
    int rc = dalloc(b, &b->coef, b->total_pixels * 9);
    if (rc) return rc;
    drop_graph(b);
  }
  for (int attempt = 0; attempt < 4; ++attempt) {
    if (use_graph) {
      std::vector<const void*> key = {gs, sh, targets, grads, shape, sh_grad, stats, images};
      if (!b->exec || key != b->graph_key || lambda != b->graph_lambda) {
        drop_graph(b);
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return ctx_fail(b->ctx, GSR_E_HIP, "hipStreamBeginCapture failed");
        issue_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, stats, images);
        hipError_t e = hipStreamEndCapture(s, &b->graph);
        if (e != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "hipStreamEndCapture failed");
        if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
          return ctx_fail(b->ctx, GSR_E_HIP, "hipGraphInstantiate failed");
        b->graph_key = key;
        b->graph_lambda = lambda;
      }
      if (hipGraphLaunch(b->exec, s) != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "hipGraphLaunch failed");
    } else {
      issue_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, stats, images);
      if (hipGetLastError() != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "kernel launch failed");
    }
    const size_t ne = size_t(b->num_passes) * b->V;
    if (hipMemcpyAsync(b->pinned, b->entries, ne * sizeof(uint64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipMemcpyAsync(b->pinned + ne, b->flags, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
      return ctx_fail(b->ctx, GSR_E_HIP, "readback failed");
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      std::string m = std::string("batch step failed: ") + hipGetErrorString(e);
      return ctx_fail(b->ctx, GSR_E_HIP, m.c_str());
    }
    uint64_t total = 0;
    for (size_t k = 0; k < ne; ++k) total += b->pinned[k];
    const uint32_t flag = uint32_t(b->pinned[ne]);
    const bool overflow = (flag & 1u) != 0 || total > b->order_cap;
    if (!overflow && !(flag & 2u)) return GSR_OK;
    if (flag & 2u) {
      // A view's depth keys were too wide for the three-pass sort, so its tile
      // lists were not in depth order: four passes from now on.
      b->wide_keys = true;
      drop_graph(b);
    }
    if (overflow) {
      // Grow the entry storage.
      uint64_t cap = total + total / 4 + 1024;
      int rc = dalloc(b, &b->order, cap);
      if (rc) return rc;
      b->order_cap = cap;
      drop_graph(b);
    }
    // and run the step again.
  }
  return ctx_fail(b->ctx, GSR_E_CAPACITY, "entry storage kept overflowing");
}

}  // namespace

extern "C" {

int gsr_batch_create(gsr_ctx* ctx, uint32_t num_views, const gsr_camera* cams, uint32_t n, int sh_degree,
                     gsr_batch** out) {
  if (!ctx) return GSR_E_INVALID;
  if (!out || !cams || num_views == 0) return ctx_fail(ctx, GSR_E_INVALID, "no cameras");
  if (sh_degree < 0 || sh_degree > 3) return ctx_fail(ctx, GSR_E_INVALID, "sh_degree not in [0, 3]");
  for (uint32_t v = 0; v < num_views; ++v) {
    const gsr_camera& c = cams[v];
    if (!(c.width >= 1.0f && c.height >= 1.0f && c.width <= 65536.0f && c.height <= 65536.0f) ||
        c.width != float(uint32_t(c.width)) || c.height != float(uint32_t(c.height)))
      return ctx_fail(ctx, GSR_E_INVALID, "camera size is not a positive integer size");
  }
  gsr_batch* b = new gsr_batch();
  b->ctx = ctx;
  b->V = num_views;
  b->n = n;
  b->deg = sh_degree;
  b->cams.assign(cams, cams + num_views);
  int rc = layout(b);
  if (!rc) rc = alloc_all(b);
  uint64_t items = uint64_t(n) * num_views;
  uint32_t chunk = 256;
  while (chunk < 4096 && uint64_t(chunk) * 2048 < items) chunk <<= 1;
  if (!rc) rc = set_chunk(b, chunk);
  // Depth slices: most tiles of an opaque scene saturate within the nearest
  // eighth of the Gaussians, and the rest rarely saturate at all.
  if (!rc) set_passes(b, n >= 65536 ? 2 : 1, n >= 65536 ? 8 : 1);
  if (!rc && hipHostMalloc(reinterpret_cast<void**>(&b->pinned), (kMaxPasses * num_views + 8) * sizeof(uint64_t),
                           hipHostMallocDefault) != hipSuccess)
    rc = ctx_fail(ctx, GSR_E_HIP, "hipHostMalloc failed");
  if (!rc) rc = upload_cams(b);
  if (!rc) {
    // Persistent grids: every CU filled to the kernels' occupancy.
    int per_cu_f = 0, per_cu_b = 0;
    raster_occupancy(&per_cu_f, &per_cu_b);
    b->grid_fwd = uint32_t(std::max(1, per_cu_f) * ctx_num_cus(ctx));
    b->grid_bwd = uint32_t(std::max(1, per_cu_b) * ctx_num_cus(ctx));
    // A first guess of the entry storage: 16 entries per Gaussian and view.
    uint64_t cap = std::max<uint64_t>(1024, items * 16);
    rc = dalloc(b, &b->order, cap);
    b->order_cap = cap;
  }
  if (rc) {
    gsr_batch_destroy(b);
    return rc;
  }
  *out = b;
  return GSR_OK;
}

void gsr_batch_destroy(gsr_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(ctx_stream(b->ctx));
  drop_graph(b);
  void* ptrs[] = {b->d_cams, b->d_views, b->d_work, b->splats, b->keys, b->keys_alt, b->vals, b->vals_alt,
                  b->rects, b->rect_sorted, b->hist, b->piece, b->matrix, b->tile_tot, b->entries, b->ebase,
                  b->run, b->ranges, b->tstate, b->done, b->order, b->flags, b->image, b->dl, b->coef, b->vgrads,
                  b->band_open, b->walk, b->heavy, b->split, b->prev_walk, b->live_work, b->live_counts,
                  b->key_range, b->valid_mask, b->open_sat, b->slice_mask, b->slice_piece, b->slice_kept};
  for (void* p : ptrs) dfree(p);
  if (b->pinned) (void)hipHostFree(b->pinned);
  delete b;
}

int gsr_batch_set_cameras(gsr_batch* b, const gsr_camera* cams) {
  if (!b || !cams) return GSR_E_INVALID;
  for (uint32_t v = 0; v < b->V; ++v)
    if (cams[v].width != b->cams[v].width || cams[v].height != b->cams[v].height)
      return ctx_fail(b->ctx, GSR_E_INVALID, "camera resolution changed");
  b->cams.assign(cams, cams + b->V);
  return upload_cams(b);
}

int gsr_batch_step_stats(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda,
                         float* grads, float* shape, float* sh_grad, float* stats, float* images, int use_graph) {
  if (!b) return GSR_E_INVALID;
  if ((b->n && (!gs || !grads || !shape)) || !targets) return ctx_fail(b->ctx, GSR_E_INVALID, "NULL argument");
  if (b->deg > 0 && (!sh || !sh_grad)) return ctx_fail(b->ctx, GSR_E_INVALID, "sh / sh_grad is NULL");
  if (!(lambda >= 0.0f && lambda <= 1.0f)) return ctx_fail(b->ctx, GSR_E_INVALID, "lambda not in [0, 1]");
  return run_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, stats, images, use_graph);
}

int gsr_batch_step(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda, float* grads,
                   float* shape, float* sh_grad, float* images, int use_graph) {
  return gsr_batch_step_stats(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, images, use_graph);
}

int gsr_batch_set_culling(gsr_batch* b, int enabled) {
  if (!b) return GSR_E_INVALID;
  const int no_cull = enabled ? 0 : 1;
  if (no_cull != b->no_cull) {
    b->no_cull = no_cull;
    drop_graph(b);  // the flag is a launch argument of the captured step
  }
  return GSR_OK;
}

int gsr_batch_tile_lists(gsr_batch* b, uint32_t view, uint32_t pass, uint32_t* order, uint64_t order_capacity,
                         uint32_t* ranges, uint64_t* num_entries) {
  if (!b) return GSR_E_INVALID;
  if (!ranges || !num_entries) return ctx_fail(b->ctx, GSR_E_INVALID, "NULL argument");
  if (view >= b->V || pass >= b->num_passes) return ctx_fail(b->ctx, GSR_E_INVALID, "view or pass out of range");
  if (order_capacity && !order) return ctx_fail(b->ctx, GSR_E_INVALID, "order is NULL with a nonzero capacity");
  hipStream_t s = ctx_stream(b->ctx);
  const size_t k = size_t(pass) * b->V + view;
  uint64_t ebase = 0, count = 0;
  if (hipMemcpyAsync(&ebase, b->ebase + k, sizeof(uint64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&count, b->entries + k, sizeof(uint64_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return ctx_fail(b->ctx, GSR_E_HIP, "reading the entry counts failed");
  *num_entries = count;
  const ViewDesc& vd = b->views[view];
  const size_t rbytes = size_t(vd.num_tiles) * 2 * sizeof(uint32_t);
  if (hipMemcpyAsync(ranges, b->ranges + (size_t(pass) * b->total_tiles + vd.tile_base) * 2, rbytes,
                     hipMemcpyDeviceToDevice, s) != hipSuccess)
    return ctx_fail(b->ctx, GSR_E_HIP, "copying the ranges failed");
  if (count > order_capacity) return ctx_fail(b->ctx, GSR_E_CAPACITY, "order is too small for the view's entries");
  if (ebase + count > b->order_cap) return ctx_fail(b->ctx, GSR_E_INVALID, "no complete step to read");
  if (count && hipMemcpyAsync(order, b->order + ebase, count * sizeof(uint32_t), hipMemcpyDeviceToDevice, s) !=
                   hipSuccess)
    return ctx_fail(b->ctx, GSR_E_HIP, "copying the order failed");
  if (hipStreamSynchronize(s) != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "copying the tile lists failed");
  return GSR_OK;
}

int gsr_batch_entries(gsr_batch* b, uint64_t* entries) {
  if (!b || !entries) return GSR_E_INVALID;
  for (uint32_t v = 0; v < b->V; ++v) {
    entries[v] = 0;
    for (uint32_t p = 0; p < b->num_passes; ++p) entries[v] += b->pinned[size_t(p) * b->V + v];
  }
  return GSR_OK;
}

int gsr_batch_set_passes(gsr_batch* b, uint32_t passes, uint32_t growth) {
  if (!b) return GSR_E_INVALID;
  if (passes < 1 || passes > uint32_t(kMaxPasses)) return ctx_fail(b->ctx, GSR_E_INVALID, "passes not in [1, 8]");
  if (passes > 1 && (growth < 2 || growth > 16)) return ctx_fail(b->ctx, GSR_E_INVALID, "growth not in [2, 16]");
  if (passes == 1) growth = 1;
  uint64_t div = 1;
  for (uint32_t p = 1; p < passes; ++p) div *= growth;
  if (div > b->n) return ctx_fail(b->ctx, GSR_E_INVALID, "too few Gaussians for the passes");
  if (passes != b->num_passes || growth != b->growth) set_passes(b, passes, growth);
  return GSR_OK;
}

int gsr_batch_set_ssim_window(gsr_batch* b, int window) {
  if (!b) return GSR_E_INVALID;
  if (window != GSR_SSIM_BOX7 && window != GSR_SSIM_GAUSS11)
    return ctx_fail(b->ctx, GSR_E_INVALID, "unknown SSIM window");
  if (window != b->ssim_window) {
    b->ssim_window = window;
    drop_graph(b);
  }
  return GSR_OK;
}

int gsr_batch_configure(gsr_batch* b, uint32_t tile_batch, uint32_t chunk) {
  if (!b) return GSR_E_INVALID;
  if (tile_batch) {
    if (tile_batch > 256) return ctx_fail(b->ctx, GSR_E_INVALID, "tile_batch above 256");
    if (tile_batch != b->tile_batch) drop_graph(b);
    b->tile_batch = tile_batch;
  }
  if (chunk && chunk != b->chunk) {
    if (chunk < 64 || chunk > 65536) return ctx_fail(b->ctx, GSR_E_INVALID, "chunk not in [64, 65536]");
    return set_chunk(b, chunk);
  }
  return GSR_OK;
}

int gsr_batch_profile(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda,
                      float* grads, float* shape, float* sh_grad, float* stage_ms, uint64_t* counters) {
  if (!b || !stage_ms) return GSR_E_INVALID;
  // A regular step first: sizes the entry storage for these inputs.
  int rc = gsr_batch_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, 0);
  if (rc) return rc;
  hipStream_t s = ctx_stream(b->ctx);
  Prof prof{};
  const size_t nt = std::max<size_t>(1, b->total_tiles);
  const size_t np = size_t(b->num_passes) * nt;
  std::vector<uint32_t> tl(np), tw(np), touched(b->V);
  if (ctx_malloc(b->ctx, reinterpret_cast<void**>(&prof.last), size_t(kMaxPasses) * nt * 4) != hipSuccess ||
      ctx_malloc(b->ctx, reinterpret_cast<void**>(&prof.touched), b->V * 4) != hipSuccess) {
    dfree(prof.last);
    return ctx_fail(b->ctx, GSR_E_NOMEM, "profile buffers");
  }
  launch_zero(s, prof.touched, b->V * 4);
  hipError_t e = hipSuccess;
  auto keep = [&](hipError_t r) {  // the first failure wins
    if (e == hipSuccess) e = r;
  };
  for (auto& ev : prof.ev) keep(hipEventCreate(&ev));
  // Stage times come from a step that runs the same kernels as a regular step;
  // the counters from a second one that also tracks each tile's last hit.
  uint32_t* last = prof.last;
  prof.last = nullptr;
  if (e == hipSuccess) {
    issue_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, nullptr, &prof);
    keep(hipStreamSynchronize(s));
  }
  float binned[3] = {prof.binned_ms[0], prof.binned_ms[1], prof.binned_ms[2]};
  float marks[6] = {0, 0, 0, 0, 0, 0};
  if (e == hipSuccess) {
    keep(hipEventElapsedTime(marks + 0, prof.ev[0], prof.ev[1]));
    keep(hipEventElapsedTime(marks + 1, prof.ev[1], prof.ev[2]));
    keep(hipEventElapsedTime(marks + 2, prof.ev[3], prof.ev[4]));
    keep(hipEventElapsedTime(marks + 3, prof.ev[4], prof.ev[5]));
    keep(hipEventElapsedTime(marks + 4, prof.ev[5], prof.ev[6]));
    prof.last = last;
    launch_zero(s, prof.touched, b->V * 4);
    issue_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, nullptr, &prof);
    keep(hipStreamSynchronize(s));
  }
  if (e == hipSuccess) {
    // Stage order: project, depth sort, bin count, bin emit, forward, loss, backward, chain.
    stage_ms[0] = marks[0];
    stage_ms[1] = marks[1];
    stage_ms[2] = binned[0];
    stage_ms[3] = binned[1];
    stage_ms[4] = binned[2];
    stage_ms[5] = marks[2];
    stage_ms[6] = marks[3];
    stage_ms[7] = marks[4];
    std::vector<uint64_t> pe(size_t(b->num_passes) * b->V), ent(b->V, 0);
    if (counters) {
      keep(hipMemcpy(tl.data(), prof.last, np * 4, hipMemcpyDeviceToHost));
      keep(hipMemcpy(tw.data(), b->walk, np * 4, hipMemcpyDeviceToHost));
      keep(hipMemcpy(touched.data(), prof.touched, b->V * 4, hipMemcpyDeviceToHost));
      keep(hipMemcpy(pe.data(), b->entries, pe.size() * 8, hipMemcpyDeviceToHost));
    }
    if (counters && e == hipSuccess) {
      for (uint32_t p = 0; p < b->num_passes; ++p)
        for (uint32_t v = 0; v < b->V; ++v) ent[v] += pe[size_t(p) * b->V + v];
      for (uint32_t v = 0; v < b->V; ++v) {
        // Per tile: W = entries examined over the passes; F = the saturated
        // prefix, counted from the start of the first pass to the furthest last
        // hit (the walks of the passes before the last one with a hit, plus that
        // pass's furthest hit).
        uint64_t f = 0, w = 0;
        for (uint32_t t = 0; t < b->views[v].num_tiles; ++t) {
          const size_t g = b->views[v].tile_base + t;
          uint64_t before = 0, tile_f = 0;
          for (uint32_t p = 0; p < b->num_passes; ++p) {
            if (tl[p * nt + g]) tile_f = before + tl[p * nt + g];
            before += tw[p * nt + g];
          }
          f += tile_f;
          w += before;
        }
        counters[4 * v] = ent[v];
        counters[4 * v + 1] = f;
        counters[4 * v + 2] = touched[v];
        counters[4 * v + 3] = w;
      }
    }
  }
  for (auto& ev : prof.ev) (void)hipEventDestroy(ev);
  dfree(last);  // prof.last is still null when the first step failed
  dfree(prof.touched);
  if (e != hipSuccess) return ctx_fail(b->ctx, GSR_E_HIP, "profiled step failed");
  return GSR_OK;
}

int gsr_batch_set_split(gsr_batch* b, uint32_t forward_first, uint32_t forward_later, uint32_t backward) {
  if (!b) return GSR_E_INVALID;
  if (forward_first != b->split_first || forward_later != b->split_later || backward != b->split_backward) {
    b->split_first = forward_first;
    b->split_later = forward_later;
    b->split_backward = backward;
    drop_graph(b);  // the thresholds are launch arguments of the captured step
  }
  return GSR_OK;
}

int gsr_batch_tuning(gsr_batch* b, uint32_t* tuning) {
  if (!b || !tuning) return GSR_E_INVALID;
  tuning[0] = b->tile_batch;
  tuning[1] = b->chunk;
  tuning[2] = b->split_first;
  tuning[3] = b->split_later;
  tuning[4] = b->split_backward;
  tuning[5] = b->emit_band;
  return GSR_OK;
}

int gsr_batch_passes(gsr_batch* b, uint32_t* passes, uint32_t* growth) {
  if (!b || !passes || !growth) return GSR_E_INVALID;
  *passes = b->num_passes;
  *growth = b->growth;
  return GSR_OK;
}

int gsr_batch_set_emit_band(gsr_batch* b, uint32_t tiles) {
  if (!b) return GSR_E_INVALID;
  if (tiles < 64 || tiles > 4096) return ctx_fail(b->ctx, GSR_E_INVALID, "emit band not in [64, 4096] tiles");
  if (tiles != b->emit_band) {
    b->emit_band = tiles;
    drop_graph(b);  // the band is a launch argument (and the LDS size) of the captured step
  }
  return GSR_OK;
}

int gsr_batch_autotune(gsr_batch* b, const float* gs, const float* sh, const float* targets, float lambda,
                       float* grads, float* shape, float* sh_grad, uint32_t* tile_batch, uint32_t* chunk) {
  if (!b) return GSR_E_INVALID;
  // The shortest of kReps single steps (each returns after its stream is done):
  // the minimum is steadier than a mean against the host's own hiccups.
  constexpr int kReps = 8;
  // A candidate replaces the setting in effect only when it is more than 1%
  // faster, so near-ties keep the current knobs and the choice is repeatable.
  constexpr double kMargin = 0.99;
  auto time_min = [&](double& dt) {
    dt = 1e30;
    // Warm: captures the graph for the knobs in effect.
    int rc = gsr_batch_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, 1);
    if (rc) return rc;
    for (int k = 0; k < kReps; ++k) {
      const auto t0 = std::chrono::steady_clock::now();
      rc = gsr_batch_step(b, gs, sh, targets, lambda, grads, shape, sh_grad, nullptr, 1);
      if (rc) return rc;
      dt = std::min(dt, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    return int(GSR_OK);
  };
  // One knob at a time: `apply` sets candidate k, the first candidate being the
  // setting in effect.
  auto pick = [&](size_t count, auto&& apply, size_t& best_k) {
    double best = 1e30;
    best_k = 0;
    for (size_t k = 0; k < count; ++k) {
      int rc = apply(k);
      if (rc) return rc;
      double dt;
      if ((rc = time_min(dt))) return rc;
      if (k == 0 || dt < best * kMargin) {
        best = dt;
        best_k = k;
      }
    }
    return apply(best_k);
  };
  int rc;
  size_t k;
  // Depth slices first, since every later stage depends on them: fewer, larger
  // slices bin more entries, more slices run more binning and forward passes.
  // Scenes of spread Gaussians (the 32-view 1M scene) take 2 x 8, scenes of
  // dense clusters (the 3M-Gaussian 64-view training scene) more, smaller
  // slices: 3 x 4 bins 655M instead of 1.05G entries per step there (emit 10.0
  // -> 6.9 ms, step 52.6 -> 49.8 ms), and with the later slices filtered 4 x 4
  // is a little faster again (45.6 against 45.9 ms).
  if (b->n >= 65536) {
    std::vector<std::pair<uint32_t, uint32_t>> slices = {{b->num_passes, b->growth}};
    for (const auto& c : {std::make_pair(2u, 8u), std::make_pair(3u, 4u), std::make_pair(4u, 4u)})
      if (c != slices[0]) slices.push_back(c);
    if ((rc = pick(slices.size(),
                   [&](size_t i) { return gsr_batch_set_passes(b, slices[i].first, slices[i].second); }, k)))
      return rc;
  }
  // The knobs below interact (the chunk that suits the default split thresholds
  // need not suit the chosen ones), so they are picked twice over.
  for (int round = 0; round < 2; ++round) {
    // Tile batch x binning chunk.
    std::vector<std::pair<uint32_t, uint32_t>> grid = {{b->tile_batch, b->chunk}};
    for (uint32_t ch : {1024u, 2048u, 4096u, 8192u})
      for (uint32_t tb : {1u, 2u, 4u, 8u})
        if (tb != b->tile_batch || ch != b->chunk) grid.push_back({tb, ch});
    if ((rc = pick(grid.size(), [&](size_t i) { return gsr_batch_configure(b, grid[i].first, grid[i].second); }, k)))
      return rc;
    // The emit band: small views keep more waves resident with 128 tiles, large
    // ones rescan their chunks less with 256.
    std::vector<uint32_t> bands = {b->emit_band};
    for (uint32_t eb : {128u, 256u})
      if (eb != b->emit_band) bands.push_back(eb);
    if ((rc = pick(bands.size(), [&](size_t i) { return gsr_batch_set_emit_band(b, bands[i]); }, k))) return rc;
    // The walk above which the backward blends a tile as two half-tile items: low
    // enough to shorten the launch's tail, high enough not to repeat each entry's
    // staging for tiles that are merely long (scenes of dense clusters prefer a
    // higher one).
    auto values = [](uint32_t current, std::initializer_list<uint32_t> cands) {
      std::vector<uint32_t> v = {current};
      for (uint32_t c : cands)
        if (c != current) v.push_back(c);
      return v;
    };
    const std::vector<uint32_t> backs = values(b->split_backward, {512u, 1024u, 2048u, 4096u});
    if ((rc = pick(backs.size(),
                   [&](size_t i) { return gsr_batch_set_split(b, b->split_first, b->split_later, backs[i]); }, k)))
      return rc;
    // Whether (and from which list length) the forward's later slices split their
    // tiles: their unsaturated tiles walk whole segments, so a few long lists can
    // end the launch late; splitting them pays on some scenes only.
    const std::vector<uint32_t> laters = values(b->split_later, {0u, 1024u, 2048u, 4096u});
    if ((rc = pick(laters.size(),
                   [&](size_t i) { return gsr_batch_set_split(b, b->split_first, laters[i], b->split_backward); }, k)))
      return rc;
    // The first slice by the previous step's walk (its tiles mostly saturate early,
    // so their list lengths say little); time_min's warm step fills that estimate.
    const std::vector<uint32_t> firsts = values(b->split_first, {0u, 1024u, 2048u});
    if ((rc = pick(firsts.size(),
                   [&](size_t i) { return gsr_batch_set_split(b, firsts[i], b->split_later, b->split_backward); }, k)))
      return rc;
  }
  if (tile_batch) *tile_batch = b->tile_batch;
  if (chunk) *chunk = b->chunk;
  return GSR_OK;
}

}  // extern "C"
