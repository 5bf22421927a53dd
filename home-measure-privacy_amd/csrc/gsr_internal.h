// Internal declarations shared by the HIP translation units of libgsr.
//
// Device code follows the WGSL of trainer/gpu.js (cited per function).  The
// blends (gsr_raster.hip) are built with -ffp-contract=off: the backward pass
// replays the forward's blend arithmetic bit for bit, so every fused
// multiply-add is written out explicitly and identically in both.  So are the
// sort and the optimizer (reference op order).  The projection and the chain
// are built with contraction and hardware-accuracy division (see build.py).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gsr.h"

namespace gsr {

// Constants of the reference (gpu.js line in brackets).
constexpr int kTile = 16;                 // tile edge [118]
constexpr float kNearZ = 0.2f;            // near-plane reject [36]
constexpr float kDilation = 0.3f;         // 2D covariance dilation [45,47]
constexpr float kAlphaMax = 0.99f;        // alpha clamp [100]
constexpr float kAlphaMin = 1.0f / 255.0f;  // alpha threshold [101]
constexpr float kTStop = 1e-4f;           // transmittance stop [124]

constexpr uint32_t kInvalidKey = 0xFFFFFFFFu;

// Per-view data of a launch over several views.  Offsets are in elements of the
// buffer named: Gaussians for splats/rects/keys, pixels for images, entries for
// order.  ranges are per tile, indexed by the global tile id.
struct ViewDesc {
  uint32_t width, height, tiles_x, tiles_y;
  uint32_t tile_base;   // first global tile id of this view
  uint32_t num_tiles;
  uint64_t pixel_base;  // first pixel of this view in concatenated images
  uint64_t splat_base;  // view * n
};

// One unit of blend work: bands [first, last] of a 16x16 tile of one view, band
// b being tile rows 4b .. 4b + 3 (a whole tile is bands 0-3).  view packs the
// view index (bits 0-23), the first band (24-25) and the last band (26-27).
struct TileWork {
  uint32_t view;
  uint32_t xy;  // tx | ty << 16
};

__host__ __device__ inline TileWork make_work(uint32_t view, uint32_t tx, uint32_t ty, uint32_t first, uint32_t last) {
  return TileWork{view | (first << 24) | (last << 26), tx | (ty << 16)};
}
__host__ __device__ inline uint32_t work_view(const TileWork& w) { return w.view & 0xFFFFFFu; }
__host__ __device__ inline uint32_t work_first(const TileWork& w) { return (w.view >> 24) & 3u; }
__host__ __device__ inline uint32_t work_last(const TileWork& w) { return (w.view >> 26) & 3u; }

// ---- launch helpers (host side), implemented per translation unit -------------

// Projection (gsr_project.hip).  cams: device array of num_views cameras.
// Tile rectangles of the binning, x0 x1 y0 y1 inclusive (empty: x0 > x1): in a
// [V][n] buffer of uint2 (16 bits per coordinate), or packed into its first
// [V][n] uint32 words (8 bits per coordinate) when no view has more than
// kPackedRectTiles tiles along an axis (rect_packed), which halves the rectangle
// traffic of the sort and the binning.
constexpr uint32_t kPackedRectTiles = 256;
inline bool rect_packed(uint32_t max_tiles_axis) { return max_tiles_axis <= kPackedRectTiles; }
// splats [V][n][12]; keys/rects/counts optional [V][n] (binning inputs).  With
// hitbox the rectangles cover the tiles of each splat's hit box (where alpha can
// reach 1/255) inside the reference square, else the square itself (tileLists'
// rectangle).  valid_mask (optional, [V][ceil(n / 64)]): bit i of word
// [v][i / 64] set when Gaussian i projects validly in view v.
void launch_project(hipStream_t s, const gsr_camera* cams, const ViewDesc* views, uint32_t num_views,
                    const float* gs, const float* sh, int sh_degree, uint32_t n, float* splats, uint32_t* keys,
                    uint2* rects, uint32_t* counts, bool packed = false, bool hitbox = true,
                    uint64_t* valid_mask = nullptr);
__host__ __device__ inline size_t valid_mask_words(uint32_t n) { return (size_t(n) + 63) / 64; }
// Binning inputs from existing splat records (drop-in path).
void launch_prep(hipStream_t s, const ViewDesc* views, uint32_t num_views, const float* splats, uint32_t n,
                 uint32_t* keys, uint2* rects, bool packed);

// Sorting and binning (gsr_sort.hip).
struct SortScratch {
  uint32_t* keys_alt;   // [V][n]
  uint32_t* vals;       // [V][n]
  uint32_t* vals_alt;   // [V][n]
  uint32_t* hist;       // [V][256 or 512][chunks]
  uint32_t* piece;      // scan partials
};
// Histogram elements of the depth sort: 256 digits per chunk, or 512 (wide) for
// the three-pass sort's offset pass.
size_t radix_hist_elems(uint32_t n, uint32_t num_views, bool wide = false);
// Elements of the three-pass sort's range scratch: [V][2], then [V][chunks][2].
size_t depth_sort_range_elems(uint32_t n, uint32_t num_views);
size_t scan_piece_elems(size_t seg_len, uint32_t num_segs);
// Stable sort of (keys[v][i], i) per view by key; result indices in vals, and
// rect_sorted[v][p] = rects[v][vals[v][p]] (the sorted keys themselves are not
// kept: keys is scratch afterwards).  Four 8-bit passes; with range (scratch of
// depth_sort_range_elems; hist sized wide) three, valid while each view's keys'
// top 16 bits span at most 511 values: a wider view sets bit 1 of *range_flag,
// and its order must not be used.
void launch_depth_sort(hipStream_t s, uint32_t* keys, uint32_t n, uint32_t num_views, const SortScratch& sc,
                       const uint2* rects, uint2* rect_sorted, bool packed, uint32_t* range = nullptr,
                       uint32_t* range_flag = nullptr);

struct BinScratch {
  uint32_t* matrix;      // counts of one pass: [view][chunk][tile] at chunks * tile_base
  uint32_t* tile_tot;    // [total tiles]
  uint2* rect_sorted;    // [V][n] tile rectangles in depth order
  uint64_t* run;         // device scalar: order slots used by the passes so far
  bool packed;           // rectangle format (rect_packed)
  uint32_t emit_band = 256;  // tiles per band of the emit kernel (a knob: results do not depend on it)
};
// One binning pass: the depth slice [pbeg, pend) of every view's sorted
// Gaussians.  Its entries go to order[ebase[v] + ranges[...]].
struct BinPass {
  uint32_t pbeg, pend;
  const uint8_t* done;  // per global tile, nonzero: saturated, gets no entries (may be null)
  uint32_t* ranges;     // [total tiles][2], view-local positions within the pass
  uint64_t* ebase;      // [V] first order slot of each view's entries in this pass
  uint64_t* entries;    // [V] entries per view in this pass
  bool first;           // first pass of the step: order slots start at 0
  // A filtered pass (launch_slice_filter): the kept Gaussians of the slice, in
  // depth order, rects[v][0 .. kept[v]) and vals[v][0 .. kept[v]) (view stride n),
  // read instead of the sorted arrays' [pbeg, pend).  Null: not filtered.
  const uint2* rects = nullptr;
  const uint32_t* vals = nullptr;
  const uint32_t* kept = nullptr;
};
// Before a later slice is binned: keeps, in depth order, the Gaussians of sorted
// positions [pbeg, pend) whose rectangle meets a tile that done leaves open,
// into rects / vals (view stride n), kept[v] of them per view.  Only for views
// of at most slice_filter_fits' tiles.
struct SliceFilter {
  uint16_t* sat;   // [total tiles]: summed-area table of the open tiles
  uint64_t* mask;  // [V][ceil(n / 64)]: kept bits
  uint32_t* piece; // [V][slice_filter_pieces(n)]
};
bool slice_filter_fits(uint32_t max_tiles);
size_t slice_filter_pieces(uint32_t n);
void launch_slice_filter(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t max_tiles, uint32_t n,
                         uint32_t pbeg, uint32_t pend, const uint2* rect_sorted, const uint32_t* sorted_vals,
                         bool packed, const uint8_t* done, const SliceFilter& f, uint2* rects, uint32_t* vals,
                         uint32_t* kept);
// Pass part 1: per-(chunk, tile) counts, tile totals, ranges, entry counts and
// entry bases.  host_views: the views on the host (grid sizing).
void launch_bin_count(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                      uint32_t n, uint32_t chunk, const BinScratch& bs, const BinPass& pass);
// Pass part 2: the order itself.  Writes beyond order_capacity are dropped and
// set *overflow (device) nonzero.
void launch_bin_emit(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                     uint32_t n, uint32_t chunk, const uint32_t* vals, const BinScratch& bs, const BinPass& pass,
                     uint32_t* order, uint64_t order_capacity, uint32_t* overflow);
size_t bin_matrix_elems(const ViewDesc* host_views, uint32_t num_views, uint32_t n, uint32_t chunk);

// Blending (gsr_raster.hip).  A tile's list may come in several passes (depth
// slices); pass p's entries of tile g are order[ebase[p][view] + ranges[p][2g ..
// 2g+1]], and the passes follow each other in depth order.
constexpr int kMaxPasses = 8;
// Words between the work counters of a persistent launch: one 128-byte line
// each.  Atomics on one line are serialised, and 5120 waves popping counters
// that shared a line took 34% of the forward waves' lifetime (forward 3.30 ->
// 2.65 ms, backward 4.14 -> 3.90 ms on the 32-view 1M-Gaussian step once apart).
constexpr uint32_t kQueueStride = 32;
struct RasterArgs {
  const ViewDesc* views;
  const TileWork* work;  // [num_work]
  uint32_t num_work;
  const uint32_t* work_count;  // optional: the number of items of `work`, on the device (replaces num_work)
  const float* splats;
  const uint32_t* order;
  uint64_t order_capacity;  // slots in order; segments ending beyond it are skipped
  int num_passes;
  const uint32_t* ranges[kMaxPasses];
  const uint64_t* ebase[kMaxPasses];  // null: 0 for every view
  float* image;          // concatenated views
  float* state;          // optional (T, last), single pass only
  const float* dl;       // backward: dLoss/dPixel, concatenated views
  // backward, optional: the targets (concatenated views); dLoss/dPixel is then
  // the L1 gradient sign(image - target) / (3 pixels) and dl is not read.
  const float* l1_target;
  float* grads;          // backward: [V][n][16]
  uint64_t grads_view_stride;  // elements between views' gradient records (0: shared)
  uint32_t grads_record;       // floats per gradient record (0: the reference's 16)
  uint32_t* queue;       // persistent mode: 9 work counters kQueueStride words apart (zeroed before launch)
  uint32_t tile_batch;   // tiles claimed per queue pop
  // Tile splitting (batched step), on when split != null: a tile with split[g]
  // set (g: global tile id) is blended by two waves, one per half (bands 0-1,
  // 2-3).  Items [0, heavy_count[0]) are those half-tile items, then come the
  // num_work tiles of `work` in order (any order of the tiles), the split ones
  // returning at once.  Both are written by launch_split_list.
  const uint8_t* split;          // [total tiles]
  const TileWork* heavy;         // [2 num_work]
  const uint32_t* heavy_count;
  // Forward over passes: pass blends segment `pass`, resuming from tstate (per
  // pixel transmittance) and image, and skips tiles flagged done (saturated by
  // earlier passes).  band_open[4 g + b] receives whether band b of tile g still
  // has a pixel with T > 1e-4 after this pass.
  int pass;
  float* tstate;         // null for a single pass
  const uint8_t* done;   // null for a single pass
  uint8_t* band_open;    // null for a single pass
  uint32_t* walk;        // forward, optional: per global tile, entries examined in this pass (max over items)
  uint32_t* last;        // forward, optional: per global tile, furthest last hit - segment start (max over items)
  // Nonzero: no hit-box culling.  Every entry is evaluated at every pixel of
  // its tile, as alphaAt does; results must not change (verification knob).
  int no_cull;
  // Backward, nonzero: the records receive the entry's pixel moments instead of
  // its gradient terms (record layout kMomentSlots); launch_shape with
  // moments != 0 turns them into gradients.
  int moments;
};
// Moment record of one (Gaussian, view): with dP = dLoss/dpower and (dx, dy) the
// pixel centre less the splat centre (u, v), summed over the pixels it blends:
//   0-2  dLoss/dcolour (r, g, b)      3  sum dP
//   4    sum dP dx                    5  sum dP dy
//   6    sum dP dx^2                  7  sum dP dx dy      8  sum dP dy^2
// The screen gradients are linear in them (gpu.js:266-281): d opLogit = (1 -
// op) m3, du = A m4 + B m5, dv = B m4 + C m5, dA = -m6 / 2, dB = -m7, dC = -m8 / 2.
constexpr int kMomentSlots = 9;
// done[g] = (no band of tile g open) with assign != 0 (after the first pass, which
// blends every tile), else done[g] |= (no band open).
void launch_done_update(hipStream_t s, uint32_t total_tiles, const uint8_t* band_open, uint8_t* done, int assign);
// Marks the tiles of a blend launch whose cost exceeds split_cost and lists
// their half-tile items (heavy_count zeroed before).  Cost of tile g: with walk (the
// backward), the entries the forward examined summed over num_passes
// (walk[p][g]); otherwise its entries over ranges[0 .. num_passes).
struct SplitArgs {
  const TileWork* tiles;  // [total_tiles], whole-tile items
  uint32_t total_tiles;
  uint32_t split_cost;
  const uint32_t* ranges[kMaxPasses];
  int num_passes;
  const uint32_t* walk;        // optional, [num_passes][total_tiles]
  uint8_t* split;
  TileWork* heavy;
  uint32_t* heavy_count;
};
void launch_split_list(hipStream_t s, const SplitArgs& a);
// The whole-tile items a batched blend launch needs, in tile order: tiles with
// entries in any of ranges[0 .. num_passes), not flagged done (optional), not
// split into half-tile items (optional).  A stable compaction of `tiles` into
// out[0 .. *out_count); counts: [ceil(total_tiles / 256)] scratch.
struct LiveArgs {
  const TileWork* tiles;  // [total_tiles]
  uint32_t total_tiles;
  const uint32_t* ranges[kMaxPasses];
  int num_passes;
  const uint8_t* done;   // optional
  const uint8_t* split;  // optional
  uint32_t* counts;
  TileWork* out;
  uint32_t* out_count;
};
void launch_live_list(hipStream_t s, const LiveArgs& a);
void launch_forward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid);
void launch_backward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid);
// Resident blocks per CU of the persistent forward and backward kernels.
void raster_occupancy(int* forward_per_cu, int* backward_per_cu);
// window 7: the reference's 7x7 box SSIM; 11: the 11x11 Gaussian (sigma 1.5).
void launch_loss_grad(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                      const float* image, const float* target, float lambda, float* coef, float* grad,
                      int window = 7);
// Device copy / zero fill of 4-byte aligned memory, bytes a multiple of 4.
void launch_copy(hipStream_t s, void* dst, const void* src, size_t bytes);
void launch_zero(hipStream_t s, void* p, size_t bytes);

// Chain rule to the 3D parameters (gsr_chain.hip); Adam and stats (gsr_optim.hip).
// grads: per-view 16-float records, view v at grads + v * grads_view_stride.
// shape / sh_grad are written with the sum over views; grads_out (optional)
// receives the summed colour (0-2) and opacity-logit (3) gradients.  With
// consume != 0 every per-view record that held a gradient is zeroed after use,
// which leaves the per-view accumulators clear for the next backward.
// touched (optional, [num_views]) counts the Gaussians with any gradient per view.
// record: floats per per-view gradient record (16, or 12 for the batch's packed
// records; the nine terms sit at 0-8 either way).  moments: the records hold
// pixel moments (RasterArgs::moments), mapped to gradients with the view's conic
// and the Gaussian's opacity.
// stats (optional, [n][2], with valid_mask from launch_project): STATS
// (gpu.js:483-497) for every view in turn, stats[i] += (|(du, dv)|, 1) where
// Gaussian i projects validly in the view; left unchanged when the device word
// stats_guard (optional) is nonzero.
void launch_shape(hipStream_t s, const gsr_camera* cams, uint32_t num_views, const float* gs, const float* sh,
                  int sh_degree, uint32_t n, float* grads, uint64_t grads_view_stride, float* shape,
                  float* sh_grad, float* grads_out, int consume, uint32_t* touched = nullptr, uint32_t record = 16,
                  const uint64_t* valid_mask = nullptr, float* stats = nullptr,
                  const uint32_t* stats_guard = nullptr, int moments = 0);
void launch_adam_sh(hipStream_t s, const gsr_step& step, float lr_sh_rest, float* gs, float* sh, const float* grads,
                    const float* shape, const float* sh_grad, float* moments);
void launch_adam(hipStream_t s, const gsr_step& step, float* gs, const float* grads, const float* shape,
                 float* moments);
void launch_stats(hipStream_t s, const gsr_step& step, const float* grads, const float* splats, float* stats);
// Fills the blend work list: one entry per tile of every view, view-major.
void launch_fill_work(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t total_tiles,
                      TileWork* work);

}  // namespace gsr
