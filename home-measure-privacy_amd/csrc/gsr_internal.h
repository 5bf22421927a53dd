// Internal declarations shared by the HIP translation units of libgsr.
//
// Device code follows the WGSL of trainer/gpu.js (cited per function).  All HIP
// sources are built with -ffp-contract=off: the backward pass replays the
// forward's blend arithmetic bit for bit, so every fused multiply-add is written
// out explicitly and identically in both.
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
  uint64_t entry_base;  // first slot of this view in the concatenated order
};

// One unit of blend work: a 16x16 tile of one view.
struct TileWork {
  uint32_t view;
  uint32_t xy;  // tx | ty << 16
};

// ---- launch helpers (host side), implemented per translation unit -------------

// Projection (gsr_project.hip).  cams: device array of num_views cameras.
// splats [V][n][12]; keys/rects/counts optional [V][n] (binning inputs).
void launch_project(hipStream_t s, const gsr_camera* cams, const ViewDesc* views, uint32_t num_views,
                    const float* gs, const float* sh, int sh_degree, uint32_t n, float* splats, uint32_t* keys,
                    uint2* rects, uint32_t* counts);
// Binning inputs from existing splat records (drop-in path).
void launch_prep(hipStream_t s, const ViewDesc* views, uint32_t num_views, const float* splats, uint32_t n,
                 uint32_t* keys, uint2* rects);

// Sorting and binning (gsr_sort.hip).
struct SortScratch {
  uint32_t* keys_alt;   // [V][n]
  uint32_t* vals;       // [V][n]
  uint32_t* vals_alt;   // [V][n]
  uint32_t* hist;       // [V][256][chunks]
  uint32_t* piece;      // scan partials
};
size_t radix_hist_elems(uint32_t n, uint32_t num_views);
size_t scan_piece_elems(size_t seg_len, uint32_t num_segs);
// Stable sort of (keys[v][i], i) per view by key; result indices in vals (and
// keys sorted in place).  Four 8-bit passes.
void launch_depth_sort(hipStream_t s, uint32_t* keys, uint32_t n, uint32_t num_views, const SortScratch& sc);

struct BinScratch {
  uint32_t* matrix;      // [sum_v chunks_v * T_v]
  uint64_t* matrix_off;  // [V] offset of each view's matrix
  uint32_t* tile_tot;    // [total tiles]
  uint64_t* entries;     // [V] entry count per view
};
// Pass 1 of binning: per-(chunk, tile) counts, tile totals, ranges (view-local
// entry positions), entry counts and views[v].entry_base.  keys/vals: the
// depth-sorted (key, Gaussian) pairs of each view.
void launch_bin_count(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t n, uint32_t chunk,
                      const uint32_t* keys, const uint32_t* vals, const uint2* rects, const BinScratch& bs,
                      uint32_t* ranges, uint32_t max_tiles_per_view);
// Pass 2: the order itself.  Writes beyond order_capacity are dropped and set
// *overflow (device) nonzero.
void launch_bin_emit(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t n, uint32_t chunk,
                     const uint32_t* keys, const uint32_t* vals, const uint2* rects, const BinScratch& bs,
                     const uint32_t* ranges, uint32_t* order, uint64_t order_capacity, uint32_t* overflow,
                     uint32_t max_tiles_per_view);
size_t bin_matrix_elems(const ViewDesc* host_views, uint32_t num_views, uint32_t n, uint32_t chunk);

// Blending (gsr_raster.hip).
struct RasterArgs {
  const ViewDesc* views;
  const TileWork* work;  // [num_work]
  uint32_t num_work;
  const float* splats;
  const uint32_t* order;
  uint64_t order_capacity;  // slots in order; tiles whose range ends beyond it are skipped
  const uint32_t* ranges;
  float* image;          // concatenated views
  float* state;          // optional
  const float* dl;       // backward: dLoss/dPixel, concatenated views
  float* grads;          // backward: [V][n][16]
  uint64_t grads_view_stride;  // elements between views' gradient records (0: shared)
  uint32_t* queue;       // persistent mode: work counter (zeroed before launch)
  uint32_t tile_batch;   // tiles claimed per queue pop
  uint32_t* tile_last;   // forward, optional: per global tile, furthest last hit - start
};
void launch_forward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid);
void launch_backward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid);
// Resident blocks per CU of the persistent forward and backward kernels.
void raster_occupancy(int* forward_per_cu, int* backward_per_cu);
void launch_loss_grad(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                      const float* image, const float* target, float lambda, float* coef, float* grad);
void launch_zero(hipStream_t s, void* p, size_t bytes);

// Chain rule to the 3D parameters, Adam, stats (gsr_chain.hip).
// grads: per-view 16-float records, view v at grads + v * grads_view_stride.
// shape / sh_grad are written with the sum over views; grads_out (optional)
// receives the summed colour (0-2) and opacity-logit (3) gradients.  With
// consume != 0 every per-view record that held a gradient is zeroed after use,
// which leaves the per-view accumulators clear for the next backward.
// touched (optional, [num_views]) counts the Gaussians with any gradient per view.
void launch_shape(hipStream_t s, const gsr_camera* cams, uint32_t num_views, const float* gs, const float* sh,
                  int sh_degree, uint32_t n, float* grads, uint64_t grads_view_stride, float* shape,
                  float* sh_grad, float* grads_out, int consume, uint32_t* touched = nullptr);
void launch_adam(hipStream_t s, const gsr_step& step, float* gs, const float* grads, const float* shape,
                 float* moments);
void launch_stats(hipStream_t s, const gsr_step& step, const float* grads, const float* splats, float* stats);
// Fills the blend work list: one entry per tile of every view, view-major.
void launch_fill_work(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t total_tiles,
                      TileWork* work);

}  // namespace gsr
