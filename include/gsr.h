/*
 * gsr.h — C ABI of the MI355X (gfx950) differentiable Gaussian-splat rasterizer.
 *
 * The operator surface is the pipeline set of trainer/gpu.js (project -> bin/sort
 * -> render forward -> loss gradient -> render backward -> shape chain -> Adam),
 * with the same buffer roles and layouts, so a JavaScript trainer binds it through
 * a thin N-API addon (home-measure-privacy_amd/js/) and any other host through
 * plain C calls.  Citations are trainer/gpu.js:line in the reference.
 *
 * Layouts (all little-endian float32 / uint32, device memory):
 *   gs       [n][16]  0-2 mean, 3 pad, 4-6 logScale, 7 opacityLogit,
 *                     8-11 rotation wxyz, 12-14 colour, 15 pad   (gpu.js:5-6)
 *   sh       [n][16][3] SH coefficients (extension, degree <= 3): coefficient 0
 *                     is the colour (DC) in colour units, 1..15 the real SH
 *                     bands; colour = dc + sum_{l>=1} Y_lm(dir) f_lm.
 *   splats   [n][12]  u, v, conicA, conicB, conicC, opacity, radius, depth,
 *                     r, g, b, valid                            (gpu.js:7-8)
 *   order    [E]      Gaussian indices, tile-major then (depth, index)
 *   ranges   [T][2]   per 16x16 tile: [start, end) into order   (gpu.js:468)
 *   image    [H][W][3] RGB, black background                    (gpu.js:133)
 *   state    [H][W][2] final transmittance, one past last hit   (gpu.js:135-136)
 *   grads    [n][16]  0-2 colour, 3 opacityLogit, 4-8 screen u, v, conic
 *                     A, B, C, 9-15 unused                      (gpu.js:9-10)
 *   shape    [n][10]  mean 3, logScale 3, rotation 4            (gpu.js:11)
 *   moments  [n][28]  Adam m[14] then v[14]                     (gpu.js:431-434)
 *
 * Conventions: every function returns GSR_OK (0) or a negative status and then
 * records a message for gsr_last_error().  Work is issued in order on the
 * context's single HIP stream; functions that hand a value back to the host
 * (gsr_download, gsr_bin_sort's entry count, gsr_synchronize) wait for it.
 * A context is used from one host thread at a time.
 */
#ifndef GSR_H_
#define GSR_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSR_OK 0
#define GSR_E_INVALID (-1)  /* bad argument */
#define GSR_E_HIP (-2)      /* HIP runtime error */
#define GSR_E_NOMEM (-3)    /* device allocation failed */
#define GSR_E_CAPACITY (-4) /* an output buffer is too small; see the call */

/* 80-byte camera uniform: Camera gpu.js:16, cameraUniform gpu.js:696-701.
 * view = world->camera rows [R | t] (COLMAP axes: x right, y down, z forward). */
typedef struct gsr_camera {
  float view[12];
  float fx, fy, cx, cy;
  float width, height;
  uint32_t count;  /* Gaussians */
  uint32_t pixels; /* width * height */
} gsr_camera;

/* 32-byte step uniform of ADAM / STATS: Step gpu.js:412. */
typedef struct gsr_step {
  float t;
  uint32_t count;
  float lr_color, lr_opacity, lr_mean, lr_scale, lr_rotation, pad;
} gsr_step;

typedef struct gsr_ctx gsr_ctx;
typedef struct gsr_batch gsr_batch;

/* ---- context and buffers (createRenderer gpu.js:704-714, read 716-726, clear 773-777) */

/* Opens device `device` and creates the context's stream.  Replaces
 * createRenderer's adapter/device setup (gpu.js:705-712). */
int gsr_create(int device, gsr_ctx** out);
void gsr_destroy(gsr_ctx* ctx);
/* Message of the last failed call on ctx ("" if none).  ctx may be NULL for
 * failures of gsr_create. */
const char* gsr_last_error(const gsr_ctx* ctx);
/* Zero-filled device buffer of max(16, bytes) bytes, as buffer() (gpu.js:713-714). */
int gsr_alloc(gsr_ctx* ctx, size_t bytes, void** dptr);
int gsr_free(gsr_ctx* ctx, void* dptr);
/* Host -> device copy, as device.queue.writeBuffer. */
int gsr_upload(gsr_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device -> host copy; waits for all queued work first, as read() (gpu.js:716-726). */
int gsr_download(gsr_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Zero bytes of a device buffer, as clear() (gpu.js:773-777).  dptr and bytes
 * are multiples of 4, as WebGPU's clearBuffer requires. */
int gsr_clear(gsr_ctx* ctx, void* dptr, size_t bytes);
/* Device -> device copy of non-overlapping regions, as copyBufferToBuffer
 * (gpu.js:719); pointers and bytes are multiples of 4.  Queued on the stream. */
int gsr_copy(gsr_ctx* ctx, void* dst, const void* src, size_t bytes);
int gsr_synchronize(gsr_ctx* ctx);
/* Hit-box culling of gsr_forward / gsr_backward on ctx (default enabled).  No
 * counterpart in the reference: the blends skip entries and 4-row bands whose
 * conservative hit box (where alpha can reach 1/255) misses the pixels still
 * open, which alphaAt (gpu.js:91-103) would evaluate to a miss.  Disabled, every
 * entry is evaluated at every pixel of its tile; images, states and gradients
 * are the same (verification knob). */
int gsr_set_culling(gsr_ctx* ctx, int enabled);

/* ---- operators (one per pipeline of gpu.js:712) */

/* PREPROCESS (gpu.js:64-88): one splat record per Gaussian.  sh may be NULL
 * (sh_degree 0: colour copied from gs[12..14]); with sh_degree 1..3 the colour
 * is evaluated from sh toward the Gaussian from the camera centre.  Rejected
 * Gaussians (z < 0.2 or det <= 0) get valid = 0 and zeros elsewhere. */
int gsr_project(gsr_ctx* ctx, const gsr_camera* cam, const float* gs, const float* sh, int sh_degree,
                uint32_t n, float* splats);

/* Tile lists with the semantics of tileLists (gpu.js:450-481), replacing BIN +
 * SORT (gpu.js:616-685) without their 1024-entry cap: valid splats in (depth,
 * index) order appended to every tile their float64 bounding box covers.
 * order needs num_entries slots; when order_capacity is smaller the call
 * returns GSR_E_CAPACITY with *num_entries set and order untouched (ranges are
 * written).  ranges has 2 * ceil(W/16) * ceil(H/16) entries.  Waits for the
 * entry count. */
int gsr_bin_sort(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, uint32_t n, uint32_t* order,
                 uint64_t order_capacity, uint32_t* ranges, uint64_t* num_entries);

/* gsr_bin_sort that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  Same arguments, results and
 * errors as gsr_bin_sort.  Its scratch is a 48 MiB static device array per GPU
 * (held until the process ends), enough for about 1M Gaussians on a 1920x1080
 * view; a larger input returns GSR_E_CAPACITY before anything is launched.
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * (two contexts, or two threads) race. */
int gsr_bin_sort_noalloc(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, uint32_t n, uint32_t* order,
                         uint64_t order_capacity, uint32_t* ranges, uint64_t* num_entries);

/* FORWARD (gpu.js:105-137): front-to-back alpha blend, stop once T <= 1e-4.
 * state may be NULL (the backward recomputes transmittance and needs none). */
int gsr_forward(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, const uint32_t* order,
                const uint32_t* ranges, float* image, float* state);

/* SSIM_COEF + LOSS_GRAD (gpu.js:336-409, host lossGradient 742-753):
 * dLoss/dPixel of (1-lambda) L1 + lambda (1 - SSIM over 7x7 windows). */
int gsr_loss_grad(gsr_ctx* ctx, const gsr_camera* cam, const float* image, const float* target, float lambda,
                  float* loss_grad);

/* SSIM windows of the loss gradient: the reference's 7x7 box (gpu.js:342-372),
 * and an 11x11 Gaussian window with sigma 1.5 (extension; the window of the
 * usual 3D Gaussian splatting loss).  Both use the valid windows (wholly inside
 * the image), C1 = 1e-4, C2 = 9e-4, and the mean over windows and channels. */
#define GSR_SSIM_BOX7 7
#define GSR_SSIM_GAUSS11 11
/* gsr_loss_grad with the SSIM window chosen (GSR_SSIM_BOX7 or GSR_SSIM_GAUSS11). */
int gsr_loss_grad_window(gsr_ctx* ctx, const gsr_camera* cam, const float* image, const float* target,
                         float lambda, int window, float* loss_grad);

/* BACKWARD_SG (gpu.js:204-293) semantics, walked front to back with the
 * transmittance recomputed from the splats (no per-pixel state): adds the
 * colour, opacity-logit and screen (u, v, A, B, C) gradients into grads
 * (accumulates; clear it first for a fresh pass, as gpu.js:922-923 does).
 * image is the forward's output for the same inputs. */
int gsr_backward(gsr_ctx* ctx, const gsr_camera* cam, const float* splats, const uint32_t* order,
                 const uint32_t* ranges, const float* image, const float* loss_grad, float* grads);

/* SHAPE (gpu.js:297-328), computed analytically instead of by central
 * differences: shape[n][10] = d loss / d(mean, logScale, rotation) from
 * grads[4..8].  With sh_degree > 0, sh_grad[n][16][3] receives the SH
 * coefficient gradients of grads[0..2] and the view-direction term is added to
 * the mean gradient; sh_grad may be NULL for sh_degree 0.  Both are written. */
int gsr_shape(gsr_ctx* ctx, const gsr_camera* cam, const float* gs, const float* sh, int sh_degree, uint32_t n,
              const float* grads, float* shape, float* sh_grad);

/* ADAM (gpu.js:411-443): the 14 trainable numbers per Gaussian, colour clamp
 * and quaternion renormalisation; gs and moments updated in place. */
int gsr_adam(gsr_ctx* ctx, const gsr_step* step, float* gs, const float* grads, const float* shape,
             float* moments);

/* ADAM over the 59 numbers of an SH3 Gaussian (extension of gpu.js:411-443, for
 * SH training): mean (lr_mean), log-scale (lr_scale), rotation (lr_rotation),
 * opacity logit (lr_opacity, gradient grads[3]), SH DC (lr_color) and SH bands
 * 1-3 (lr_sh_rest), with the reference's update and quaternion renormalisation;
 * SH coefficients are not clamped.  gs, sh [n][16][3] and moments (120 n
 * floats, zero at the start: SH m [n][48] at 0 and v at 48 n; parameter k of
 * mean 0-2, logScale 3-5, rotation 6-9, opacity 10 has m at 96 n + k n + i and
 * v at 107 n + k n + i) are updated in place; sh_grad as written by gsr_shape /
 * gsr_batch_step.  gs, sh, sh_grad and moments are 16-byte aligned. */
int gsr_adam_sh(gsr_ctx* ctx, const gsr_step* step, float lr_sh_rest, float* gs, float* sh, const float* grads,
                const float* shape, const float* sh_grad, float* moments);

/* STATS (gpu.js:483-497): stats[2i] += |(du, dv)|, stats[2i+1] += 1 for valid splats. */
int gsr_stats(gsr_ctx* ctx, const gsr_step* step, const float* grads, const float* splats, float* stats);

/* ---- batched views (no counterpart in the reference, which renders one view
 * per step, gpu.js:915-933): many cameras of any resolution against one shared
 * Gaussian set in single launches per stage, with a per-view tile work queue. */

/* Creates a batch for num_views cameras and n Gaussians of SH degree sh_degree.
 * Device buffers are sized from the cameras; entry storage grows on demand. */
int gsr_batch_create(gsr_ctx* ctx, uint32_t num_views, const gsr_camera* cams, uint32_t n, int sh_degree,
                     gsr_batch** out);
void gsr_batch_destroy(gsr_batch* batch);
/* Replaces the cameras (same count and resolutions). */
int gsr_batch_set_cameras(gsr_batch* batch, const gsr_camera* cams);
/* One forward + backward over every view.  targets: the views' HWC images
 * concatenated in view order (sum of W*H*3 floats).  Outputs (summed over views,
 * written): grads [n][16] (colour 0-2 for sh_degree 0, opacityLogit 3), shape
 * [n][10], sh_grad [n][16][3] (sh_degree > 0, else may be NULL).  images may be
 * NULL or receive the rendered views, concatenated like targets.  With
 * use_graph != 0 the step is captured once into a HIP graph and replayed while
 * the pointer arguments stay the same. */
int gsr_batch_step(gsr_batch* batch, const float* gs, const float* sh, const float* targets, float lambda,
                   float* grads, float* shape, float* sh_grad, float* images, int use_graph);
/* gsr_batch_step that also accumulates the densify statistics of STATS
 * (gpu.js:483-497) for every view in turn: stats[n][2] += (|(du, dv)|, 1) with
 * the view's screen-position gradient, for each Gaussian that projects validly
 * in the view (the sum of one STATS pass per view).  stats may be NULL. */
int gsr_batch_step_stats(gsr_batch* batch, const float* gs, const float* sh, const float* targets, float lambda,
                         float* grads, float* shape, float* sh_grad, float* stats, float* images, int use_graph);
/* Per-view entry counts of the last step (host array of num_views). */
int gsr_batch_entries(gsr_batch* batch, uint64_t* entries);
/* Hit-box binning and culling of the batched step (default enabled).  Disabled,
 * each splat is binned to the reference square's tiles (tileLists, gpu.js:455-
 * 462) and the blends evaluate every entry at every pixel; results are the same
 * (verification knob; no counterpart in the reference). */
int gsr_batch_set_culling(gsr_batch* batch, int enabled);
/* The last step's tile lists of one view and depth slice (inspection; no
 * counterpart in the reference): order (device, order_capacity slots) receives
 * the slice's entries of the view, tile-major in (depth, index) order, and
 * ranges (device, 2 * tiles of the view) their [start, end) per tile, as
 * gsr_bin_sort writes them.  *num_entries is set; GSR_E_CAPACITY when order is
 * too small (ranges are written). */
int gsr_batch_tile_lists(gsr_batch* batch, uint32_t view, uint32_t pass, uint32_t* order, uint64_t order_capacity,
                         uint32_t* ranges, uint64_t* num_entries);
/* Tuning knobs: tiles claimed per queue pop (0 = keep) and Gaussians per
 * binning chunk (0 = keep).  gsr_batch_autotune times candidates on the given
 * inputs and keeps the fastest; it writes the choice to *tile_batch / *chunk.
 * It also picks the depth slices of gsr_batch_set_passes (2 x 8, 3 x 4 or
 * 4 x 4, for n >= 65536), the emit band and the split thresholds of
 * gsr_batch_set_split.
 * Results do not depend on any of these. */
int gsr_batch_configure(gsr_batch* batch, uint32_t tile_batch, uint32_t chunk);
/* Walk lengths above which a tile is blended as two half-tile items: in the
 * forward's first depth slice (by the previous step's walk of the tile), in its
 * later slices (by list length) and in the backward (by the entries the forward
 * walked).  0 = never; defaults 0, 0 and 1024.  gsr_batch_autotune picks all
 * three.  Results do not depend on them. */
int gsr_batch_set_split(gsr_batch* batch, uint32_t forward_first, uint32_t forward_later, uint32_t backward);
/* Tiles per band of the tile-order emission (64 to 4096, default 128):
 * gsr_batch_autotune picks 128 or 256.  Results do not depend on it. */
int gsr_batch_set_emit_band(gsr_batch* batch, uint32_t tiles);
/* The knobs in effect (after gsr_batch_autotune, say), so a later run can set
 * them directly: tuning[6] = tile_batch, chunk, the three split thresholds in
 * gsr_batch_set_split's order, and the emit band. */
int gsr_batch_tuning(gsr_batch* batch, uint32_t* tuning);
/* SSIM window of the batched step's loss (default GSR_SSIM_BOX7). */
int gsr_batch_set_ssim_window(gsr_batch* batch, int window);
/* Depth-slice passes of the batched step (1 to 8; default 2 with growth 8 for
 * n >= 65536, else 1).  With k > 1 passes the depth-sorted Gaussians are binned
 * in slices whose ends grow by `growth` (2 to 16; the first slice holds
 * ceil(n / growth^(k-1))), and a slice is binned only for tiles whose pixels
 * the slices before it left unsaturated (T > 1e-4).  Images and gradients are
 * those of the full tile lists: a saturated tile's blend stops before any later
 * entry.  growth is ignored for passes == 1. */
int gsr_batch_set_passes(gsr_batch* batch, uint32_t passes, uint32_t growth);
/* The depth slices in effect (growth is 1 for one pass). */
int gsr_batch_passes(gsr_batch* batch, uint32_t* passes, uint32_t* growth);
/* Runs one step without the graph, timing each stage with HIP events on the
 * context's stream, then a second step that collects the counters.  stage_ms[GSR_STAGES] receives milliseconds for: project,
 * depth sort, bin count, bin emit, forward, loss gradient, backward, chain.
 * counters[4 * num_views] receives per view: entries E, saturated-prefix entries
 * F (sum over tiles of the furthest pixel's last hit minus the tile start),
 * touched Gaussians G (any gradient in that view) and walked entries W (entries
 * the forward examined before its tiles saturated or ended). */
#define GSR_STAGES 8
int gsr_batch_profile(gsr_batch* batch, const float* gs, const float* sh, const float* targets, float lambda,
                      float* grads, float* shape, float* sh_grad, float* stage_ms, uint64_t* counters);
int gsr_batch_autotune(gsr_batch* batch, const float* gs, const float* sh, const float* targets, float lambda,
                       float* grads, float* shape, float* sh_grad, uint32_t* tile_batch, uint32_t* chunk);

#ifdef __cplusplus
}
#endif

#endif /* GSR_H_ */
