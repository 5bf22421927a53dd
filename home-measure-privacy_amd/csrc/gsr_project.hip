// Per-Gaussian projection: PREPROCESS (trainer/gpu.js:64-88) with projectShape
// (gpu.js:30-60), plus SH colour (extension) and the tile-binning inputs of
// tileLists (gpu.js:450-481): depth key, float64 tile rectangle, tile count.
// The single-view binning (prep_kernel) uses the reference's square rectangle;
// the batched step's projection uses the tighter hit-box rectangle.
//
// One thread per Gaussian reads its 64-byte record (and 192 bytes of SH) once as
// float4s and loops over every view of the launch, so a batch of V views reads
// the parameters once instead of V times.
#include "gsr_internal.h"
#include "gsr_math.h"

namespace gsr {

namespace {

struct GaussLocal {
  float mean[3];
  float S[3][3];  // world covariance (R diag(s))(R diag(s))^T
  float colour[3];
  float opacity;
  float sh[48];
};

__device__ inline void load_gaussian(const float* __restrict__ gs, const float* __restrict__ sh, int deg, uint32_t i,
                                     GaussLocal& g) {
  const float4* g4 = reinterpret_cast<const float4*>(gs) + size_t(i) * 4;
  float4 a = g4[0], b = g4[1], c = g4[2], d = g4[3];
  g.mean[0] = a.x; g.mean[1] = a.y; g.mean[2] = a.z;
  float q[10] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z, c.w};
  world_cov(q, g.S);
  g.colour[0] = d.x; g.colour[1] = d.y; g.colour[2] = d.z;
  // sigmoid, gpu.js:85
  g.opacity = 1.0f / (1.0f + expf(-b.w));
  if (deg > 0) {
    const float4* s4 = reinterpret_cast<const float4*>(sh) + size_t(i) * 12;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      float4 v = s4[k];
      g.sh[4 * k] = v.x; g.sh[4 * k + 1] = v.y; g.sh[4 * k + 2] = v.z; g.sh[4 * k + 3] = v.w;
    }
  }
}

__device__ inline void write_splat(float* __restrict__ splats, size_t idx, const Projected& p, float opacity,
                                   const float rgb[3]) {
  float4* o = reinterpret_cast<float4*>(splats + idx * 12);
  if (!p.ok) {
    o[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    o[1] = make_float4(0.f, 0.f, 0.f, 0.f);
    o[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  o[0] = make_float4(p.u, p.v, p.a, p.b);
  o[1] = make_float4(p.c, opacity, p.radius, p.depth);
  o[2] = make_float4(rgb[0], rgb[1], rgb[2], 1.0f);
}

// Depth key and tile rectangle of half-widths (rx, ry) around (u, v); splats
// without tiles get the invalid key.
__device__ inline void write_bin_inputs(uint32_t* keys, uint2* rects, bool packed, uint32_t* counts, size_t idx,
                                        bool valid,
                                        float u, float v, float rx, float ry, float depth, uint32_t tiles_x,
                                        uint32_t tiles_y) {
  int x0 = 0, x1 = -1, y0 = 0, y1 = -1;
  bool any = valid && tile_rect(u, v, rx, ry, tiles_x, tiles_y, x0, x1, y0, y1);
  uint32_t cnt = any ? uint32_t(x1 - x0 + 1) * uint32_t(y1 - y0 + 1) : 0u;
  if (keys) keys[idx] = any ? __float_as_uint(depth) : kInvalidKey;
  if (rects && packed)
    reinterpret_cast<uint32_t*>(rects)[idx] =
        any ? uint32_t(x0) | (uint32_t(x1) << 8) | (uint32_t(y0) << 16) | (uint32_t(y1) << 24) : 1u;
  else if (rects)
    rects[idx] = any ? make_uint2(uint32_t(x0) | (uint32_t(x1) << 16), uint32_t(y0) | (uint32_t(y1) << 16))
                     : make_uint2(1u, 0u);
  if (counts) counts[idx] = cnt;
}

__global__ __launch_bounds__(256) void project_kernel(const gsr_camera* __restrict__ cams,
                                                      const ViewDesc* __restrict__ views, uint32_t num_views,
                                                      const float* __restrict__ gs, const float* __restrict__ sh,
                                                      int deg, uint32_t n, float* __restrict__ splats,
                                                      uint32_t* __restrict__ keys, uint2* __restrict__ rects,
                                                      uint32_t* __restrict__ counts, bool packed, bool hitbox,
                                                      uint64_t* __restrict__ valid_mask) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t mask_words = valid_mask_words(n);
  GaussLocal g;
  load_gaussian(gs, sh, deg, i, g);
  for (uint32_t v = 0; v < num_views; ++v) {
    CamF cam = load_cam(cams[v]);
    Projected p = project_cov(cam, g.mean, g.S);
    float rgb[3] = {g.colour[0], g.colour[1], g.colour[2]};
    if (deg > 0 && p.ok) sh_colour(deg, g.sh, g.mean, cam, rgb);
    size_t idx = size_t(v) * n + i;
    write_splat(splats, idx, p, g.opacity, rgb);
    if (valid_mask) {
      // Waves cover 64 consecutive Gaussians from a multiple of 64 (256-thread
      // blocks), so one ballot is one mask word; lanes past n are inactive.
      const uint64_t m = __ballot(p.ok);
      if (__lane_id() == 0) valid_mask[size_t(v) * mask_words + i / 64] = m;
    }
    if (keys || rects || counts) {
      // The batched step bins each splat to the tiles of its hit box (where
      // alpha can reach 1/255, padded by half a pixel for the float32 pixel
      // offsets) within the reference square: the tiles it leaves out hold no
      // pixel the splat can hit, so renders and gradients are unchanged.
      // Without hitbox: the reference square, as tileLists bins.
      float rx = p.radius, ry = p.radius;
      if (hitbox && p.ok && fabsf(p.u) < 0x1p20f && fabsf(p.v) < 0x1p20f) {
        const float2 e = hit_extent(p.a, p.b, p.c, g.opacity, p.radius);
        rx = e.x < 0.0f ? e.x : fminf(p.radius, e.x + 0.5f);
        ry = e.y < 0.0f ? e.y : fminf(p.radius, e.y + 0.5f);
      }
      write_bin_inputs(keys, rects, packed, counts, idx, p.ok, p.u, p.v, rx, ry, p.depth, views[v].tiles_x,
                       views[v].tiles_y);
    }
  }
}

// Binning inputs from splat records written elsewhere: valid = record[11] > 0.5
// (gpu.js:454), u, v, radius, depth from the record.
__global__ __launch_bounds__(256) void prep_kernel(const ViewDesc* __restrict__ views, uint32_t num_views,
                                                   const float* __restrict__ splats, uint32_t n,
                                                   uint32_t* __restrict__ keys, uint2* __restrict__ rects,
                                                   bool packed) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t v = blockIdx.y;
  if (i >= n || v >= num_views) return;
  size_t idx = size_t(v) * n + i;
  const float4* r4 = reinterpret_cast<const float4*>(splats + idx * 12);
  float4 a = r4[0], b = r4[1], c = r4[2];
  bool valid = c.w > 0.5f;
  write_bin_inputs(keys, rects, packed, nullptr, idx, valid, a.x, a.y, b.z, b.z, b.w, views[v].tiles_x,
                   views[v].tiles_y);
}

}  // namespace

void launch_project(hipStream_t s, const gsr_camera* cams, const ViewDesc* views, uint32_t num_views,
                    const float* gs, const float* sh, int sh_degree, uint32_t n, float* splats, uint32_t* keys,
                    uint2* rects, uint32_t* counts, bool packed, bool hitbox, uint64_t* valid_mask) {
  if (n == 0 || num_views == 0) return;
  dim3 grid((n + 255) / 256);
  hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, s, cams, views, num_views, gs, sh, sh_degree, n, splats,
                     keys, rects, counts, packed, hitbox, valid_mask);
}

void launch_prep(hipStream_t s, const ViewDesc* views, uint32_t num_views, const float* splats, uint32_t n,
                 uint32_t* keys, uint2* rects, bool packed) {
  if (n == 0 || num_views == 0) return;
  dim3 grid((n + 255) / 256, num_views);
  hipLaunchKernelGGL(prep_kernel, grid, dim3(256), 0, s, views, num_views, splats, n, keys, rects, packed);
}

}  // namespace gsr
