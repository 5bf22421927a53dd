// ADAM (trainer/gpu.js:411-443), its SH3 extension, and STATS (gpu.js:483-497).
// Built without fused multiply-adds, so the updates follow the reference's
// float operation order (the tests compare them with a float restatement).
#include "gsr_internal.h"

namespace gsr {

namespace {

// ADAM.main (gpu.js:422-443): SLOT maps the 14 trainable numbers (gpu.js:420).
__global__ __launch_bounds__(256) void adam_kernel(gsr_step st, float* __restrict__ gs,
                                                   const float* __restrict__ grads, const float* __restrict__ shape,
                                                   float* __restrict__ moments) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.count) return;
  const int slot[14] = {12, 13, 14, 7, 0, 1, 2, 4, 5, 6, 8, 9, 10, 11};
  const float bc1 = 1.0f - powf(0.9f, st.t);
  const float bc2 = 1.0f - powf(0.999f, st.t);
  float* b = gs + size_t(i) * 16;
  float* m = moments + size_t(i) * 28;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    float g = k < 4 ? grads[size_t(i) * 16 + k] : shape[size_t(i) * 10 + k - 4];
    float lr = st.lr_rotation;
    if (k < 3) lr = st.lr_color;
    else if (k == 3) lr = st.lr_opacity;
    else if (k < 7) lr = st.lr_mean;
    else if (k < 10) lr = st.lr_scale;
    float mk = 0.9f * m[k] + 0.1f * g;
    float vk = 0.999f * m[14 + k] + 0.001f * g * g;
    m[k] = mk;
    m[14 + k] = vk;
    float update = lr * (mk / bc1) / (sqrtf(vk / bc2) + 1e-8f);
    b[slot[k]] -= update;
  }
#pragma unroll
  for (int c = 12; c < 15; ++c) b[c] = fminf(1.0f, fmaxf(0.0f, b[c]));
  float qv[4] = {b[8], b[9], b[10], b[11]};
  float len = fmaxf(sqrtf(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2] + qv[3] * qv[3]), 1e-12f);
#pragma unroll
  for (int k = 0; k < 4; ++k) b[8 + k] = qv[k] / len;
}

// ADAM over the 59 numbers of an SH3 Gaussian (extension of gpu.js:411-443): the
// reference's update (beta 0.9 / 0.999, eps 1e-8, pow bias correction with t as
// f32) for mean, log-scale, rotation, opacity logit and the 48 SH coefficients,
// then the quaternion renormalisation.  The SH coefficients are not clamped (the
// colour clamp of gpu.js:439 applies to the SH0 colour only).  Moments [n][120]:
// m of parameter k at k, v at 60 + k, parameters ordered mean 0-2, logScale 3-5,
// rotation 6-9, opacity logit 10, SH 11-58 ([16][3]).
__device__ inline float adam_update(float g, float lr, float bc1, float bc2, float& m, float& v) {
  m = 0.9f * m + 0.1f * g;
  v = 0.999f * v + 0.001f * g * g;
  return lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// Adam over the 59 numbers of an SH3 Gaussian in two coalesced sweeps.  Moments
// (120 n floats): SH m [n][48] at 0, SH v [n][48] at 48 n, then the 11 base
// parameters (mean 0-2, log-scale 3-5, rotation 6-9, opacity logit 10) m at
// 96 n + k n + i and v at 107 n + k n + i; the last 2 n floats are unused.
__global__ __launch_bounds__(256) void adam_base_kernel(gsr_step st, float* __restrict__ gs,
                                                        const float* __restrict__ grads,
                                                        const float* __restrict__ shape,
                                                        float* __restrict__ moments) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = st.count;
  if (i >= n) return;
  const float bc1 = 1.0f - powf(0.9f, st.t);
  const float bc2 = 1.0f - powf(0.999f, st.t);
  float4* rec = reinterpret_cast<float4*>(gs) + size_t(i) * 4;
  float4 a = rec[0], b = rec[1], c = rec[2];
  float* mb = moments + size_t(96) * n + i;
  float* vb = moments + size_t(107) * n + i;
  const float* shp = shape + size_t(i) * 10;
  float* p[11] = {&a.x, &a.y, &a.z, &b.x, &b.y, &b.z, &c.x, &c.y, &c.z, &c.w, &b.w};
#pragma unroll
  for (int k = 0; k < 11; ++k) {
    const float g = k < 10 ? shp[k] : grads[size_t(i) * 16 + 3];
    const float lr = k < 3 ? st.lr_mean : k < 6 ? st.lr_scale : k < 10 ? st.lr_rotation : st.lr_opacity;
    float m = mb[size_t(k) * n], v = vb[size_t(k) * n];
    *p[k] -= adam_update(g, lr, bc1, bc2, m, v);
    mb[size_t(k) * n] = m;
    vb[size_t(k) * n] = v;
  }
  const float len = fmaxf(sqrtf(c.x * c.x + c.y * c.y + c.z * c.z + c.w * c.w), 1e-12f);
  c = make_float4(c.x / len, c.y / len, c.z / len, c.w / len);
  rec[0] = a;
  rec[1] = b;
  rec[2] = c;
}

// SH coefficients: element e of sh [n][48] is coefficient (e mod 48) / 3; the DC
// triple uses lr_color, bands 1-3 lr_sh_rest.  Four elements per thread.
__global__ __launch_bounds__(256) void adam_sh_coef_kernel(gsr_step st, float lr_sh_rest, float* __restrict__ sh,
                                                           const float* __restrict__ sh_grad,
                                                           float* __restrict__ moments) {
  const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t quads = size_t(st.count) * 12;
  if (j >= quads) return;
  const float bc1 = 1.0f - powf(0.9f, st.t);
  const float bc2 = 1.0f - powf(0.999f, st.t);
  float4* s4 = reinterpret_cast<float4*>(sh) + j;
  float4* m4 = reinterpret_cast<float4*>(moments) + j;
  float4* v4 = reinterpret_cast<float4*>(moments) + quads + j;
  float4 x = *s4, m = *m4, v = *v4;
  const float4 g = reinterpret_cast<const float4*>(sh_grad)[j];
  const uint32_t e0 = uint32_t(j % 12) * 4;  // element of the first lane within its Gaussian's 48
  float* xs[4] = {&x.x, &x.y, &x.z, &x.w};
  float* ms[4] = {&m.x, &m.y, &m.z, &m.w};
  float* vs[4] = {&v.x, &v.y, &v.z, &v.w};
  const float gg[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float lr = e0 + c < 3 ? st.lr_color : lr_sh_rest;
    *xs[c] -= adam_update(gg[c], lr, bc1, bc2, *ms[c], *vs[c]);
  }
  *s4 = x;
  *m4 = m;
  *v4 = v;
}

// STATS.main (gpu.js:491-497).
__global__ __launch_bounds__(256) void stats_kernel(uint32_t n, const float* __restrict__ grads,
                                                    const float* __restrict__ splats, float* __restrict__ st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || splats[size_t(i) * 12 + 11] < 0.5f) return;
  float gu = grads[size_t(i) * 16 + 4], gv = grads[size_t(i) * 16 + 5];
  st[size_t(i) * 2] += sqrtf(gu * gu + gv * gv);
  st[size_t(i) * 2 + 1] += 1.0f;
}

}  // namespace

void launch_adam(hipStream_t s, const gsr_step& step, float* gs, const float* grads, const float* shape,
                 float* moments) {
  if (step.count == 0) return;
  hipLaunchKernelGGL(adam_kernel, dim3((step.count + 255) / 256), dim3(256), 0, s, step, gs, grads, shape, moments);
}

void launch_adam_sh(hipStream_t s, const gsr_step& step, float lr_sh_rest, float* gs, float* sh, const float* grads,
                    const float* shape, const float* sh_grad, float* moments) {
  if (step.count == 0) return;
  hipLaunchKernelGGL(adam_base_kernel, dim3((step.count + 255) / 256), dim3(256), 0, s, step, gs, grads, shape,
                     moments);
  const size_t quads = size_t(step.count) * 12;
  hipLaunchKernelGGL(adam_sh_coef_kernel, dim3(uint32_t((quads + 255) / 256)), dim3(256), 0, s, step, lr_sh_rest, sh,
                     sh_grad, moments);
}

void launch_stats(hipStream_t s, const gsr_step& step, const float* grads, const float* splats, float* stats) {
  if (step.count == 0) return;
  hipLaunchKernelGGL(stats_kernel, dim3((step.count + 255) / 256), dim3(256), 0, s, step.count, grads, splats,
                     stats);
}

}  // namespace gsr
