// From screen-space gradients to the 3D parameters: the analytic derivative of
// projectShape (trainer/gpu.js:30-60), replacing the central differences of
// SHAPE (gpu.js:297-328), plus the SH colour chain (extension).  Built with
// fused multiply-adds (its results are checked against the FP64 oracle); the
// optimizer kernels, which follow the reference's op order, are in gsr_optim.hip.
//
// One thread per Gaussian loops over the views of a launch.  The covariance
// gradient is linear in the per-view terms, so the views' world-space covariance
// gradients are summed first and pushed through M = R(q) diag(exp(s)) once.
#include "gsr_internal.h"
#include "gsr_math.h"

namespace gsr {

namespace {

// Three floats at 4-byte alignment (the batch's 9-float moment records lie 36
// bytes apart): one 12-byte access.
struct __attribute__((aligned(4))) Float3u {
  float x, y, z;
};

// The last link for Gaussian i, after its views: dL/dM = 2 Gw M for Sigma = M M^T,
// M = R diag(s), to log-scales and the quaternion; writes shape[i] (mean, log
// scale, quaternion), sh_grad[i] (degree > 0) and the colour / opacity part of
// grads_out[i].
__device__ inline void finish_gaussian(uint32_t i, const float q[10], const float (&Gw)[3][3], const float gmean[3],
                                       const float gsh[48], const float gcol[3], float gop, int deg,
                                       float* __restrict__ shape, float* __restrict__ sh_grad,
                                       float* __restrict__ grads_out) {
  const float s[3] = {expf(q[3]), expf(q[4]), expf(q[5])};
  float R[3][3];
  const float qlen = quat_to_mat(q + 6, R);
  float gR[3][3], gls[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    float acc = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float gM = 2.0f * (Gw[a][0] * R[0][b] * s[b] + Gw[a][1] * R[1][b] * s[b] + Gw[a][2] * R[2][b] * s[b]);
      gR[a][b] = gM * s[b];
      acc += gM * R[a][b];
    }
    gls[b] = acc * s[b];
  }
  const float w = q[6] / qlen, qx = q[7] / qlen, qy = q[8] / qlen, qz = q[9] / qlen;
  float gn[4];
  gn[0] = 2.0f * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] + qx * gR[2][1]);
  gn[1] = 2.0f * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - 2.0f * qx * gR[1][1] - w * gR[1][2] +
                  qz * gR[2][0] + w * gR[2][1] - 2.0f * qx * gR[2][2]);
  gn[2] = 2.0f * (-2.0f * qy * gR[0][0] + qx * gR[0][1] + w * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] -
                  w * gR[2][0] + qz * gR[2][1] - 2.0f * qy * gR[2][2]);
  gn[3] = 2.0f * (-2.0f * qz * gR[0][0] - w * gR[0][1] + qx * gR[0][2] + w * gR[1][0] - 2.0f * qz * gR[1][1] +
                  qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
  const float nq[4] = {w, qx, qy, qz};
  const float dotn = nq[0] * gn[0] + nq[1] * gn[1] + nq[2] * gn[2] + nq[3] * gn[3];
  const float rawlen = sqrtf(q[6] * q[6] + q[7] * q[7] + q[8] * q[8] + q[9] * q[9]);
  float gq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) gq[k] = (rawlen > 1e-12f ? gn[k] - nq[k] * dotn : gn[k]) / qlen;
  float* o = shape + size_t(i) * 10;
  o[0] = gmean[0]; o[1] = gmean[1]; o[2] = gmean[2];
  o[3] = gls[0]; o[4] = gls[1]; o[5] = gls[2];
  o[6] = gq[0]; o[7] = gq[1]; o[8] = gq[2]; o[9] = gq[3];
  if (deg > 0 && sh_grad) {
    float4* so = reinterpret_cast<float4*>(sh_grad + size_t(i) * 48);
#pragma unroll
    for (int k = 0; k < 12; ++k) so[k] = make_float4(gsh[4 * k], gsh[4 * k + 1], gsh[4 * k + 2], gsh[4 * k + 3]);
  }
  if (grads_out) {
    float4* go = reinterpret_cast<float4*>(grads_out + size_t(i) * 16);
    go[0] = make_float4(gcol[0], gcol[1], gcol[2], gop);
  }
}

template <bool kMoments>
__global__ __launch_bounds__(256) void shape_kernel(const gsr_camera* __restrict__ cams, uint32_t num_views,
                                                    const float* __restrict__ gs, const float* __restrict__ sh,
                                                    int deg, uint32_t n, float* __restrict__ grads,
                                                    uint64_t stride, float* __restrict__ shape,
                                                    float* __restrict__ sh_grad, float* __restrict__ grads_out,
                                                    int consume, uint32_t* __restrict__ touched, uint32_t record,
                                                    const uint64_t* __restrict__ valid_mask,
                                                    float* __restrict__ stats,
                                                    const uint32_t* __restrict__ stats_guard) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // A step whose entries overflowed their storage is run again: its statistics
  // are left out, so they are added once.
  if (stats && stats_guard && *stats_guard) stats = nullptr;
  // STATS (gpu.js:491-497) per view, in view order: the running sums start from
  // the stored ones, so V views add exactly as V single-view STATS passes would.
  float st_norm = 0.0f, st_count = 0.0f;
  if (stats) {
    st_norm = stats[size_t(i) * 2];
    st_count = stats[size_t(i) * 2 + 1];
  }
  const size_t mask_words = valid_mask_words(n);
  const float4* g4 = reinterpret_cast<const float4*>(gs) + size_t(i) * 4;
  const float4 pa = g4[0], pb = g4[1], pc = g4[2];
  const float mean[3] = {pa.x, pa.y, pa.z};
  const float q[10] = {pa.x, pa.y, pa.z, pb.x, pb.y, pb.z, pc.x, pc.y, pc.z, pc.w};
  float S[3][3];
  world_cov(q, S);
  float shc[48];
  float gsh[48];
  if (deg > 0) {
    const float4* s4 = reinterpret_cast<const float4*>(sh) + size_t(i) * 12;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      float4 v = s4[k];
      shc[4 * k] = v.x; shc[4 * k + 1] = v.y; shc[4 * k + 2] = v.z; shc[4 * k + 3] = v.w;
    }
#pragma unroll
    for (int k = 0; k < 48; ++k) gsh[k] = 0.0f;
  }
  float Gw[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  float gmean[3] = {0.f, 0.f, 0.f};
  float gcol[3] = {0.f, 0.f, 0.f};
  float gop = 0.0f;
  // Moment records: d opLogit = (1 - op) m3, op the sigmoid of gpu.js:85 as
  // the projection computes it.
  const float op_scale = kMoments ? 1.0f - 1.0f / (1.0f + expf(-pb.w)) : 1.0f;
  // Most (Gaussian, view) records are zero, so the walk over views is bound by
  // load latency: the records of kAhead views are loaded before any is used.
  constexpr uint32_t kAhead = 4;
  // STATS (gpu.js:491-497) of one view: |(du, dv)| where the Gaussian projects
  // validly (valid_mask), in view order.
  auto add_stats = [&](bool valid, float gu, float gv) __attribute__((always_inline)) {
    if (!valid) return;
    st_norm += sqrtf(gu * gu + gv * gv);
    st_count += 1.0f;
  };
  // A zero (du, dv): the norm adds +0, which leaves the sum as it is.
  auto add_zero_stats = [&](bool valid) __attribute__((always_inline)) { st_count += valid ? 1.0f : 0.0f; };
  auto view_step = [&](uint32_t v, const float (&gr)[9], bool valid) __attribute__((always_inline)) {
    float* rp = grads + size_t(v) * stride + size_t(i) * record;
    gcol[0] += gr[0]; gcol[1] += gr[1]; gcol[2] += gr[2];
    if (kMoments) gop = fmaf(op_scale, gr[3], gop);
    else gop += gr[3];
    if (!kMoments) add_stats(valid, gr[4], gr[5]);
    bool screen = gr[4] != 0.0f || gr[5] != 0.0f || gr[6] != 0.0f || gr[7] != 0.0f || gr[8] != 0.0f;
    if (kMoments && !screen) add_zero_stats(valid);
    bool colour = deg > 0 && (gr[0] != 0.0f || gr[1] != 0.0f || gr[2] != 0.0f);
    const bool any = screen || gr[0] != 0.0f || gr[1] != 0.0f || gr[2] != 0.0f || gr[3] != 0.0f;
    if (consume && any) {
      if (kMoments) {
        Float3u* r = reinterpret_cast<Float3u*>(rp);
        const Float3u z{0.f, 0.f, 0.f};
        r[0] = z; r[1] = z; r[2] = z;
      } else {
        float4* r = reinterpret_cast<float4*>(rp);
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        r[0] = z; r[1] = z; r[2] = z;
      }
    }
    if (touched) {  // profiling only: one atomic per wave
      const uint64_t m = __ballot(any);
      if (m && __lane_id() == uint32_t(__ffsll(static_cast<unsigned long long>(m)) - 1))
        atomicAdd(touched + v, uint32_t(__popcll(m)));
    }
    if (!screen && !colour) return;
    // A screen record with no valid projection here (it cannot have one: the
    // blends only see splats that projected) adds a zero norm.
    const bool stats_open = kMoments && screen;
    const CamF cam = load_cam(cams[v]);
    float p[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      p[k] = cam.V[k][0] * mean[0] + cam.V[k][1] * mean[1] + cam.V[k][2] * mean[2] + cam.t[k];
    if (p[2] < 0.2f) {
      if (stats_open) add_zero_stats(valid);
      return;
    }
    // Camera-frame covariance, as project_cov.
    float T[3][3], C[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) T[a][b] = cam.V[a][0] * S[0][b] + cam.V[a][1] * S[1][b] + cam.V[a][2] * S[2][b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) C[a][b] = T[a][0] * cam.V[b][0] + T[a][1] * cam.V[b][1] + T[a][2] * cam.V[b][2];
    const float x = p[0], y = p[1], z = p[2];
    const float iz = 1.0f / z, iz2 = iz * iz;
    const float j0[3] = {cam.fx * iz, 0.0f, -cam.fx * x * iz2};
    const float j1[3] = {0.0f, cam.fy * iz, -cam.fy * y * iz2};
    float Cj0[3], Cj1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Cj0[a] = C[a][0] * j0[0] + C[a][1] * j0[1] + C[a][2] * j0[2];
      Cj1[a] = C[a][0] * j1[0] + C[a][1] * j1[1] + C[a][2] * j1[2];
    }
    const float ca = j0[0] * Cj0[0] + j0[1] * Cj0[1] + j0[2] * Cj0[2] + 0.3f;
    const float cb = j0[0] * Cj1[0] + j0[1] * Cj1[1] + j0[2] * Cj1[2];
    const float cc = j1[0] * Cj1[0] + j1[1] * Cj1[1] + j1[2] * Cj1[2] + 0.3f;
    const float D = ca * cc - cb * cb;
    if (!(D > 0.0f)) {
      if (stats_open) add_zero_stats(valid);
      return;
    }
    if (screen) {
      // Conic (A, B, C) = (c, -b, a) / D.
      const float iD = 1.0f / D, iD2 = iD * iD;
      float gu = gr[4], gv = gr[5], gA = gr[6], gB = gr[7], gC = gr[8];
      if (kMoments) {
        const float A = cc * iD, B = -cb * iD, C = ca * iD;
        gu = fmaf(A, gr[4], B * gr[5]);
        gv = fmaf(B, gr[4], C * gr[5]);
        gA = -0.5f * gr[6];
        gB = -gr[7];
        gC = -0.5f * gr[8];
        add_stats(valid, gu, gv);
      }
      const float g_a = (-gA * cc * cc + gB * cb * cc - gC * cb * cb) * iD2;
      const float g_b = (2.0f * gA * cb * cc + 2.0f * gC * ca * cb - 2.0f * gB * cb * cb) * iD2 - gB * iD;
      const float g_c = (-gA * cb * cb + gB * ca * cb - gC * ca * ca) * iD2;
      // dL/dC (camera frame, symmetric) and dL/dj0, dL/dj1.
      float G[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
          G[a][b] = g_a * j0[a] * j0[b] + 0.5f * g_b * (j0[a] * j1[b] + j1[a] * j0[b]) + g_c * j1[a] * j1[b];
      float gj0[3], gj1[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        gj0[a] = 2.0f * g_a * Cj0[a] + g_b * Cj1[a];
        gj1[a] = 2.0f * g_c * Cj1[a] + g_b * Cj0[a];
      }
      float gp[3];
      gp[0] = gu * cam.fx * iz - gj0[2] * cam.fx * iz2;
      gp[1] = gv * cam.fy * iz - gj1[2] * cam.fy * iz2;
      gp[2] = -gu * cam.fx * x * iz2 - gv * cam.fy * y * iz2 - gj0[0] * cam.fx * iz2 +
              gj0[2] * 2.0f * cam.fx * x * iz2 * iz - gj1[1] * cam.fy * iz2 + gj1[2] * 2.0f * cam.fy * y * iz2 * iz;
#pragma unroll
      for (int b = 0; b < 3; ++b) gmean[b] += cam.V[0][b] * gp[0] + cam.V[1][b] * gp[1] + cam.V[2][b] * gp[2];
      // World covariance gradient V^T G V.
      float T1[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) T1[a][b] = cam.V[0][a] * G[0][b] + cam.V[1][a] * G[1][b] + cam.V[2][a] * G[2][b];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Gw[a][b] += T1[a][0] * cam.V[0][b] + T1[a][1] * cam.V[1][b] + T1[a][2] * cam.V[2][b];
    }
    if (colour) {
      float dir[3] = {mean[0] - cam.centre[0], mean[1] - cam.centre[1], mean[2] - cam.centre[2]};
      float len = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
      if (len > 0.0f) {
        float il = 1.0f / len;
        float d[3] = {dir[0] * il, dir[1] * il, dir[2] * il};
        float Y[16], GY[16][3];
        sh_basis(deg, d, Y);
        sh_basis_grad(deg, d, GY);
        const int K = (deg + 1) * (deg + 1);
        float gd[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          gsh[c] += gr[c];
#pragma unroll
          for (int k = 1; k < 16; ++k) {
            if (k >= K) continue;
            gsh[k * 3 + c] += Y[k] * gr[c];
            float w = gr[c] * shc[k * 3 + c];
            gd[0] += w * GY[k][0];
            gd[1] += w * GY[k][1];
            gd[2] += w * GY[k][2];
          }
        }
        float dd = d[0] * gd[0] + d[1] * gd[1] + d[2] * gd[2];
#pragma unroll
        for (int e = 0; e < 3; ++e) gmean[e] += (gd[e] - d[e] * dd) * il;
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) gsh[c] += gr[c];
      }
    }
  };
  for (uint32_t v0 = 0; v0 < num_views; v0 += kAhead) {
    float rec[kAhead][9];
#pragma unroll
    for (uint32_t k = 0; k < kAhead; ++k) {
      // Past the last view the last view's record is loaded again and not used.
      const uint32_t v = min(v0 + k, num_views - 1);
      const float* rp = grads + size_t(v) * stride + size_t(i) * record;
      if (kMoments) {  // 9-float moment records, 4-byte aligned: three 12-byte loads
        const Float3u* r = reinterpret_cast<const Float3u*>(rp);
        const Float3u a = r[0], b = r[1], c = r[2];
        rec[k][0] = a.x; rec[k][1] = a.y; rec[k][2] = a.z;
        rec[k][3] = b.x; rec[k][4] = b.y; rec[k][5] = b.z;
        rec[k][6] = c.x; rec[k][7] = c.y; rec[k][8] = c.z;
      } else {  // the reference's 16-float records
        const float4* r = reinterpret_cast<const float4*>(rp);
        const float4 a = r[0], b = r[1];
        rec[k][0] = a.x; rec[k][1] = a.y; rec[k][2] = a.z; rec[k][3] = a.w;
        rec[k][4] = b.x; rec[k][5] = b.y; rec[k][6] = b.z; rec[k][7] = b.w;
        rec[k][8] = rp[8];
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kAhead; ++k) {
      if (v0 + k >= num_views) continue;
      const bool valid = stats && ((valid_mask[size_t(v0 + k) * mask_words + i / 64] >> (i % 64)) & 1u);
      view_step(v0 + k, rec[k], valid);
    }
  }
  if (stats) {
    stats[size_t(i) * 2] = st_norm;
    stats[size_t(i) * 2 + 1] = st_count;
  }
  finish_gaussian(i, q, Gw, gmean, gsh, gcol, gop, deg, shape, sh_grad, grads_out);
}

// ---- many views: the per-view chain run lane-dense ---------------------------
// In shape_kernel a lane owns a Gaussian and the wave runs the per-view chain for
// a view whenever one of its 64 Gaussians has a nonzero record there.  About 5%
// of the records are nonzero on the 32-view 1M-Gaussian bench scene, so a wave
// runs the chain for nearly every view with a few lanes active.  shape_dense_kernel
// queues the (Gaussian, view) records that need the chain in LDS and runs it one
// record per lane; each Gaussian's lane then adds its records' terms in view
// order, as shape_kernel does.

// What the per-view chain of one moment record adds to its Gaussian.
constexpr uint32_t kTermScreen = 1;   // Gw and gmean_s (a screen record that projects)
constexpr uint32_t kTermCount = 2;    // STATS count + 1 (valid projection)
constexpr uint32_t kTermNorm = 4;     // STATS norm + st_norm (with kTermCount)
constexpr uint32_t kTermColourY = 8;  // SH: the DC term, Y[k] gr[c] and gmean_c
constexpr uint32_t kTermColourDc = 16;  // SH: the DC term only (the camera at the mean)
constexpr int kTermWords = 33;  // Gw 9, gmean_s 3, gmean_c 3, st_norm, flags, Y[1..15]; odd: no bank conflicts

// The chain of one view for a moment record gr (view_step of shape_kernel with
// moments): mean, S and shc are the Gaussian's, valid its projection flag in the
// view (STATS).  Writes the terms to out[kTermWords].
__device__ inline void view_terms(const CamF& cam, const float mean[3], const float (&S)[3][3], const float gr[9],
                                  bool screen, bool colour, bool valid, int deg, const float* __restrict__ shc,
                                  float* out) {
  uint32_t flags = 0;
  float st_norm = 0.0f;
  float Gw[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  float gms[3] = {0.f, 0.f, 0.f}, gmc[3] = {0.f, 0.f, 0.f};
  float Y[16] = {};
  float p[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) p[k] = cam.V[k][0] * mean[0] + cam.V[k][1] * mean[1] + cam.V[k][2] * mean[2] + cam.t[k];
  bool ok = p[2] >= 0.2f;
  float ca = 0.f, cb = 0.f, cc = 0.f, D = 0.f;
  float j0[3] = {}, j1[3] = {}, Cj0[3] = {}, Cj1[3] = {};
  float iz = 0.f, iz2 = 0.f;
  if (ok) {
    // Camera-frame covariance, as project_cov.
    float T[3][3], C[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) T[a][b] = cam.V[a][0] * S[0][b] + cam.V[a][1] * S[1][b] + cam.V[a][2] * S[2][b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) C[a][b] = T[a][0] * cam.V[b][0] + T[a][1] * cam.V[b][1] + T[a][2] * cam.V[b][2];
    iz = 1.0f / p[2];
    iz2 = iz * iz;
    j0[0] = cam.fx * iz; j0[2] = -cam.fx * p[0] * iz2;
    j1[1] = cam.fy * iz; j1[2] = -cam.fy * p[1] * iz2;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Cj0[a] = C[a][0] * j0[0] + C[a][1] * j0[1] + C[a][2] * j0[2];
      Cj1[a] = C[a][0] * j1[0] + C[a][1] * j1[1] + C[a][2] * j1[2];
    }
    ca = j0[0] * Cj0[0] + j0[1] * Cj0[1] + j0[2] * Cj0[2] + 0.3f;
    cb = j0[0] * Cj1[0] + j0[1] * Cj1[1] + j0[2] * Cj1[2];
    cc = j1[0] * Cj1[0] + j1[1] * Cj1[1] + j1[2] * Cj1[2] + 0.3f;
    D = ca * cc - cb * cb;
    ok = D > 0.0f;
  }
  // A screen record without a projection here (it cannot have one: the blends
  // only see splats that projected) adds a zero norm.
  if (screen && !ok && valid) flags |= kTermCount;
  if (ok && screen) {
    const float x = p[0], y = p[1];
    // Conic (A, B, C) = (c, -b, a) / D.
    const float iD = 1.0f / D, iD2 = iD * iD;
    const float A = cc * iD, B = -cb * iD, Cn = ca * iD;
    const float gu = fmaf(A, gr[4], B * gr[5]);
    const float gv = fmaf(B, gr[4], Cn * gr[5]);
    const float gA = -0.5f * gr[6], gB = -gr[7], gC = -0.5f * gr[8];
    if (valid) {
      flags |= kTermCount | kTermNorm;
      st_norm = sqrtf(gu * gu + gv * gv);
    }
    const float g_a = (-gA * cc * cc + gB * cb * cc - gC * cb * cb) * iD2;
    const float g_b = (2.0f * gA * cb * cc + 2.0f * gC * ca * cb - 2.0f * gB * cb * cb) * iD2 - gB * iD;
    const float g_c = (-gA * cb * cb + gB * ca * cb - gC * ca * ca) * iD2;
    // dL/dC (camera frame, symmetric) and dL/dj0, dL/dj1.
    float G[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b)
        G[a][b] = g_a * j0[a] * j0[b] + 0.5f * g_b * (j0[a] * j1[b] + j1[a] * j0[b]) + g_c * j1[a] * j1[b];
    float gj0[3], gj1[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      gj0[a] = 2.0f * g_a * Cj0[a] + g_b * Cj1[a];
      gj1[a] = 2.0f * g_c * Cj1[a] + g_b * Cj0[a];
    }
    float gp[3];
    gp[0] = gu * cam.fx * iz - gj0[2] * cam.fx * iz2;
    gp[1] = gv * cam.fy * iz - gj1[2] * cam.fy * iz2;
    gp[2] = -gu * cam.fx * x * iz2 - gv * cam.fy * y * iz2 - gj0[0] * cam.fx * iz2 +
            gj0[2] * 2.0f * cam.fx * x * iz2 * iz - gj1[1] * cam.fy * iz2 + gj1[2] * 2.0f * cam.fy * y * iz2 * iz;
#pragma unroll
    for (int b = 0; b < 3; ++b) gms[b] = cam.V[0][b] * gp[0] + cam.V[1][b] * gp[1] + cam.V[2][b] * gp[2];
    // World covariance gradient V^T G V.
    float T1[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) T1[a][b] = cam.V[0][a] * G[0][b] + cam.V[1][a] * G[1][b] + cam.V[2][a] * G[2][b];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) Gw[a][b] = T1[a][0] * cam.V[0][b] + T1[a][1] * cam.V[1][b] + T1[a][2] * cam.V[2][b];
    flags |= kTermScreen;
  }
  if (ok && colour) {
    float dir[3] = {mean[0] - cam.centre[0], mean[1] - cam.centre[1], mean[2] - cam.centre[2]};
    float len = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (len > 0.0f) {
      float il = 1.0f / len;
      float d[3] = {dir[0] * il, dir[1] * il, dir[2] * il};
      float GY[16][3];
      sh_basis(deg, d, Y);
      sh_basis_grad(deg, d, GY);
      const int K = (deg + 1) * (deg + 1);
      float gd[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int k = 1; k < 16; ++k) {
          if (k >= K) continue;
          float w = gr[c] * shc[k * 3 + c];
          gd[0] += w * GY[k][0];
          gd[1] += w * GY[k][1];
          gd[2] += w * GY[k][2];
        }
      }
      float dd = d[0] * gd[0] + d[1] * gd[1] + d[2] * gd[2];
#pragma unroll
      for (int e = 0; e < 3; ++e) gmc[e] = (gd[e] - d[e] * dd) * il;
      flags |= kTermColourY;
    } else {
      flags |= kTermColourDc;
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) out[a * 3 + b] = Gw[a][b];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    out[9 + e] = gms[e];
    out[12 + e] = gmc[e];
  }
  out[15] = st_norm;
  out[16] = __uint_as_float(flags);
#pragma unroll
  for (int k = 1; k < 16; ++k) out[16 + k] = Y[k];
}

// One wave per 64 Gaussians (lane = the Gaussian's owner), moment records, many
// views.  Owners walk the views four at a time: colour and opacity terms, STATS
// zero-norm counts and the consumed records' zeroing as in shape_kernel; records
// that need the per-view chain are queued (owner, view, record).  Before a view
// could overflow the queue, lane j runs the chain for queued record j (the
// owner's mean and covariance from LDS, its SH coefficients from memory), and
// each owner adds its records' terms, in queue (= view) order.
constexpr uint32_t kDenseQueue = 64;

__global__ __launch_bounds__(64) void shape_dense_kernel(const gsr_camera* __restrict__ cams, uint32_t num_views,
                                                         const float* __restrict__ gs, const float* __restrict__ sh,
                                                         int deg, uint32_t n, float* __restrict__ grads,
                                                         uint64_t stride, float* __restrict__ shape,
                                                         float* __restrict__ sh_grad, float* __restrict__ grads_out,
                                                         int consume, uint32_t* __restrict__ touched,
                                                         uint32_t record, const uint64_t* __restrict__ valid_mask,
                                                         float* __restrict__ stats,
                                                         const uint32_t* __restrict__ stats_guard) {
  __shared__ float geo[64][9];                 // owners' mean and world covariance (xx xy xz yy yz zz)
  __shared__ uint32_t q_item[kDenseQueue];     // owner lane | valid << 6 | screen << 7 | colour << 8 | view << 9
  __shared__ float q_rec[kDenseQueue][9];
  __shared__ float q_out[kDenseQueue][kTermWords];
  const uint32_t lane = threadIdx.x, base = blockIdx.x * 64u, i = base + lane;
  const bool own = i < n;
  if (stats && stats_guard && *stats_guard) stats = nullptr;  // an overflowed step adds its statistics once
  float st_norm = 0.0f, st_count = 0.0f;
  if (stats && own) {
    st_norm = stats[size_t(i) * 2];
    st_count = stats[size_t(i) * 2 + 1];
  }
  float q[10] = {}, S[3][3] = {};
  float op_scale = 0.0f;
  if (own) {
    const float4* g4 = reinterpret_cast<const float4*>(gs) + size_t(i) * 4;
    const float4 pa = g4[0], pb = g4[1], pc = g4[2];
    const float qq[10] = {pa.x, pa.y, pa.z, pb.x, pb.y, pb.z, pc.x, pc.y, pc.z, pc.w};
#pragma unroll
    for (int k = 0; k < 10; ++k) q[k] = qq[k];
    world_cov(q, S);
    // d opLogit = (1 - op) m3, op the sigmoid of gpu.js:85 as the projection computes it.
    op_scale = 1.0f - 1.0f / (1.0f + expf(-pb.w));
  }
  geo[lane][0] = q[0]; geo[lane][1] = q[1]; geo[lane][2] = q[2];
  geo[lane][3] = S[0][0]; geo[lane][4] = S[0][1]; geo[lane][5] = S[0][2];
  geo[lane][6] = S[1][1]; geo[lane][7] = S[1][2]; geo[lane][8] = S[2][2];
  float Gw[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
  float gmean[3] = {0.f, 0.f, 0.f}, gcol[3] = {0.f, 0.f, 0.f}, gsh[48];
#pragma unroll
  for (int k = 0; k < 48; ++k) gsh[k] = 0.0f;
  float gop = 0.0f;
  const size_t mask_words = valid_mask_words(n);
  const uint64_t below = (uint64_t(1) << lane) - 1;
  uint32_t queued = 0;  // wave-uniform
  uint64_t mine = 0;    // this owner's queue slots

  // Runs the chain for queue slots [0, queued) and adds the terms to their owners.
  auto flush = [&]() __attribute__((always_inline)) {
    __syncthreads();
    if (lane < queued) {
      const uint32_t it = q_item[lane], o = it & 63u, v = it >> 9;
      const float mn[3] = {geo[o][0], geo[o][1], geo[o][2]};
      const float So[3][3] = {{geo[o][3], geo[o][4], geo[o][5]},
                              {geo[o][4], geo[o][6], geo[o][7]},
                              {geo[o][5], geo[o][7], geo[o][8]}};
      float gr[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) gr[k] = q_rec[lane][k];
      const bool colour = (it >> 8) & 1u;
      float shc[48];
      if (colour) {
        const float4* s4 = reinterpret_cast<const float4*>(sh) + size_t(base + o) * 12;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
          const float4 x = s4[k];
          shc[4 * k] = x.x; shc[4 * k + 1] = x.y; shc[4 * k + 2] = x.z; shc[4 * k + 3] = x.w;
        }
      }
      view_terms(load_cam(cams[v]), mn, So, gr, (it >> 7) & 1u, colour, (it >> 6) & 1u, deg, shc, q_out[lane]);
    }
    __syncthreads();
    uint64_t m = mine;
    while (m) {
      const uint32_t j = uint32_t(__builtin_ctzll(m));
      m &= m - 1;
      const float* t = q_out[j];
      const uint32_t flags = __float_as_uint(t[16]);
      if (flags & kTermCount) st_count += 1.0f;
      if (flags & kTermNorm) st_norm += t[15];
      if (flags & kTermScreen) {
#pragma unroll
        for (int b = 0; b < 3; ++b) gmean[b] += t[9 + b];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int b = 0; b < 3; ++b) Gw[a][b] += t[a * 3 + b];
      }
      if (flags & (kTermColourY | kTermColourDc)) {
        const float* rc = q_rec[j];
#pragma unroll
        for (int c = 0; c < 3; ++c) gsh[c] += rc[c];
        if (flags & kTermColourY) {
#pragma unroll
          for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 1; k < 16; ++k) gsh[k * 3 + c] += t[16 + k] * rc[c];
#pragma unroll
          for (int e = 0; e < 3; ++e) gmean[e] += t[12 + e];
        }
      }
    }
    mine = 0;
    queued = 0;
    __syncthreads();
  };

  constexpr uint32_t kAhead = 4;
  for (uint32_t v0 = 0; v0 < num_views; v0 += kAhead) {
    float rec[kAhead][9];
#pragma unroll
    for (uint32_t k = 0; k < kAhead; ++k) {
      // Past the last view (or the last Gaussian) nothing is loaded: zeros.
      const uint32_t v = v0 + k;
#pragma unroll
      for (int e = 0; e < 9; ++e) rec[k][e] = 0.0f;
      if (own && v < num_views) {
        const Float3u* r = reinterpret_cast<const Float3u*>(grads + size_t(v) * stride + size_t(i) * record);
        const Float3u a = r[0], b = r[1], c = r[2];
        rec[k][0] = a.x; rec[k][1] = a.y; rec[k][2] = a.z;
        rec[k][3] = b.x; rec[k][4] = b.y; rec[k][5] = b.z;
        rec[k][6] = c.x; rec[k][7] = c.y; rec[k][8] = c.z;
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kAhead; ++k) {
      const uint32_t v = v0 + k;
      if (v >= num_views) continue;  // wave-uniform
      const float* gr = rec[k];
      const bool valid = stats && own && ((valid_mask[size_t(v) * mask_words + i / 64] >> (i % 64)) & 1u);
      gcol[0] += gr[0]; gcol[1] += gr[1]; gcol[2] += gr[2];
      gop = fmaf(op_scale, gr[3], gop);
      const bool screen = gr[4] != 0.0f || gr[5] != 0.0f || gr[6] != 0.0f || gr[7] != 0.0f || gr[8] != 0.0f;
      const bool colour = deg > 0 && (gr[0] != 0.0f || gr[1] != 0.0f || gr[2] != 0.0f);
      const bool any = screen || gr[0] != 0.0f || gr[1] != 0.0f || gr[2] != 0.0f || gr[3] != 0.0f;
      if (!screen && valid) st_count += 1.0f;  // a zero (du, dv): the norm adds +0
      if (consume && any) {
        Float3u* r = reinterpret_cast<Float3u*>(grads + size_t(v) * stride + size_t(i) * record);
        const Float3u z{0.f, 0.f, 0.f};
        r[0] = z; r[1] = z; r[2] = z;
      }
      if (touched) {  // profiling only: one atomic per wave
        const uint64_t m = __ballot(any);
        if (m && lane == uint32_t(__ffsll(static_cast<unsigned long long>(m)) - 1))
          atomicAdd(touched + v, uint32_t(__popcll(m)));
      }
      const bool need = screen || colour;
      const uint64_t needs = __ballot(need);
      if (needs == 0) continue;
      if (queued + uint32_t(__popcll(needs)) > kDenseQueue) flush();
      if (need) {
        const uint32_t slot = queued + uint32_t(__popcll(needs & below));
        q_item[slot] = lane | (valid ? 1u << 6 : 0u) | (screen ? 1u << 7 : 0u) | (colour ? 1u << 8 : 0u) | (v << 9);
#pragma unroll
        for (int e = 0; e < 9; ++e) q_rec[slot][e] = gr[e];
        mine |= uint64_t(1) << slot;
      }
      queued += uint32_t(__popcll(needs));
    }
  }
  if (queued > 0) flush();
  if (!own) return;
  if (stats) {
    stats[size_t(i) * 2] = st_norm;
    stats[size_t(i) * 2 + 1] = st_count;
  }
  finish_gaussian(i, q, Gw, gmean, gsh, gcol, gop, deg, shape, sh_grad, grads_out);
}

}  // namespace

void launch_shape(hipStream_t s, const gsr_camera* cams, uint32_t num_views, const float* gs, const float* sh,
                  int sh_degree, uint32_t n, float* grads, uint64_t grads_view_stride, float* shape,
                  float* sh_grad, float* grads_out, int consume, uint32_t* touched, uint32_t record,
                  const uint64_t* valid_mask, float* stats, const uint32_t* stats_guard, int moments) {
  if (n == 0) return;
  if (!valid_mask) stats = nullptr;
  // The dense chain's queue word holds the view in its top 23 bits.
  if (moments && num_views > 1 && num_views < (1u << 23))
    hipLaunchKernelGGL(shape_dense_kernel, dim3((n + 63) / 64), dim3(64), 0, s, cams, num_views, gs, sh, sh_degree,
                       n, grads, grads_view_stride, shape, sh_grad, grads_out, consume, touched, record, valid_mask,
                       stats, stats_guard);
  else if (moments)
    hipLaunchKernelGGL(shape_kernel<true>, dim3((n + 255) / 256), dim3(256), 0, s, cams, num_views, gs, sh,
                       sh_degree, n, grads, grads_view_stride, shape, sh_grad, grads_out, consume, touched, record,
                       valid_mask, stats, stats_guard);
  else
    hipLaunchKernelGGL(shape_kernel<false>, dim3((n + 255) / 256), dim3(256), 0, s, cams, num_views, gs, sh,
                       sh_degree, n, grads, grads_view_stride, shape, sh_grad, grads_out, consume, touched, record,
                       valid_mask, stats, stats_guard);
}

}  // namespace gsr
