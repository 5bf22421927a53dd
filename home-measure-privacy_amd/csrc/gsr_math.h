// Device-side mathematics shared by the projection, binning and chain kernels.
// Operation order follows the WGSL of trainer/gpu.js so that, built without FMA
// contraction, results stay within a few ulp of a float restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "gsr.h"

namespace gsr {

struct CamF {
  float V[3][3];
  float t[3];
  float fx, fy, cx, cy;
  float centre[3];  // -V^T t
};

__device__ inline CamF load_cam(const gsr_camera& c) {
  CamF o;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) o.V[i][j] = c.view[i * 4 + j];
    o.t[i] = c.view[i * 4 + 3];
  }
  o.fx = c.fx; o.fy = c.fy; o.cx = c.cx; o.cy = c.cy;
#pragma unroll
  for (int j = 0; j < 3; ++j) o.centre[j] = -(o.V[0][j] * o.t[0] + o.V[1][j] * o.t[1] + o.V[2][j] * o.t[2]);
  return o;
}

// quatToMat (gpu.js:20-27), R[row][col]; returns max(|q|, 1e-12).
__device__ inline float quat_to_mat(const float q[4], float R[3][3]) {
  float len = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  float n = fmaxf(len, 1e-12f);
  float w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
  R[0][0] = 1.0f - 2.0f * (y * y + z * z);
  R[1][0] = 2.0f * (x * y + w * z);
  R[2][0] = 2.0f * (x * z - w * y);
  R[0][1] = 2.0f * (x * y - w * z);
  R[1][1] = 1.0f - 2.0f * (x * x + z * z);
  R[2][1] = 2.0f * (y * z + w * x);
  R[0][2] = 2.0f * (x * z + w * y);
  R[1][2] = 2.0f * (y * z - w * x);
  R[2][2] = 1.0f - 2.0f * (x * x + y * y);
  return n;
}

// World covariance M M^T with M = R(q) diag(exp(logScale)) (gpu.js:37-39).
// q = mean.xyz, logScale.xyz, rotation.wxyz.
__device__ inline void world_cov(const float q[10], float S[3][3]) {
  float s[3] = {expf(q[3]), expf(q[4]), expf(q[5])};
  float R[3][3];
  quat_to_mat(q + 6, R);
  float M[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) M[i][j] = R[i][j] * s[j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) S[i][j] = M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2];
}

struct Projected {
  bool ok;
  float u, v, a, b, c, radius, depth;
};

// projectShape (gpu.js:30-60) from the mean and world covariance.
__device__ inline Projected project_cov(const CamF& cam, const float mean[3], const float S[3][3]) {
  Projected o;
  o.ok = false;
  o.u = o.v = o.a = o.b = o.c = o.radius = o.depth = 0.0f;
  float p[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    p[i] = cam.V[i][0] * mean[0] + cam.V[i][1] * mean[1] + cam.V[i][2] * mean[2] + cam.t[i];
  if (p[2] < 0.2f) return o;
  float T[3][3], C[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) T[i][j] = cam.V[i][0] * S[0][j] + cam.V[i][1] * S[1][j] + cam.V[i][2] * S[2][j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) C[i][j] = T[i][0] * cam.V[j][0] + T[i][1] * cam.V[j][1] + T[i][2] * cam.V[j][2];
  float j0[3] = {cam.fx / p[2], 0.0f, -cam.fx * p[0] / (p[2] * p[2])};
  float j1[3] = {0.0f, cam.fy / p[2], -cam.fy * p[1] / (p[2] * p[2])};
  float Cj0[3], Cj1[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Cj0[i] = C[i][0] * j0[0] + C[i][1] * j0[1] + C[i][2] * j0[2];
    Cj1[i] = C[i][0] * j1[0] + C[i][1] * j1[1] + C[i][2] * j1[2];
  }
  float a = j0[0] * Cj0[0] + j0[1] * Cj0[1] + j0[2] * Cj0[2] + 0.3f;
  float b = j0[0] * Cj1[0] + j0[1] * Cj1[1] + j0[2] * Cj1[2];
  float c = j1[0] * Cj1[0] + j1[1] * Cj1[1] + j1[2] * Cj1[2] + 0.3f;
  float det = a * c - b * b;
  if (det <= 0.0f) return o;  // gpu.js:49 (a NaN det passes, as in WGSL)
  float mid = (a + c) / 2.0f;
  o.ok = true;
  o.u = cam.fx * (p[0] / p[2]) + cam.cx;
  o.v = cam.fy * (p[1] / p[2]) + cam.cy;
  o.a = c / det;
  o.b = -b / det;
  o.c = a / det;
  o.radius = ceilf(3.0f * sqrtf(mid + sqrtf(fmaxf(0.1f, mid * mid - det))));
  o.depth = p[2];
  return o;
}

// ---- spherical harmonics (extension; the reference is degree 0, gpu.js:6) ----
// Real SH with the sign convention of the 3D Gaussian Splatting code (Kerbl et
// al. 2023; constants from Sloan, "Stupid spherical harmonics tricks").  Y[0] = 1
// because the DC coefficient is stored in colour units.
#define GSR_SH_C1 0.4886025119029199f
#define GSR_SH_C2_0 1.0925484305920792f
#define GSR_SH_C2_1 (-1.0925484305920792f)
#define GSR_SH_C2_2 0.31539156525252005f
#define GSR_SH_C2_3 (-1.0925484305920792f)
#define GSR_SH_C2_4 0.5462742152960396f
#define GSR_SH_C3_0 (-0.5900435899266435f)
#define GSR_SH_C3_1 2.890611442640554f
#define GSR_SH_C3_2 (-0.4570457994644658f)
#define GSR_SH_C3_3 0.3731763325901154f
#define GSR_SH_C3_4 (-0.4570457994644658f)
#define GSR_SH_C3_5 1.445305721320277f
#define GSR_SH_C3_6 (-0.5900435899266435f)

__device__ inline void sh_basis(int deg, const float d[3], float Y[16]) {
  float x = d[0], y = d[1], z = d[2];
  Y[0] = 1.0f;
#pragma unroll
  for (int k = 1; k < 16; ++k) Y[k] = 0.0f;
  if (deg < 1) return;
  Y[1] = -GSR_SH_C1 * y;
  Y[2] = GSR_SH_C1 * z;
  Y[3] = -GSR_SH_C1 * x;
  if (deg < 2) return;
  float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  Y[4] = GSR_SH_C2_0 * xy;
  Y[5] = GSR_SH_C2_1 * yz;
  Y[6] = GSR_SH_C2_2 * (2.0f * zz - xx - yy);
  Y[7] = GSR_SH_C2_3 * xz;
  Y[8] = GSR_SH_C2_4 * (xx - yy);
  if (deg < 3) return;
  Y[9] = GSR_SH_C3_0 * y * (3.0f * xx - yy);
  Y[10] = GSR_SH_C3_1 * xy * z;
  Y[11] = GSR_SH_C3_2 * y * (4.0f * zz - xx - yy);
  Y[12] = GSR_SH_C3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  Y[13] = GSR_SH_C3_4 * x * (4.0f * zz - xx - yy);
  Y[14] = GSR_SH_C3_5 * z * (xx - yy);
  Y[15] = GSR_SH_C3_6 * x * (xx - 3.0f * yy);
}

// dY_k / d(direction), k = 1..15 (partials of the polynomials above).
__device__ inline void sh_basis_grad(int deg, const float d[3], float G[16][3]) {
  float x = d[0], y = d[1], z = d[2];
#pragma unroll
  for (int k = 0; k < 16; ++k) G[k][0] = G[k][1] = G[k][2] = 0.0f;
  if (deg < 1) return;
  G[1][1] = -GSR_SH_C1;
  G[2][2] = GSR_SH_C1;
  G[3][0] = -GSR_SH_C1;
  if (deg < 2) return;
  G[4][0] = GSR_SH_C2_0 * y; G[4][1] = GSR_SH_C2_0 * x;
  G[5][1] = GSR_SH_C2_1 * z; G[5][2] = GSR_SH_C2_1 * y;
  G[6][0] = -2.0f * GSR_SH_C2_2 * x; G[6][1] = -2.0f * GSR_SH_C2_2 * y; G[6][2] = 4.0f * GSR_SH_C2_2 * z;
  G[7][0] = GSR_SH_C2_3 * z; G[7][2] = GSR_SH_C2_3 * x;
  G[8][0] = 2.0f * GSR_SH_C2_4 * x; G[8][1] = -2.0f * GSR_SH_C2_4 * y;
  if (deg < 3) return;
  float xx = x * x, yy = y * y, zz = z * z;
  G[9][0] = GSR_SH_C3_0 * 6.0f * x * y;
  G[9][1] = GSR_SH_C3_0 * (3.0f * xx - 3.0f * yy);
  G[10][0] = GSR_SH_C3_1 * y * z; G[10][1] = GSR_SH_C3_1 * x * z; G[10][2] = GSR_SH_C3_1 * x * y;
  G[11][0] = -GSR_SH_C3_2 * 2.0f * x * y;
  G[11][1] = GSR_SH_C3_2 * (4.0f * zz - xx - 3.0f * yy);
  G[11][2] = GSR_SH_C3_2 * 8.0f * y * z;
  G[12][0] = -GSR_SH_C3_3 * 6.0f * x * z;
  G[12][1] = -GSR_SH_C3_3 * 6.0f * y * z;
  G[12][2] = GSR_SH_C3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy);
  G[13][0] = GSR_SH_C3_4 * (4.0f * zz - 3.0f * xx - yy);
  G[13][1] = -GSR_SH_C3_4 * 2.0f * x * y;
  G[13][2] = GSR_SH_C3_4 * 8.0f * x * z;
  G[14][0] = GSR_SH_C3_5 * 2.0f * x * z;
  G[14][1] = -GSR_SH_C3_5 * 2.0f * y * z;
  G[14][2] = GSR_SH_C3_5 * (xx - yy);
  G[15][0] = GSR_SH_C3_6 * (3.0f * xx - 3.0f * yy);
  G[15][1] = -GSR_SH_C3_6 * 6.0f * x * y;
}

// colour_c = sh[c] + sum_{k>=1} Y_k(dir) sh[3k + c], dir = normalise(mean - centre).
__device__ inline void sh_colour(int deg, const float sh[48], const float mean[3], const CamF& cam, float rgb[3]) {
  // The colour feeds no discrete rule of the blend, so its sum may fuse.
#pragma clang fp contract(fast)
  float dir[3] = {mean[0] - cam.centre[0], mean[1] - cam.centre[1], mean[2] - cam.centre[2]};
  float len = sqrtf(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  float inv = len > 0.0f ? 1.0f / len : 0.0f;
  float d[3] = {dir[0] * inv, dir[1] * inv, dir[2] * inv};
  float Y[16];
  sh_basis(deg, d, Y);
  int K = (deg + 1) * (deg + 1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float acc = sh[c];
#pragma unroll
    for (int k = 1; k < 16; ++k)
      if (k < K) acc += Y[k] * sh[k * 3 + c];
    rgb[c] = acc;
  }
}

// Tile rectangle of tileLists (gpu.js:456-462), computed in float64 from the
// float32 record as JavaScript does.  False when the rectangle is empty, which
// includes every NaN bound (Math.max/Math.min propagate NaN and the loops then
// run zero times).
__device__ inline bool tile_rect(float u, float v, float rx, float ry, uint32_t tiles_x, uint32_t tiles_y, int& x0,
                                 int& x1, int& y0, int& y1) {
  double du = u, dv = v, drx = rx, dry = ry;
  double fx0 = floor((du - drx - 0.5) / 16.0);
  double fx1 = floor((du + drx - 0.5) / 16.0);
  double fy0 = floor((dv - dry - 0.5) / 16.0);
  double fy1 = floor((dv + dry - 0.5) / 16.0);
  if (isnan(fx0) || isnan(fx1) || isnan(fy0) || isnan(fy1)) return false;
  fx0 = fmax(0.0, fx0);
  fx1 = fmin(double(tiles_x) - 1.0, fx1);
  fy0 = fmax(0.0, fy0);
  fy1 = fmin(double(tiles_y) - 1.0, fy1);
  if (!(fx0 <= fx1) || !(fy0 <= fy1)) return false;
  x0 = int(fx0); x1 = int(fx1); y0 = int(fy0); y1 = int(fy1);
  return true;
}

constexpr float kNever = -1e30f;  // extent of an entry that cannot hit any pixel

// Half-extents (ex, ey) of the pixel offsets (dx, dy) at which alphaAt
// (gpu.js:92-103) can return a hit: inside its square |dx|, |dy| <= r and inside
// the ellipse where opacity * exp(power) >= 1/255, i.e. -2 power <= 2 ln(255 op)
// with -2 power = [dx dy] [[A B] [B C]] [dx dy]^T, whose x half-width is
// sqrt(2 t C / det) and y half-width sqrt(2 t A / det).  The threshold t is raised
// by 0.01 plus a bound on the float32 rounding of the exponent's terms, det is
// Kahan's compensated A C - B^2, and the widths are padded, so every pixel outside
// the box evaluates to a miss in the blend's own arithmetic too and skipping it
// changes no result.  Opacities below ~1/255 get kNever;
// degenerate conics fall back to the square.
__device__ inline float hit_threshold(float A, float B, float C, float op, float r) {
  const float r1 = r + 1.0f;
  const float slack = 0.01f + 1e-6f * (fabsf(A) + 2.0f * fabsf(B) + fabsf(C)) * r1 * r1;
  return __logf(255.0f * op) + slack;
}

__device__ inline float2 hit_extent(float A, float B, float C, float op, float r) {
  const float t = hit_threshold(A, B, C, op, r);
  const float bb = B * B;
  const float det = fmaf(A, C, -bb) - fmaf(B, B, -bb);
  if (t != t || !(det > 0.0f) || !(A > 0.0f) || !(C > 0.0f)) return make_float2(r, r);
  if (!(t > 0.0f)) return make_float2(kNever, kNever);
  const float k = 2.0f * t / det;
  const float ex = fmaf(sqrtf(k * C), 1.0001f, 1e-3f);
  const float ey = fmaf(sqrtf(k * A), 1.0001f, 1e-3f);
  return make_float2(fminf(r, ex), fminf(r, ey));
}

// Least value of -2 power = A x^2 + 2 B x y + C y^2 (A, C > 0, A C > B^2) over
// the rectangle [x0, x1] x [y0, y1] of pixel offsets (x0 <= x1, y0 <= y1): 0 when
// the rectangle holds the centre, else the least of its four edges' minima, each
// at the edge's stationary point (y = -B x / C on an edge x = const, x = -B y / A
// on y = const) clamped into the edge.  A rounding error in that point raises
// the value only to second order in the error.
__device__ inline float conic_min_rect(float A, float B, float C, float x0, float x1, float y0, float y1) {
  const float bc = __fdividef(-B, C), ba = __fdividef(-B, A), b2 = 2.0f * B;
  const float ya = fminf(fmaxf(bc * x0, y0), y1), yb = fminf(fmaxf(bc * x1, y0), y1);
  const float xa = fminf(fmaxf(ba * y0, x0), x1), xb = fminf(fmaxf(ba * y1, x0), x1);
  const float qa = fmaf(x0, fmaf(A, x0, b2 * ya), C * ya * ya);
  const float qb = fmaf(x1, fmaf(A, x1, b2 * yb), C * yb * yb);
  const float qc = fmaf(y0, fmaf(C, y0, b2 * xa), A * xa * xa);
  const float qd = fmaf(y1, fmaf(C, y1, b2 * xb), A * xb * xb);
  const bool inside = x0 <= 0.0f && x1 >= 0.0f && y0 <= 0.0f && y1 >= 0.0f;
  return inside ? 0.0f : fminf(fminf(qa, qb), fminf(qc, qd));
}

}  // namespace gsr
