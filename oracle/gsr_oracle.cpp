// CPU oracle for the Gaussian-splat rasterizer.  TEST INFRASTRUCTURE ONLY: the
// tests, the smoke check and the bench's cpu_baseline leg may load it, as the
// checker or as the timed CPU baseline; the product path never calls it.
//
// Every routine is a template over the arithmetic type R and is exported twice:
//   *_f32  "reference-semantics restatement": the WGSL kernels of trainer/gpu.js
//          restated in float, same op order, built with -ffp-contract=off;
//   *_f64  the FP64 oracle: the same mathematics in double precision.
// The tile-list routine always computes its bounds in double, as the reference's
// JavaScript does (numbers are doubles, Float32Array reads are exact).
//
// Reference citations are trainer/gpu.js:line unless another file is named.
//
// Pinning (tests/test_oracle.py):
//   project      closed forms (isotropic Gaussian, conic*cov = I, q == -q,
//                near-plane / det rejects), brute-force radius definition.
//   SH basis     orthonormality by quadrature on the sphere, degree-0 reduction.
//   tile_lists   brute force on small scenes (per-tile membership by the f64
//                bounding box, per-tile (depth, index) order).
//   forward      closed forms for one and two Gaussians, T-stop and 1/255 rules.
//   loss         sign(x-y)/(3P) known answer (ssim-test.html:52-53) and central
//                differences of loss_value (ssim-test.html:41-51).
//   backward     central differences of the FP64 forward (weighted-sum loss).
//   shape chain  central differences of FP64 projectShape, and of the full FP64
//                forward with respect to every parameter group.
//   adam         Kingma & Ba closed form at t = 1 and t = 2 (also the SH3 variant).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// 80-byte camera uniform, gpu.js:16 and cameraUniform gpu.js:696-701.
struct CamU {
  float view[12];
  float fx, fy, cx, cy;
  float width, height;
  uint32_t count, pixels;
};
static_assert(sizeof(CamU) == 80, "camera uniform is 80 bytes");

// Step uniform of ADAM / STATS, gpu.js:412.
struct StepU {
  float t;
  uint32_t count;
  float lrColor, lrOpacity, lrMean, lrScale, lrRotation, pad;
};
static_assert(sizeof(StepU) == 32, "step uniform is 32 bytes");

template <class R>
struct Cam {
  R V[3][3];
  R t[3];
  R fx, fy, cx, cy;
  int W, H;
  explicit Cam(const CamU& c) {
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) V[i][j] = R(c.view[i * 4 + j]);
      t[i] = R(c.view[i * 4 + 3]);
    }
    fx = R(c.fx); fy = R(c.fy); cx = R(c.cx); cy = R(c.cy);
    W = int(c.width); H = int(c.height);
  }
  // Camera centre in world space: -V^T t (colmap.js:66-71).
  void centre(R c[3]) const {
    for (int j = 0; j < 3; ++j) c[j] = -(V[0][j] * t[0] + V[1][j] * t[1] + V[2][j] * t[2]);
  }
};

template <class R>
struct Proj {
  bool ok;
  R u, v, a, b, c, radius, depth;
};

// quatToMat, gpu.js:20-27.  WGSL builds the matrix from columns; here R[row][col].
template <class R>
void quat_to_mat(const R q[4], R M[3][3], R* len_out = nullptr) {
  R len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  R n = std::max(len, R(1e-12));
  if (len_out) *len_out = n;
  R w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
  M[0][0] = R(1) - R(2) * (y * y + z * z);
  M[1][0] = R(2) * (x * y + w * z);
  M[2][0] = R(2) * (x * z - w * y);
  M[0][1] = R(2) * (x * y - w * z);
  M[1][1] = R(1) - R(2) * (x * x + z * z);
  M[2][1] = R(2) * (y * z + w * x);
  M[0][2] = R(2) * (x * z + w * y);
  M[1][2] = R(2) * (y * z - w * x);
  M[2][2] = R(1) - R(2) * (x * x + y * y);
}

// Camera-frame 3D covariance W (R S)(R S)^T W^T of one Gaussian (gpu.js:37-41).
template <class R>
void cam_cov(const Cam<R>& cam, const R q[10], R C[3][3]) {
  R s[3] = {std::exp(q[3]), std::exp(q[4]), std::exp(q[5])};
  R Rm[3][3];
  quat_to_mat(q + 6, Rm);
  R M[3][3], S[3][3], T[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = Rm[i][j] * s[j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = M[i][0] * M[j][0] + M[i][1] * M[j][1] + M[i][2] * M[j][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T[i][j] = cam.V[i][0] * S[0][j] + cam.V[i][1] * S[1][j] + cam.V[i][2] * S[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = T[i][0] * cam.V[j][0] + T[i][1] * cam.V[j][1] + T[i][2] * cam.V[j][2];
}

// projectShape, gpu.js:30-60.  q = mean.xyz, logScale.xyz, rotation.wxyz.
template <class R>
Proj<R> project_shape(const Cam<R>& cam, const R q[10]) {
  Proj<R> o{};
  o.ok = false;
  R p[3];
  for (int i = 0; i < 3; ++i) p[i] = cam.V[i][0] * q[0] + cam.V[i][1] * q[1] + cam.V[i][2] * q[2] + cam.t[i];
  if (p[2] < R(0.2)) return o;
  R C[3][3];
  cam_cov(cam, q, C);
  R j0[3] = {cam.fx / p[2], R(0), -cam.fx * p[0] / (p[2] * p[2])};
  R j1[3] = {R(0), cam.fy / p[2], -cam.fy * p[1] / (p[2] * p[2])};
  R Cj0[3], Cj1[3];
  for (int i = 0; i < 3; ++i) {
    Cj0[i] = C[i][0] * j0[0] + C[i][1] * j0[1] + C[i][2] * j0[2];
    Cj1[i] = C[i][0] * j1[0] + C[i][1] * j1[1] + C[i][2] * j1[2];
  }
  R a = j0[0] * Cj0[0] + j0[1] * Cj0[1] + j0[2] * Cj0[2] + R(0.3);
  R b = j0[0] * Cj1[0] + j0[1] * Cj1[1] + j0[2] * Cj1[2];
  R c = j1[0] * Cj1[0] + j1[1] * Cj1[1] + j1[2] * Cj1[2] + R(0.3);
  R det = a * c - b * b;
  if (det <= R(0)) return o;
  R mid = (a + c) / R(2);
  o.ok = true;
  o.u = cam.fx * (p[0] / p[2]) + cam.cx;
  o.v = cam.fy * (p[1] / p[2]) + cam.cy;
  o.a = c / det;
  o.b = -b / det;
  o.c = a / det;
  o.radius = std::ceil(R(3) * std::sqrt(mid + std::sqrt(std::max(R(0.1), mid * mid - det))));
  o.depth = p[2];
  return o;
}

// ---- Spherical harmonics (build extension; the reference is degree 0 only, gpu.js:6).
// Real SH basis with the Condon-Shortley sign convention of the 3D Gaussian
// Splatting code (Kerbl et al. 2023; constants from Sloan, "Stupid SH tricks").
// The DC coefficient is stored in colour units: colour = dc + sum_{l>=1} Y_lm(d) f_lm,
// so degree 0 reduces to the reference's colour copy (gpu.js:87).
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                             -1.0925484305920792, 0.5462742152960396};
constexpr double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                             0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                             -0.5900435899266435};

// Basis values Y[1..15] for unit direction d (Y[0] = 1 in colour units).
template <class R>
void sh_basis(int deg, const R d[3], R Y[16]) {
  R x = d[0], y = d[1], z = d[2];
  Y[0] = R(1);
  if (deg < 1) return;
  Y[1] = -R(SH_C1) * y;
  Y[2] = R(SH_C1) * z;
  Y[3] = -R(SH_C1) * x;
  if (deg < 2) return;
  R xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  Y[4] = R(SH_C2[0]) * xy;
  Y[5] = R(SH_C2[1]) * yz;
  Y[6] = R(SH_C2[2]) * (R(2) * zz - xx - yy);
  Y[7] = R(SH_C2[3]) * xz;
  Y[8] = R(SH_C2[4]) * (xx - yy);
  if (deg < 3) return;
  Y[9] = R(SH_C3[0]) * y * (R(3) * xx - yy);
  Y[10] = R(SH_C3[1]) * xy * z;
  Y[11] = R(SH_C3[2]) * y * (R(4) * zz - xx - yy);
  Y[12] = R(SH_C3[3]) * z * (R(2) * zz - R(3) * xx - R(3) * yy);
  Y[13] = R(SH_C3[4]) * x * (R(4) * zz - xx - yy);
  Y[14] = R(SH_C3[5]) * z * (xx - yy);
  Y[15] = R(SH_C3[6]) * x * (xx - R(3) * yy);
}

// Jacobian dY[k]/dd (unnormalised partials in x, y, z), k = 1..15.
template <class R>
void sh_basis_grad(int deg, const R d[3], R G[16][3]) {
  R x = d[0], y = d[1], z = d[2];
  for (int k = 0; k < 16; ++k) G[k][0] = G[k][1] = G[k][2] = R(0);
  if (deg < 1) return;
  G[1][1] = -R(SH_C1);
  G[2][2] = R(SH_C1);
  G[3][0] = -R(SH_C1);
  if (deg < 2) return;
  G[4][0] = R(SH_C2[0]) * y; G[4][1] = R(SH_C2[0]) * x;
  G[5][1] = R(SH_C2[1]) * z; G[5][2] = R(SH_C2[1]) * y;
  G[6][0] = -R(2) * R(SH_C2[2]) * x; G[6][1] = -R(2) * R(SH_C2[2]) * y; G[6][2] = R(4) * R(SH_C2[2]) * z;
  G[7][0] = R(SH_C2[3]) * z; G[7][2] = R(SH_C2[3]) * x;
  G[8][0] = R(2) * R(SH_C2[4]) * x; G[8][1] = -R(2) * R(SH_C2[4]) * y;
  if (deg < 3) return;
  R xx = x * x, yy = y * y, zz = z * z;
  G[9][0] = R(SH_C3[0]) * R(6) * x * y;
  G[9][1] = R(SH_C3[0]) * (R(3) * xx - R(3) * yy);
  G[10][0] = R(SH_C3[1]) * y * z; G[10][1] = R(SH_C3[1]) * x * z; G[10][2] = R(SH_C3[1]) * x * y;
  G[11][0] = -R(SH_C3[2]) * R(2) * x * y;
  G[11][1] = R(SH_C3[2]) * (R(4) * zz - xx - R(3) * yy);
  G[11][2] = R(SH_C3[2]) * R(8) * y * z;
  G[12][0] = -R(SH_C3[3]) * R(6) * x * z;
  G[12][1] = -R(SH_C3[3]) * R(6) * y * z;
  G[12][2] = R(SH_C3[3]) * (R(6) * zz - R(3) * xx - R(3) * yy);
  G[13][0] = R(SH_C3[4]) * (R(4) * zz - R(3) * xx - yy);
  G[13][1] = -R(SH_C3[4]) * R(2) * x * y;
  G[13][2] = R(SH_C3[4]) * R(8) * x * z;
  G[14][0] = R(SH_C3[5]) * R(2) * x * z;
  G[14][1] = -R(SH_C3[5]) * R(2) * y * z;
  G[14][2] = R(SH_C3[5]) * (xx - yy);
  G[15][0] = R(SH_C3[6]) * (R(3) * xx - R(3) * yy);
  G[15][1] = -R(SH_C3[6]) * R(6) * x * y;
}

// Colour of Gaussian i seen from camera centre cc.  sh is [n][16][3], coefficient
// major, channel minor; entries beyond (deg+1)^2 are ignored.
template <class R>
void sh_colour(int deg, const R* sh_i, const R mean[3], const R cc[3], R rgb[3]) {
  R dir[3] = {mean[0] - cc[0], mean[1] - cc[1], mean[2] - cc[2]};
  R len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  R inv = len > R(0) ? R(1) / len : R(0);
  R d[3] = {dir[0] * inv, dir[1] * inv, dir[2] * inv};
  R Y[16];
  sh_basis(deg, d, Y);
  int K = (deg + 1) * (deg + 1);
  for (int c = 0; c < 3; ++c) {
    R acc = R(sh_i[c]);
    for (int k = 1; k < K; ++k) acc += Y[k] * R(sh_i[k * 3 + c]);
    rgb[c] = acc;
  }
}

template <class R>
void load_shape(const R* g, R q[10]) {
  // shapeOf, gpu.js:69-72.
  const int idx[10] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 11};
  for (int k = 0; k < 10; ++k) q[k] = R(g[idx[k]]);
}

// PREPROCESS.main, gpu.js:74-88 (+ SH colour when deg > 0).  Invalid splats get
// valid = 0 and zeros elsewhere (the reference leaves those fields stale).
template <class R>
void project(const CamU* cu, const R* gs, const R* sh, int deg, uint32_t n, R* splats) {
  Cam<R> cam(*cu);
  R cc[3];
  cam.centre(cc);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(n); ++i) {
    const R* g = gs + i * 16;
    R q[10];
    load_shape(g, q);
    Proj<R> p = project_shape(cam, q);
    R* o = splats + i * 12;
    for (int k = 0; k < 12; ++k) o[k] = R(0);
    if (!p.ok) continue;
    o[0] = p.u; o[1] = p.v; o[2] = p.a; o[3] = p.b; o[4] = p.c;
    o[5] = R(1) / (R(1) + std::exp(-R(g[7])));
    o[6] = p.radius; o[7] = p.depth;
    if (deg > 0 && sh) {
      R rgb[3];
      sh_colour<R>(deg, sh + i * 48, q, cc, rgb);
      o[8] = rgb[0]; o[9] = rgb[1]; o[10] = rgb[2];
    } else {
      o[8] = R(g[12]); o[9] = R(g[13]); o[10] = R(g[14]);
    }
    o[11] = R(1);
  }
}

// tileLists, gpu.js:450-481: valid splats stable-sorted by depth (index order on
// ties), each appended to every tile its f64 bounding box covers.  Returns the
// entry count; writes list (when cap allows) and ranges[2T].
template <class S>
uint32_t tile_lists(const S* splats, uint32_t n, int W, int H, uint32_t* list, uint64_t cap, uint32_t* ranges) {
  const int tilesX = (W + 15) / 16, tilesY = (H + 15) / 16;
  std::vector<uint32_t> valid;
  valid.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!(double(splats[i * 12 + 11]) > 0.5)) continue;
    if (std::isnan(double(splats[i * 12 + 7]))) continue;  // no entries either way
    valid.push_back(i);
  }
  std::stable_sort(valid.begin(), valid.end(), [&](uint32_t a, uint32_t b) {
    return double(splats[a * 12 + 7]) < double(splats[b * 12 + 7]);
  });
  auto bounds = [&](uint32_t i, double* bx) {
    double u = double(splats[i * 12]), v = double(splats[i * 12 + 1]), r = double(splats[i * 12 + 6]);
    // Math.max / Math.min propagate NaN; a NaN bound makes the loops empty.
    bx[0] = std::floor((u - r - 0.5) / 16);
    bx[1] = std::floor((u + r - 0.5) / 16);
    bx[2] = std::floor((v - r - 0.5) / 16);
    bx[3] = std::floor((v + r - 0.5) / 16);
    if (!std::isnan(bx[0])) bx[0] = std::max(0.0, bx[0]);
    if (!std::isnan(bx[1])) bx[1] = std::min(double(tilesX - 1), bx[1]);
    if (!std::isnan(bx[2])) bx[2] = std::max(0.0, bx[2]);
    if (!std::isnan(bx[3])) bx[3] = std::min(double(tilesY - 1), bx[3]);
  };
  const size_t T = size_t(tilesX) * tilesY;
  std::vector<uint64_t> counts(T, 0);
  for (uint32_t i : valid) {
    double b[4];
    bounds(i, b);
    if (!(b[0] <= b[1]) || !(b[2] <= b[3])) continue;
    for (int y = int(b[2]); y <= int(b[3]); ++y)
      for (int x = int(b[0]); x <= int(b[1]); ++x) counts[size_t(y) * tilesX + x]++;
  }
  uint64_t total = 0;
  std::vector<uint64_t> cursor(T);
  for (size_t t = 0; t < T; ++t) {
    ranges[t * 2] = uint32_t(total);
    cursor[t] = total;
    total += counts[t];
    ranges[t * 2 + 1] = uint32_t(total);
  }
  if (list && total <= cap) {
    for (uint32_t i : valid) {
      double b[4];
      bounds(i, b);
      if (!(b[0] <= b[1]) || !(b[2] <= b[3])) continue;
      for (int y = int(b[2]); y <= int(b[3]); ++y)
        for (int x = int(b[0]); x <= int(b[1]); ++x) list[cursor[size_t(y) * tilesX + x]++] = i;
    }
  }
  return uint32_t(total);
}

// alphaAt, gpu.js:92-103.  Returns false on a miss.
template <class R>
inline bool alpha_at(const R* o, R px, R py, R& alpha, R& dx, R& dy) {
  if (o[11] < R(0.5)) return false;
  dx = px - o[0];
  dy = py - o[1];
  R r = o[6];
  if (std::fabs(dx) > r || std::fabs(dy) > r) return false;
  R power = R(-0.5) * (o[2] * dx * dx + o[4] * dy * dy) - o[3] * dx * dy;
  if (power > R(0)) return false;
  alpha = std::min(R(0.99), o[5] * std::exp(power));
  if (alpha < R(1) / R(255)) return false;
  return true;
}

// For each (tile, splat) pair: the largest opacity * exp(power) of alphaAt
// (gpu.js:92-103, clamped to 0.99 as there) over the tile's pixels that pass its
// square and power <= 0 tests, or 0 when none does.  Below 1/255 the splat hits
// no pixel of the tile.  Float64 from the float32 records.
void tile_reach(const float* splats, uint64_t pairs, const uint32_t* tiles, const uint32_t* gidx, int W, int H,
                double* out) {
  const int tilesX = (W + 15) / 16;
#pragma omp parallel for schedule(dynamic, 4096)
  for (long k = 0; k < long(pairs); ++k) {
    const float* o = splats + size_t(gidx[k]) * 12;
    const int tx = int(tiles[k] % uint32_t(tilesX)), ty = int(tiles[k] / uint32_t(tilesX));
    const double u = o[0], v = o[1], A = o[2], B = o[3], C = o[4], op = o[5], r = o[6];
    double best = -1e300;
    bool any = false;
    for (int y = ty * 16; y < std::min(H, ty * 16 + 16); ++y) {
      const double dy = (double(y) + 0.5) - v;
      if (std::fabs(dy) > r) continue;
      for (int x = tx * 16; x < std::min(W, tx * 16 + 16); ++x) {
        const double dx = (double(x) + 0.5) - u;
        if (std::fabs(dx) > r) continue;
        const double power = -0.5 * (A * dx * dx + C * dy * dy) - B * dx * dy;
        if (power > 0.0) continue;
        if (!any || power > best) best = power;
        any = true;
      }
    }
    out[k] = (any && o[11] >= 0.5f) ? std::min(0.99, op * std::exp(best)) : 0.0;
  }
}

// FORWARD.main, gpu.js:113-137.  image HWC RGB, state (T, last) per pixel.
template <class R>
void forward(const R* splats, const uint32_t* list, const uint32_t* ranges, int W, int H, R* image, R* state) {
  const int tilesX = (W + 15) / 16;
#pragma omp parallel for schedule(dynamic, 16)
  for (long pix = 0; pix < long(W) * H; ++pix) {
    int x = int(pix % W), y = int(pix / W);
    R px = R(x) + R(0.5), py = R(y) + R(0.5);
    int tile = (y / 16) * tilesX + x / 16;
    uint32_t start = ranges[tile * 2], end = ranges[tile * 2 + 1];
    R T = R(1), rgb[3] = {R(0), R(0), R(0)};
    uint32_t last = start;
    for (uint32_t k = start; k < end; ++k) {
      if (T <= R(1e-4)) break;
      const R* o = splats + size_t(list[k]) * 12;
      R alpha, dx, dy;
      if (!alpha_at(o, px, py, alpha, dx, dy)) continue;
      for (int c = 0; c < 3; ++c) rgb[c] += o[8 + c] * alpha * T;
      T *= R(1) - alpha;
      last = k + 1;
    }
    for (int c = 0; c < 3; ++c) image[pix * 3 + c] = rgb[c];
    if (state) {
      state[pix * 2] = T;
      state[pix * 2 + 1] = R(last);
    }
  }
}

template <class R>
inline R sgn(R v) { return v > R(0) ? R(1) : (v < R(0) ? R(-1) : R(0)); }

// SSIM window weights per axis: win 7 is the reference's 7x7 box (1/7 each, so
// 1/49 per tap, gpu.js:360); win 11 an 11x11 Gaussian with sigma 1.5, g
// normalised to sum 1 (extension).
template <class R>
std::vector<R> window_weights(int win) {
  std::vector<R> w(win);
  if (win == 7) {
    for (int k = 0; k < win; ++k) w[k] = R(1) / R(7);
  } else {
    R sum = 0;
    for (int k = 0; k < win; ++k) {
      R d = R(k - win / 2);
      w[k] = std::exp(-d * d / (R(2) * R(1.5) * R(1.5)));
      sum += w[k];
    }
    for (int k = 0; k < win; ++k) w[k] /= sum;
  }
  return w;
}

// SSIM_COEF (gpu.js:342-372) and LOSS_GRAD (gpu.js:383-409) with lambda > 0;
// plain L1 when lambda == 0 (gpu.js:750).  Window win (7: the reference's box,
// 11: Gaussian): per valid window, dSSIM/dx_p = w_p (a + 2 b x_p + d y_p) with
// the weighted moments; the gradient sums that over the windows holding p.
template <class R>
void loss_grad(const R* image, const R* target, int W, int H, R lambda, R* grad, int win = 7) {
  const size_t P = size_t(W) * H;
  const int rad = win / 2;
  const std::vector<R> wt = window_weights<R>(win);
  std::vector<R> coef;
  const R windows = R(std::max(0, W - 2 * rad) * std::max(0, H - 2 * rad) * 3);
  if (lambda > R(0)) {
    coef.assign(P * 9, R(0));
#pragma omp parallel for schedule(static)
    for (long pix = 0; pix < long(P); ++pix) {
      int x = int(pix % W), y = int(pix / W);
      bool inside = x >= rad && y >= rad && x + rad < W && y + rad < H;
      for (int c = 0; c < 3; ++c) {
        R a = 0, b = 0, d = 0;
        if (inside) {
          R mx = 0, my = 0, sxx = 0, syy = 0, sxy = 0;
          for (int dy = 0; dy < win; ++dy)
            for (int dx = 0; dx < win; ++dx) {
              size_t q = (size_t(y + dy - rad) * W + x + dx - rad) * 3 + c;
              R w = wt[dy] * wt[dx];
              R xv = image[q], yv = target[q];
              mx += w * xv; my += w * yv; sxx += w * xv * xv; syy += w * yv * yv; sxy += w * xv * yv;
            }
          const R c1 = R(0.0001), c2 = R(0.0009);
          R a1 = R(2) * mx * my + c1;
          R a2 = R(2) * (sxy - mx * my) + c2;
          R b1 = mx * mx + my * my + c1;
          R b2 = (sxx - mx * mx) + (syy - my * my) + c2;
          R ssim = a1 * a2 / (b1 * b2);
          a = ssim * (R(2) * my / a1 - R(2) * my / a2 - R(2) * mx / b1 + R(2) * mx / b2);
          b = -ssim / b2;
          d = R(2) * ssim / a2;
        }
        coef[pix * 9 + c * 3] = a;
        coef[pix * 9 + c * 3 + 1] = b;
        coef[pix * 9 + c * 3 + 2] = d;
      }
    }
  }
  const R nn = R(P * 3);
#pragma omp parallel for schedule(static)
  for (long pix = 0; pix < long(P); ++pix) {
    int x0 = int(pix % W), y0 = int(pix / W);
    for (int c = 0; c < 3; ++c) {
      R xv = image[pix * 3 + c], yv = target[pix * 3 + c];
      R g = (R(1) - lambda) * sgn(xv - yv) / nn;
      if (lambda > R(0) && windows > R(0)) {
        R sa = 0, sb = 0, sd = 0;
        for (int dy = -rad; dy <= rad; ++dy)
          for (int dx = -rad; dx <= rad; ++dx) {
            int cx = x0 + dx, cy = y0 + dy;
            if (cx < 0 || cy < 0 || cx >= W || cy >= H) continue;
            size_t k = (size_t(cy) * W + cx) * 9 + c * 3;
            R w = wt[rad - dy] * wt[rad - dx];
            sa += w * coef[k]; sb += w * coef[k + 1]; sd += w * coef[k + 2];
          }
        R dS = (sa + R(2) * xv * sb + yv * sd) / windows;
        g -= lambda * dS;
      }
      grad[pix * 3 + c] = g;
    }
  }
}

// Loss value as ssim-test.html:13-30 defines it (win 7), or with the 11x11
// Gaussian window (win 11).
template <class R>
R loss_value(const R* image, const R* target, int W, int H, R lambda, int win = 7) {
  const size_t n = size_t(W) * H * 3;
  const int rad = win / 2;
  const std::vector<R> wt = window_weights<R>(win);
  R l1 = 0;
  for (size_t i = 0; i < n; ++i) l1 += std::fabs(image[i] - target[i]);
  l1 /= R(n);
  R ssim = 0, windows = 0;
  const R C1 = R(0.0001), C2 = R(0.0009);
  for (int cy = rad; cy + rad < H; ++cy)
    for (int cx = rad; cx + rad < W; ++cx)
      for (int c = 0; c < 3; ++c) {
        R mx = 0, my = 0, sxx = 0, syy = 0, sxy = 0;
        for (int dy = -rad; dy <= rad; ++dy)
          for (int dx = -rad; dx <= rad; ++dx) {
            size_t q = (size_t(cy + dy) * W + cx + dx) * 3 + c;
            R w = wt[dy + rad] * wt[dx + rad];
            R xv = image[q], yv = target[q];
            mx += w * xv; my += w * yv; sxx += w * xv * xv; syy += w * yv * yv; sxy += w * xv * yv;
          }
        ssim += ((R(2) * mx * my + C1) * (R(2) * (sxy - mx * my) + C2)) /
                ((mx * mx + my * my + C1) * (sxx - mx * mx + syy - my * my + C2));
        windows += 1;
      }
  R s = windows > 0 ? ssim / windows : R(0);
  return (R(1) - lambda) * l1 + lambda * (R(1) - s);
}

// BACKWARD.main, gpu.js:159-202: per pixel, back to front from the stored state,
// T recovered by division.  Gradients accumulate into grads[16n] (+=), in pixel
// order (no atomics: deterministic).
template <class R>
void backward(const R* splats, const uint32_t* list, const uint32_t* ranges, const R* state,
              const R* dLdpix, int W, int H, R* grads) {
  const int tilesX = (W + 15) / 16;
  for (long pix = 0; pix < long(W) * H; ++pix) {
    int x = int(pix % W), y = int(pix / W);
    R px = R(x) + R(0.5), py = R(y) + R(0.5);
    R dL[3] = {dLdpix[pix * 3], dLdpix[pix * 3 + 1], dLdpix[pix * 3 + 2]};
    int tile = (y / 16) * tilesX + x / 16;
    uint32_t start = ranges[tile * 2];
    R T = state[pix * 2];
    R behind[3] = {0, 0, 0};
    uint32_t k = uint32_t(state[pix * 2 + 1]);
    while (k > start) {
      k -= 1;
      uint32_t gi = list[k];
      const R* o = splats + size_t(gi) * 12;
      R alpha, dx, dy;
      if (!alpha_at(o, px, py, alpha, dx, dy)) continue;
      R Tk = T / (R(1) - alpha);
      R* g = grads + size_t(gi) * 16;
      g[0] += dL[0] * alpha * Tk;
      g[1] += dL[1] * alpha * Tk;
      g[2] += dL[2] * alpha * Tk;
      R dAlpha = 0;
      for (int c = 0; c < 3; ++c) dAlpha += dL[c] * (Tk * o[8 + c] - behind[c] / (R(1) - alpha));
      for (int c = 0; c < 3; ++c) behind[c] += o[8 + c] * alpha * Tk;
      T = Tk;
      if (alpha < R(0.99)) {
        R op = o[5];
        g[3] += dAlpha * (alpha / op) * op * (R(1) - op);
        R dPower = dAlpha * alpha;
        R A = o[2], B = o[3], C = o[4];
        g[4] += dPower * (A * dx + B * dy);
        g[5] += dPower * (C * dy + B * dx);
        g[6] += dPower * R(-0.5) * dx * dx;
        g[7] += dPower * -dx * dy;
        g[8] += dPower * R(-0.5) * dy * dy;
      }
    }
  }
}

// The same per-pixel backward run on every host thread, gradients added with
// atomic float adds (a CAS loop, as addF, gpu.js:148-157).  Summation order then
// depends on scheduling, as on the GPU; used to time the CPU baseline.
template <class R>
void backward_mt(const R* splats, const uint32_t* list, const uint32_t* ranges, const R* state, const R* dLdpix,
                 int W, int H, R* grads) {
  const int tilesX = (W + 15) / 16;
#pragma omp parallel for schedule(dynamic, 64)
  for (long pix = 0; pix < long(W) * H; ++pix) {
    int x = int(pix % W), y = int(pix / W);
    R px = R(x) + R(0.5), py = R(y) + R(0.5);
    R dL[3] = {dLdpix[pix * 3], dLdpix[pix * 3 + 1], dLdpix[pix * 3 + 2]};
    int tile = (y / 16) * tilesX + x / 16;
    uint32_t start = ranges[tile * 2];
    R T = state[pix * 2];
    R behind[3] = {0, 0, 0};
    uint32_t k = uint32_t(state[pix * 2 + 1]);
    while (k > start) {
      k -= 1;
      uint32_t gi = list[k];
      const R* o = splats + size_t(gi) * 12;
      R alpha, dx, dy;
      if (!alpha_at(o, px, py, alpha, dx, dy)) continue;
      R Tk = T / (R(1) - alpha);
      R g[9] = {dL[0] * alpha * Tk, dL[1] * alpha * Tk, dL[2] * alpha * Tk, 0, 0, 0, 0, 0, 0};
      R dAlpha = 0;
      for (int c = 0; c < 3; ++c) dAlpha += dL[c] * (Tk * o[8 + c] - behind[c] / (R(1) - alpha));
      for (int c = 0; c < 3; ++c) behind[c] += o[8 + c] * alpha * Tk;
      T = Tk;
      if (alpha < R(0.99)) {
        R op = o[5];
        g[3] = dAlpha * (alpha / op) * op * (R(1) - op);
        R dPower = dAlpha * alpha;
        R A = o[2], B = o[3], C = o[4];
        g[4] = dPower * (A * dx + B * dy);
        g[5] = dPower * (C * dy + B * dx);
        g[6] = dPower * R(-0.5) * dx * dx;
        g[7] = dPower * -dx * dy;
        g[8] = dPower * R(-0.5) * dy * dy;
      }
      R* out = grads + size_t(gi) * 16;
      for (int q = 0; q < 9; ++q) {
        if (g[q] == R(0)) continue;
#pragma omp atomic
        out[q] += g[q];
      }
    }
  }
}

// Analytic chain from screen-space gradients (grads[4..8] = u, v, conic A, B, C)
// to mean, log-scale and rotation: the derivative of projectShape (gpu.js:30-60)
// that SHAPE (gpu.js:297-328) estimates by central differences.  With deg > 0 the
// colour gradient grads[0..2] is also chained to the SH coefficients (sh_grad,
// [n][16][3], written) and, through the view direction, to the mean.
// shape[10n] (mean3, logScale3, rotation4) is written, not accumulated.
template <class R>
void shape_grad_one(const Cam<R>& cam, const R cc[3], const R* g, const R* sh, int deg, const R* gr,
                    R* shape_out, R* sh_out) {
  R q[10];
  load_shape(g, q);
  for (int k = 0; k < 10; ++k) shape_out[k] = R(0);
  if (sh_out)
    for (int k = 0; k < 48; ++k) sh_out[k] = R(0);
  R p[3];
  for (int i = 0; i < 3; ++i) p[i] = cam.V[i][0] * q[0] + cam.V[i][1] * q[1] + cam.V[i][2] * q[2] + cam.t[i];
  if (p[2] < R(0.2)) return;
  R C[3][3];
  cam_cov(cam, q, C);
  const R x = p[0], y = p[1], z = p[2];
  R j0[3] = {cam.fx / z, R(0), -cam.fx * x / (z * z)};
  R j1[3] = {R(0), cam.fy / z, -cam.fy * y / (z * z)};
  R Cj0[3], Cj1[3];
  for (int i = 0; i < 3; ++i) {
    Cj0[i] = C[i][0] * j0[0] + C[i][1] * j0[1] + C[i][2] * j0[2];
    Cj1[i] = C[i][0] * j1[0] + C[i][1] * j1[1] + C[i][2] * j1[2];
  }
  R a = j0[0] * Cj0[0] + j0[1] * Cj0[1] + j0[2] * Cj0[2] + R(0.3);
  R b = j0[0] * Cj1[0] + j0[1] * Cj1[1] + j0[2] * Cj1[2];
  R c = j1[0] * Cj1[0] + j1[1] * Cj1[1] + j1[2] * Cj1[2] + R(0.3);
  R D = a * c - b * b;
  if (D <= R(0)) return;
  const R gu = gr[4], gv = gr[5], gA = gr[6], gB = gr[7], gC = gr[8];
  // Conic (A, B, C) = (c, -b, a) / D, D = ac - b^2.
  R D2 = D * D;
  R ga = gA * (-c * c / D2) + gB * (b * c / D2) + gC * (-b * b / D2);
  R gb = gA * (R(2) * b * c / D2) + gB * (R(-1) / D - R(2) * b * b / D2) + gC * (R(2) * a * b / D2);
  R gc = gA * (-b * b / D2) + gB * (a * b / D2) + gC * (-a * a / D2);
  // a = j0' C j0, b = j0' C j1, c = j1' C j1  ->  dL/dC (symmetric).
  R G[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      G[i][j] = ga * j0[i] * j0[j] + gb * R(0.5) * (j0[i] * j1[j] + j1[i] * j0[j]) + gc * j1[i] * j1[j];
  // dL/dj0, dL/dj1.
  R gj0[3], gj1[3];
  for (int i = 0; i < 3; ++i) {
    gj0[i] = R(2) * ga * Cj0[i] + gb * Cj1[i];
    gj1[i] = R(2) * gc * Cj1[i] + gb * Cj0[i];
  }
  // Camera-space position p -> u, v, j0, j1.
  R gp[3];
  gp[0] = gu * cam.fx / z + gj0[2] * (-cam.fx / (z * z));
  gp[1] = gv * cam.fy / z + gj1[2] * (-cam.fy / (z * z));
  gp[2] = gu * (-cam.fx * x / (z * z)) + gv * (-cam.fy * y / (z * z)) + gj0[0] * (-cam.fx / (z * z)) +
          gj0[2] * (R(2) * cam.fx * x / (z * z * z)) + gj1[1] * (-cam.fy / (z * z)) +
          gj1[2] * (R(2) * cam.fy * y / (z * z * z));
  // World-space covariance gradient: Cw-grad = V^T G V.
  R T1[3][3], Gw[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) T1[i][j] = cam.V[0][i] * G[0][j] + cam.V[1][i] * G[1][j] + cam.V[2][i] * G[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Gw[i][j] = T1[i][0] * cam.V[0][j] + T1[i][1] * cam.V[1][j] + T1[i][2] * cam.V[2][j];
  // Sigma = M M^T, M = Rq diag(s).  dL/dM = 2 Gw M.
  R s[3] = {std::exp(q[3]), std::exp(q[4]), std::exp(q[5])};
  R Rm[3][3], qlen;
  quat_to_mat(q + 6, Rm, &qlen);
  R M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = Rm[i][j] * s[j];
  R gM[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gM[i][j] = R(2) * (Gw[i][0] * M[0][j] + Gw[i][1] * M[1][j] + Gw[i][2] * M[2][j]);
  // log-scale: dL/dlogs_j = s_j * sum_i gM_ij R_ij.
  for (int j = 0; j < 3; ++j) shape_out[3 + j] = s[j] * (gM[0][j] * Rm[0][j] + gM[1][j] * Rm[1][j] + gM[2][j] * Rm[2][j]);
  // Rotation: dL/dR_ij = gM_ij s_j, then through the normalised quaternion.
  R gR[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gR[i][j] = gM[i][j] * s[j];
  R w = q[6] / qlen, qx = q[7] / qlen, qy = q[8] / qlen, qz = q[9] / qlen;
  R gn[4];
  gn[0] = R(2) * (-qz * gR[0][1] + qy * gR[0][2] + qz * gR[1][0] - qx * gR[1][2] - qy * gR[2][0] + qx * gR[2][1]);
  gn[1] = R(2) * (qy * gR[0][1] + qz * gR[0][2] + qy * gR[1][0] - R(2) * qx * gR[1][1] - w * gR[1][2] + qz * gR[2][0] +
                  w * gR[2][1] - R(2) * qx * gR[2][2]);
  gn[2] = R(2) * (-R(2) * qy * gR[0][0] + qx * gR[0][1] + w * gR[0][2] + qx * gR[1][0] + qz * gR[1][2] - w * gR[2][0] +
                  qz * gR[2][1] - R(2) * qy * gR[2][2]);
  gn[3] = R(2) * (-R(2) * qz * gR[0][0] - w * gR[0][1] + qx * gR[0][2] + w * gR[1][0] - R(2) * qz * gR[1][1] +
                  qy * gR[1][2] + qx * gR[2][0] + qy * gR[2][1]);
  // n = q / |q|: dL/dq = (gn - n (n . gn)) / |q|  (when |q| > 1e-12).
  R nq[4] = {w, qx, qy, qz};
  R dotn = nq[0] * gn[0] + nq[1] * gn[1] + nq[2] * gn[2] + nq[3] * gn[3];
  R qn = std::sqrt(q[6] * q[6] + q[7] * q[7] + q[8] * q[8] + q[9] * q[9]);
  for (int k = 0; k < 4; ++k) shape_out[6 + k] = qn > R(1e-12) ? (gn[k] - nq[k] * dotn) / qlen : gn[k] / qlen;
  // Mean: dL/dmu = V^T gp.
  for (int j = 0; j < 3; ++j) shape_out[j] = cam.V[0][j] * gp[0] + cam.V[1][j] * gp[1] + cam.V[2][j] * gp[2];
  // SH colour chain (deg > 0): colour_c = dc_c + sum_k Y_k(d) f_kc, d = (mu - cc)/|mu - cc|.
  if (deg > 0 && sh && sh_out) {
    R dir[3] = {q[0] - cc[0], q[1] - cc[1], q[2] - cc[2]};
    R len = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (len > R(0)) {
      R d[3] = {dir[0] / len, dir[1] / len, dir[2] / len};
      R Y[16], GY[16][3];
      sh_basis(deg, d, Y);
      sh_basis_grad(deg, d, GY);
      int K = (deg + 1) * (deg + 1);
      R gd[3] = {0, 0, 0};
      for (int cch = 0; cch < 3; ++cch) {
        R gcol = gr[cch];
        sh_out[cch] = gcol;
        for (int k = 1; k < K; ++k) {
          sh_out[k * 3 + cch] = Y[k] * gcol;
          for (int e = 0; e < 3; ++e) gd[e] += gcol * R(sh[k * 3 + cch]) * GY[k][e];
        }
      }
      // d = dir/len: dd/ddir = (I - d d^T)/len.
      R dd = d[0] * gd[0] + d[1] * gd[1] + d[2] * gd[2];
      for (int e = 0; e < 3; ++e) shape_out[e] += (gd[e] - d[e] * dd) / len;
    } else {
      for (int cch = 0; cch < 3; ++cch) sh_out[cch] = gr[cch];
    }
  }
}

template <class R>
void shape_grad(const CamU* cu, const R* gs, const R* sh, int deg, uint32_t n, const R* grads, R* shape,
                R* sh_grad) {
  Cam<R> cam(*cu);
  R cc[3];
  cam.centre(cc);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(n); ++i)
    shape_grad_one(cam, cc, gs + i * 16, sh ? sh + i * 48 : nullptr, deg, grads + i * 16, shape + i * 10,
                   sh_grad ? sh_grad + i * 48 : nullptr);
}

// SHAPE.main, gpu.js:308-328: central differences of projectShape, h = 1e-3
// (2e-3 for the rotation), zero when either side is rejected.
template <class R>
void shape_fd(const CamU* cu, const R* gs, uint32_t n, const R* grads, R* shape) {
  Cam<R> cam(*cu);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < long(n); ++i) {
    const R* gr = grads + i * 16;
    R base[10];
    load_shape(gs + i * 16, base);
    for (int k = 0; k < 10; ++k) {
      R h = k >= 6 ? R(2e-3) : R(1e-3);
      R up[10], down[10];
      std::memcpy(up, base, sizeof(up));
      std::memcpy(down, base, sizeof(down));
      up[k] = base[k] + h;
      down[k] = base[k] - h;
      Proj<R> A = project_shape(cam, up), B = project_shape(cam, down);
      R d = 0;
      if (A.ok && B.ok)
        d = (gr[4] * (A.u - B.u) + gr[5] * (A.v - B.v) + gr[6] * (A.a - B.a) + gr[7] * (A.b - B.b) +
             gr[8] * (A.c - B.c)) / (R(2) * h);
      shape[i * 10 + k] = d;
    }
  }
}

// Adam over the 59 numbers of an SH3 Gaussian (the extension of gpu.js:411-443):
// parameters mean 0-2, logScale 3-5, rotation 6-9, opacity logit 10, SH 11-58.
// Moments (120 n): SH m [n][48] at 0 and v at 48 n; base parameter k < 11 m at
// 96 n + k n + i and v at 107 n + k n + i.  The reference's update per
// parameter, then the quaternion renormalisation; SH coefficients unclamped.
template <class R>
void adam_sh(const StepU* st, float lr_sh_rest, float* gs, float* sh, const R* grads, const R* shape,
             const R* sh_grad, R* moments) {
  static const int SLOT[11] = {0, 1, 2, 4, 5, 6, 8, 9, 10, 11, 7};
  const uint32_t n = st->count;
  const R t = R(st->t);
  const R bc1 = R(1) - std::pow(R(0.9), t), bc2 = R(1) - std::pow(R(0.999), t);
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 59; ++k) {
      R g, lr;
      float* p;
      R* mk = k < 11 ? moments + size_t(96) * n + size_t(k) * n + i : moments + size_t(i) * 48 + (k - 11);
      R* vk = k < 11 ? moments + size_t(107) * n + size_t(k) * n + i : moments + size_t(48) * n + size_t(i) * 48 + (k - 11);
      if (k < 10) {
        g = shape[i * 10 + k];
        lr = R(k < 3 ? st->lrMean : k < 6 ? st->lrScale : st->lrRotation);
        p = gs + i * 16 + SLOT[k];
      } else if (k == 10) {
        g = grads[i * 16 + 3];
        lr = R(st->lrOpacity);
        p = gs + i * 16 + 7;
      } else {
        g = sh_grad[i * 48 + (k - 11)];
        lr = R(k - 11 < 3 ? st->lrColor : lr_sh_rest);
        p = sh + i * 48 + (k - 11);
      }
      R m = R(0.9) * *mk + R(0.1) * g;
      R v = R(0.999) * *vk + R(0.001) * g * g;
      *mk = m;
      *vk = v;
      *p = float(R(*p) - lr * (m / bc1) / (std::sqrt(v / bc2) + R(1e-8)));
    }
    float* b = gs + i * 16;
    R qv[4] = {R(b[8]), R(b[9]), R(b[10]), R(b[11])};
    R len = std::max(std::sqrt(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2] + qv[3] * qv[3]), R(1e-12));
    for (int k = 0; k < 4; ++k) b[8 + k] = float(qv[k] / len);
  }
}

// ADAM.main, gpu.js:422-443 (14 parameters, SLOT map gpu.js:420).  gs and moments
// are updated in place.
template <class R>
void adam(const StepU* st, float* gs, const R* grads, const R* shape, R* moments) {
  static const int SLOT[14] = {12, 13, 14, 7, 0, 1, 2, 4, 5, 6, 8, 9, 10, 11};
  const uint32_t n = st->count;
  const R t = R(st->t);
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 14; ++k) {
      R g = k < 4 ? grads[i * 16 + k] : shape[i * 10 + k - 4];
      R lr = R(st->lrRotation);
      if (k < 3) lr = R(st->lrColor);
      else if (k == 3) lr = R(st->lrOpacity);
      else if (k < 7) lr = R(st->lrMean);
      else if (k < 10) lr = R(st->lrScale);
      R m = R(0.9) * moments[i * 28 + k] + R(0.1) * g;
      R v = R(0.999) * moments[i * 28 + 14 + k] + R(0.001) * g * g;
      moments[i * 28 + k] = m;
      moments[i * 28 + 14 + k] = v;
      R update = lr * (m / (R(1) - std::pow(R(0.9), t))) / (std::sqrt(v / (R(1) - std::pow(R(0.999), t))) + R(1e-8));
      gs[i * 16 + SLOT[k]] = float(R(gs[i * 16 + SLOT[k]]) - update);
    }
    float* b = gs + i * 16;
    for (int c = 12; c < 15; ++c) b[c] = std::min(1.0f, std::max(0.0f, b[c]));
    R qv[4] = {R(b[8]), R(b[9]), R(b[10]), R(b[11])};
    R len = std::max(std::sqrt(qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2] + qv[3] * qv[3]), R(1e-12));
    for (int k = 0; k < 4; ++k) b[8 + k] = float(qv[k] / len);
  }
}

// STATS.main, gpu.js:491-497.
template <class R>
void stats(uint32_t n, const R* grads, const R* splats, R* st) {
  for (uint32_t i = 0; i < n; ++i) {
    if (splats[i * 12 + 11] < R(0.5)) continue;
    R gu = grads[i * 16 + 4], gv = grads[i * 16 + 5];
    st[i * 2] += std::sqrt(gu * gu + gv * gv);
    st[i * 2 + 1] += R(1);
  }
}

}  // namespace

extern "C" {

int oracle_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

#define GSR_ORACLE_EXPORTS(R, SUF)                                                                                  \
  void oracle_project_##SUF(const CamU* cam, const R* gs, const R* sh, int deg, uint32_t n, R* splats) {       \
    project<R>(cam, gs, sh, deg, n, splats);                                                                          \
  }                                                                                                                   \
  uint32_t oracle_tile_lists_##SUF(const R* splats, uint32_t n, int W, int H, uint32_t* list, uint64_t cap,         \
                                   uint32_t* ranges) {                                                               \
    return tile_lists<R>(splats, n, W, H, list, cap, ranges);                                                        \
  }                                                                                                                   \
  void oracle_forward_##SUF(const R* splats, const uint32_t* list, const uint32_t* ranges, int W, int H, R* image,  \
                            R* state) {                                                                               \
    forward<R>(splats, list, ranges, W, H, image, state);                                                            \
  }                                                                                                                   \
  void oracle_loss_grad_##SUF(const R* image, const R* target, int W, int H, R lambda, R* grad, int win) {          \
    loss_grad<R>(image, target, W, H, lambda, grad, win);                                                            \
  }                                                                                                                   \
  R oracle_loss_value_##SUF(const R* image, const R* target, int W, int H, R lambda, int win) {                     \
    return loss_value<R>(image, target, W, H, lambda, win);                                                          \
  }                                                                                                                   \
  void oracle_backward_##SUF(const R* splats, const uint32_t* list, const uint32_t* ranges, const R* state,          \
                             const R* dL, int W, int H, R* grads) {                                                  \
    backward<R>(splats, list, ranges, state, dL, W, H, grads);                                                       \
  }                                                                                                                   \
  void oracle_backward_mt_##SUF(const R* splats, const uint32_t* list, const uint32_t* ranges, const R* state,       \
                                const R* dL, int W, int H, R* grads) {                                               \
    backward_mt<R>(splats, list, ranges, state, dL, W, H, grads);                                                    \
  }                                                                                                                   \
  void oracle_shape_grad_##SUF(const CamU* cam, const R* gs, const R* sh, int deg, uint32_t n,                        \
                               const R* grads, R* shape, R* sh_grad) {                                               \
    shape_grad<R>(cam, gs, sh, deg, n, grads, shape, sh_grad);                                                       \
  }                                                                                                                   \
  void oracle_shape_fd_##SUF(const CamU* cam, const R* gs, uint32_t n, const R* grads, R* shape) {                \
    shape_fd<R>(cam, gs, n, grads, shape);                                                                           \
  }                                                                                                                   \
  void oracle_adam_##SUF(const StepU* st, float* gs, const R* grads, const R* shape, R* moments) {                  \
    adam<R>(st, gs, grads, shape, moments);                                                                          \
  }                                                                                                                   \
  void oracle_adam_sh_##SUF(const StepU* st, float lr_sh_rest, float* gs, float* sh, const R* grads,               \
                            const R* shape, const R* sh_grad, R* moments) {                                          \
    adam_sh<R>(st, lr_sh_rest, gs, sh, grads, shape, sh_grad, moments);                                              \
  }                                                                                                                   \
  void oracle_stats_##SUF(uint32_t n, const R* grads, const R* splats, R* st) { stats<R>(n, grads, splats, st); }   \
  void oracle_sh_basis_##SUF(int deg, const R* d, R* Y) { sh_basis<R>(deg, d, Y); }                                  \
  int oracle_project_shape_##SUF(const CamU* cam, const R* q, R* out) {                                             \
    Cam<R> c(*cam);                                                                                                   \
    Proj<R> p = project_shape(c, q);                                                                                 \
    out[0] = p.u; out[1] = p.v; out[2] = p.a; out[3] = p.b; out[4] = p.c; out[5] = p.radius; out[6] = p.depth;       \
    return p.ok ? 1 : 0;                                                                                              \
  }

GSR_ORACLE_EXPORTS(float, f32)
GSR_ORACLE_EXPORTS(double, f64)

void oracle_tile_reach(const float* splats, uint64_t pairs, const uint32_t* tiles, const uint32_t* gidx, int W, int H,
                       double* out) {
  tile_reach(splats, pairs, tiles, gidx, W, H, out);
}

}  // extern "C"
