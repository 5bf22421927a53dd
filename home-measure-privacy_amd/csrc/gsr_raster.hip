// Tile blending: FORWARD (trainer/gpu.js:105-137) and the backward of BACKWARD_SG
// (gpu.js:204-293), plus the loss gradient (SSIM_COEF / LOSS_GRAD, gpu.js:336-409).
//
// One wave64 owns one 16x16 tile; each lane owns four pixels (column lane & 15,
// rows (lane >> 4) + 4 i).  The tile's list is staged 64 entries at a time in
// LDS (each lane gathers one 48-byte splat record; the next batch is prefetched
// into registers while the current one is blended), every lane then reads the
// batch back as broadcasts.  A wave stops as soon as all of its 256 pixels have
// T <= 1e-4, so saturated tiles retire early.  Staging also tests each entry's
// hit region (where alpha can reach 1/255: its bounding box, then the ellipse
// itself) against the tile's live pixels, band by band; the walk visits only
// the entries with a band left, and only those bands, without changing any
// result.
//
// The backward walks the same list front to back and recomputes transmittance
// with exactly the forward's arithmetic, so it stops at the same entry for each
// pixel.  The colour behind an entry enters only as d . behind (d = dL/dC): it
// starts as d . image and loses d . c times each entry's weight as the walk
// passes it.  Per entry, the four pixels of a lane are summed in
// registers, nine values are reduced over the wave together (cross-half lane
// swaps, then DPP within rows), and nine lanes issue one float atomic each into
// the entry's record.  The values are either the reference's nine gradient terms
// (the 16-float record of the standalone backward) or, in the batched step, nine
// pixel moments the terms are linear in, so the entry's conic and opacity are
// applied once per (Gaussian, view) by the shape kernel instead of in every
// lane: 5 lane swaps per entry instead of 8 (backward 4.44 -> 4.13 ms on the
// 32-view 1M-Gaussian step).
#include "gsr_internal.h"
#include "gsr_math.h"

namespace gsr {

namespace {

constexpr int kPix = 4;  // pixels per lane

struct Tile {
  uint32_t view;
  uint32_t x0, y0;      // pixel origin
  uint32_t width, height;
  uint32_t start, end;  // entry range (view-local)
  const uint32_t* order;
  const float* splats;
  uint64_t pixel_base;
  uint32_t gtile;       // global tile id (ranges index)
  int first, last;      // bands of the tile this item covers
};

// The tile (and bands) of item w; false for the whole-tile items of split tiles.
__device__ __forceinline__ bool load_tile(const RasterArgs& a, uint32_t w, Tile& t) {
  const uint32_t nh = a.split ? a.heavy_count[0] : 0u;
  const bool whole = w >= nh;
  const TileWork tw = whole ? a.work[w - nh] : a.heavy[w];
  const uint32_t first = work_first(tw), last = work_last(tw);
  const uint32_t view = work_view(tw);
  ViewDesc vd = a.views[view];
  t.view = view;
  t.first = int(first);
  t.last = int(last);
  uint32_t tx = tw.xy & 0xFFFFu, ty = tw.xy >> 16;
  t.x0 = tx * kTile;
  t.y0 = ty * kTile;
  t.width = vd.width;
  t.height = vd.height;
  t.gtile = vd.tile_base + ty * vd.tiles_x + tx;
  // a split tile is blended by its half-tile items only (split is indexed by tile id)
  if (whole && a.split && a.split[t.gtile]) return false;
  t.start = t.end = 0;
  t.order = a.order;
  t.splats = a.splats + vd.splat_base * 12;
  t.pixel_base = vd.pixel_base;
  return true;
}

// Points t at its entries of pass p.  A batched step whose entries overflowed the
// order storage is rerun; its ranges then point past the storage, so such
// segments read nothing.
__device__ __forceinline__ void tile_segment(const RasterArgs& a, int p, Tile& t) {
  const uint64_t eb = a.ebase[p] ? a.ebase[p][t.view] : 0;
  t.start = a.ranges[p][size_t(t.gtile) * 2];
  t.end = a.ranges[p][size_t(t.gtile) * 2 + 1];
  if (t.end < t.start || eb + t.end > a.order_capacity) t.start = t.end = 0;
  t.order = a.order + eb;
}

// Records of entries [k0, k0 + 64): lane j fetches entry k0 + j.  The loads
// are unconditional (lanes past the list read its last entry and are flagged
// out), so the registers they fill need no merge and nothing waits for them
// until the batch is staged; the order indices are loaded a batch earlier
// still, so the records' addresses are ready when they are issued.
struct Batch {
  float4 r0, r1, r2;  // (u, v, A, B), (C, opacity, radius, -), (r, g, b, valid)
  uint32_t gi;
  bool in;            // entry k0 + lane exists
};

// Order indices of entries [k0, k0 + 64), clamped to the list; k0 < t.end.
__device__ inline uint32_t fetch_index(const Tile& t, uint32_t k0, uint32_t lane) {
  return t.order[min(k0 + lane, t.end - 1u)];
}

__device__ inline Batch fetch_batch(const Tile& t, uint32_t gi, uint32_t k0, uint32_t lane) {
  Batch b;
  b.gi = gi;
  b.in = k0 + lane < t.end;
  const float4* r = reinterpret_cast<const float4*>(t.splats + size_t(gi) * 12);
  b.r0 = r[0];
  const float3 r1 = *reinterpret_cast<const float3*>(r + 1);  // the depth (w) is not blended
  b.r1 = make_float4(r1.x, r1.y, r1.z, 0.0f);
  b.r2 = r[2];
  return b;
}

// The entry's hit box, or with no_cull an unbounded one: every pixel of the tile
// is then evaluated, as alphaAt (gpu.js:91-103) does.  An invalid splat misses
// everywhere either way (gpu.js:93).
template <bool kNoCull>
__device__ inline float2 entry_extent(const Batch& b) {
  if (!b.in || b.r2.w < 0.5f) return make_float2(kNever, kNever);  // past the list, or an invalid splat
  if (kNoCull) return make_float2(__builtin_huge_valf(), __builtin_huge_valf());
  return hit_extent(b.r0.z, b.r0.w, b.r1.x, b.r1.y, b.r1.z);
}

// Per-entry conic coefficients of the exponent in base 2:
// power * log2(e) = qA dx^2 + dy (qC dy - qB dx) with qA = -A/2 log2 e,
// qB = B log2 e, qC = -C/2 log2 e.  The sign (the power > 0 test) is unchanged.
struct Conic2 {
  float qA, qB, qC;
};

__device__ inline Conic2 conic2(float A, float B, float C) {
  const float l2e = 1.4426950408889634f;
  return Conic2{-0.5f * l2e * A, l2e * B, -0.5f * l2e * C};
}

// Terms of the exponent shared by the pixels of one lane column.
struct Row {
  float qadx2;  // qA dx^2
  float nqbdx;  // -qB dx
  bool missx;   // |dx| > r
};

__device__ inline Row row_terms(const Conic2& q, float dx, float rad) {
  return Row{q.qA * dx * dx, -(q.qB * dx), fabsf(dx) > rad};
}

// alphaAt (gpu.js:92-103) without branches: alpha, or 0 for a miss or for a pixel
// already done (T <= 1e-4).  The miss tests are the reference's, negated as
// written there (a NaN takes the same path).  They are combined with bitwise
// ors, so the compiler selects instead of masking lanes around the exponential
// (the masked form cost both blends about 6%).  The forward and the backward
// call this same function, so both see the same alpha bit for bit.
__device__ inline float alpha_at(float dy, const Conic2& q, const Row& w, float op, float rad, float T) {
  const float p2 = fmaf(dy, fmaf(q.qC, dy, w.nqbdx), w.qadx2);
  const float alpha = fminf(kAlphaMax, op * __builtin_amdgcn_exp2f(p2));
  const bool miss = w.missx | (fabsf(dy) > rad) | (p2 > 0.0f) | (alpha < kAlphaMin) | !(T > kTStop);
  return miss ? 0.0f : alpha;
}

__device__ inline bool any_active(const float T[kPix]) {
  bool act = false;
#pragma unroll
  for (int i = 0; i < kPix; ++i) act |= T[i] > kTStop;
  return __ballot(act) != 0;
}

// Pixel-centre bounding box of the wave's still-active pixels (empty: x0 > x1).
struct Box {
  float x0, x1, y0, y1;
};

// From the four per-slot ballots and scalar bit scans: lane l of slot i is the
// pixel (x0 + (l & 15), y0 + (l >> 4) + 4 i), so the first / last active column
// is the lowest / highest set bit of the masks folded to 16 bits, and the first
// / last active row lies in the first / last slot with an active pixel, at its
// lowest / highest set lane.  The coordinates are the pixels' own centres
// (float(x) + 0.5, float(y0 + (l >> 4)) + 0.5 + 4 i), as the walk computes them.
__device__ inline Box active_box(uint32_t x0, uint32_t y0, const float T[kPix]) {
  uint64_t m[kPix];
  uint64_t all = 0;
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    m[i] = __ballot(T[i] > kTStop);
    all |= m[i];
  }
  if (all == 0) return Box{__builtin_huge_valf(), -__builtin_huge_valf(), __builtin_huge_valf(), -__builtin_huge_valf()};
  const uint32_t half = uint32_t(all) | uint32_t(all >> 32);
  const uint32_t cols = (half | (half >> 16)) & 0xFFFFu;
  const uint32_t c0 = uint32_t(__builtin_ctz(cols)), c1 = uint32_t(31 - __builtin_clz(cols));
  int ifirst = kPix - 1, ilast = 0;
#pragma unroll
  for (int i = kPix - 1; i >= 0; --i)
    if (m[i] != 0) ifirst = i;
#pragma unroll
  for (int i = 0; i < kPix; ++i)
    if (m[i] != 0) ilast = i;
  uint64_t mf = m[kPix - 1], ml = m[0];
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    if (i == ifirst) mf = m[i];
    if (i == ilast) ml = m[i];
  }
  const uint32_t r0 = uint32_t(__builtin_ctzll(mf)) >> 4, r1 = uint32_t(63 - __builtin_clzll(ml)) >> 4;
  Box b;
  b.x0 = float(x0 + c0) + 0.5f;
  b.x1 = float(x0 + c1) + 0.5f;
  b.y0 = (float(y0 + r0) + 0.5f) + 4.0f * float(ifirst);
  b.y1 = (float(y0 + r1) + 0.5f) + 4.0f * float(ilast);
  return b;
}

// True when the entry's hit box (|dx| <= ex, |dy| <= ey) misses every active
// pixel.  Uses alphaAt's own subtraction px - u, which is monotonic in px, so the
// test agrees with the per-pixel values at the box edges.
__device__ inline bool box_culled(const Box& b, float u, float v, float ex, float ey) {
  return (b.x1 - u) < -ex || (b.x0 - u) > ex || (b.y1 - v) < -ey || (b.y0 - v) > ey;
}

// Lane rows (lane >> 4) + 4 i form band i: tile rows 4 i .. 4 i + 3.  True when
// the entry's hit box misses the band's pixel centres [ylo, ylo + 3] (the
// subtractions are the pixels' own py - v, monotonic in py).
__device__ inline bool band_culled(float ylo, float v, float ey) {
  return (ylo + 3.0f - v) < -ey || (ylo - v) > ey;
}

// Clears the bands in which the entry's hit ellipse (-2 power <= 2 t, t from
// hit_threshold) misses the rectangle of live pixel centres: the box's columns
// and the band's rows within the box, both clipped to the hit box.  The hit box
// bounds the ellipse, so it overshoots an elongated, tilted one by a wide margin;
// this test does not.  Offsets are the blend's own subtractions px - u, py - v,
// so the rectangle holds the offsets of every live pixel of the band.  The
// bound gets a margin of twice hit_threshold's rounding term beyond it, for the
// rounding of conic_min_rect, so a band is cleared only where the blend's own
// arithmetic misses at every pixel.  Degenerate conics keep their bands.
__device__ inline uint32_t ellipse_bands(uint32_t bands, const Batch& b, float band0, const Box& box, float2 e) {
  const float A = b.r0.z, B = b.r0.w, C = b.r1.x, u = b.r0.x, v = b.r0.y, r = b.r1.z;
  const float t = hit_threshold(A, B, C, b.r1.y, r);
  const float bb = B * B;
  const float det = fmaf(A, C, -bb) - fmaf(B, B, -bb);
  if (bands == 0 || !(t > 0.0f) || !(det > 0.0f) || !(A > 0.0f) || !(C > 0.0f)) return bands;
  const float r1 = r + 1.0f;
  const float lim = 2.0f * t + 4e-6f * (fabsf(A) + 2.0f * fabsf(B) + fabsf(C)) * r1 * r1;
  const float x0 = fmaxf(box.x0 - u, -e.x), x1 = fminf(box.x1 - u, e.x);
  const float by0 = fmaxf(box.y0 - v, -e.y), by1 = fminf(box.y1 - v, e.y);
#pragma unroll 1
  for (int i = 0; i < kPix; ++i) {
    const float y0 = fmaxf((band0 + 4.0f * i) - v, by0), y1 = fminf((band0 + 4.0f * i + 3.0f) - v, by1);
    const bool miss = !(y0 <= y1) || conic_min_rect(A, B, C, x0, x1, y0, y1) > lim;
    bands &= miss ? ~(1u << i) : ~0u;
  }
  return bands;
}

// LDS layout of a staged entry: (u, v, A, B), (C, opacity, radius, ex), (r, g, b,
// ey) and (qA, qB, qC, bands).  The fourth quad holds what depends only on the
// entry and the tile, computed once per entry by its own lane rather than by the
// whole wave in the walk: the base-2 conic and the bands of the tile (bit i: rows
// 4 i .. 4 i + 3, pixel centres band0 + 4 i ..) that the hit box reaches.  A band
// whose bit is clear holds no pixel the entry can hit.  With `box` (the live
// pixels when the batch is staged), an entry whose hit box misses it gets no
// bands: pixels still live later in the batch lie inside that box, so the walk
// skips such entries.  (Moving these tests out of the walk's serial chain of
// compares and branches took the forward from 6.4 to 5.7 ms and the backward
// from 6.4 to 6.0 ms on the 32-view 1M-Gaussian step.)  Returns the bands, which
// the lane keeps: the walk visits only the entries whose bits are set (a ballot)
// and takes each one's bits from its lane with readlane, so an entry costs no
// LDS round trip before its record is read, and a culled one costs nothing
// (forward 5.5 -> 4.2 ms, backward 5.1 -> 4.8 ms).
template <bool kNoCull>
__device__ inline uint32_t stage(float4 (*lds)[4], uint32_t lane, const Batch& b, float band0, const Box& box) {
  const float2 e = entry_extent<kNoCull>(b);
  lds[lane][0] = b.r0;
  lds[lane][1] = make_float4(b.r1.x, b.r1.y, b.r1.z, e.x);
  lds[lane][2] = make_float4(b.r2.x, b.r2.y, b.r2.z, e.y);
  const Conic2 q = conic2(b.r0.z, b.r0.w, b.r1.x);
  uint32_t bands = 0;
#pragma unroll
  for (int i = 0; i < kPix; ++i) bands |= band_culled(band0 + 4.0f * i, b.r0.y, e.y) ? 0u : (1u << i);
  if (box_culled(box, b.r0.x, b.r0.y, e.x, e.y)) bands = 0;
  if (!kNoCull) bands = ellipse_bands(bands, b, band0, box, e);
  lds[lane][3] = make_float4(q.qA, q.qB, q.qC, __uint_as_float(bands));
  return bands;
}

// Bands [first, last] of a work item as a bit mask.
__device__ inline uint32_t item_bands(const Tile& t) {
  return ((1u << (t.last + 1)) - 1u) & ~((1u << t.first) - 1u);
}

// ---- forward -------------------------------------------------------------------

// kLast: also track each pixel's last hit (the optional state output and the
// profiling counters need it; the batched training step does not).  kNoCull:
// evaluate every entry at every pixel (RasterArgs::no_cull), a kernel of its own
// so the regular blend keeps its register budget.
// Inlined into the kernels, so the kernel arguments are read in place.
template <bool kLast, bool kNoCull>
__device__ __forceinline__ void forward_tile(const RasterArgs& a, uint32_t w, float4 (*lds)[4]) {
  const uint32_t lane = __lane_id();
  Tile t;
  if (!load_tile(a, w, t)) return;
  const bool resume = a.pass > 0;
  if (resume && a.done[t.gtile]) return;  // saturated by an earlier pass
  tile_segment(a, a.pass, t);
  const uint32_t x = t.x0 + (lane & 15);
  const float px = float(x) + 0.5f;
  const float py0 = float(t.y0 + (lane >> 4)) + 0.5f;  // row of pixel i: py0 + 4 i
  float T[kPix], cr[kPix], cg[kPix], cb[kPix];
  uint32_t last[kPix];
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    uint32_t y = t.y0 + (lane >> 4) + 4 * i;
    bool inside = x < t.width && y < t.height && i >= t.first && i <= t.last;
    T[i] = inside ? 1.0f : 0.0f;  // pixels outside the image or the item count as done
    cr[i] = cg[i] = cb[i] = 0.0f;
    last[i] = t.start;
    if (resume && inside) {
      size_t pix = t.pixel_base + size_t(y) * t.width + x;
      T[i] = a.tstate[pix];
      cr[i] = a.image[pix * 3];
      cg[i] = a.image[pix * 3 + 1];
      cb[i] = a.image[pix * 3 + 2];
    }
  }
  const float band0 = float(t.y0) + 0.5f;
  const uint32_t items = item_bands(t);
  Batch nb{};
  uint32_t ngi = 0;  // order indices of the batch after the next
  if (t.start < t.end) {
    nb = fetch_batch(t, fetch_index(t, t.start, lane), t.start, lane);
    ngi = fetch_index(t, t.start + 64, lane);
  }
  uint32_t walked = t.start;  // kLast only: up to the entry at which the walk stopped
  uint32_t stop = t.end;      // start of the batch at which the tile had no active pixel
  for (uint32_t k0 = t.start; k0 < t.end; k0 += 64) {
    if (!any_active(T)) {
      stop = k0;
      break;
    }
    const Box sbox = active_box(t.x0, t.y0, T);
    __syncthreads();
    const uint32_t sb = stage<kNoCull>(lds, lane, nb, band0, sbox) & items;
    __syncthreads();
    // Unconditional: past the end the loads read the list's last entry and the
    // batch is flagged out (and never staged).
    nb = fetch_batch(t, ngi, k0 + 64, lane);
    ngi = fetch_index(t, k0 + 128, lane);
    uint64_t live = __ballot(sb != 0u);
    uint32_t seen = 0;
    if (kLast) walked = min(k0 + 64u, t.end);
    while (live != 0) {
      const uint32_t j = uint32_t(__builtin_ctzll(live));
      live &= live - 1;
      if ((++seen & 15u) == 0u && !any_active(T)) {
        if (kLast) walked = k0 + j;
        break;
      }
      const uint32_t bands = __builtin_amdgcn_readlane(sb, j);
      const float4 r0 = lds[j][0], r1 = lds[j][1], r2 = lds[j][2], r3 = lds[j][3];
      const float u = r0.x, v = r0.y, op = r1.y, rad = r1.z;
      const Conic2 q{r3.x, r3.y, r3.z};
      const Row row = row_terms(q, px - u, rad);
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        // A band outside the hit box or the item, a miss or a done pixel blends
        // with weight 0: colour and T unchanged.  (State updated in place per
        // band: this form compiles without copies of the pixel state.)
        if (((bands >> i) & 1u) == 0) continue;
        const float alpha = alpha_at((py0 + 4.0f * i) - v, q, row, op, rad, T[i]);
        const float wgt = alpha * T[i];
        cr[i] = fmaf(r2.x, wgt, cr[i]);
        cg[i] = fmaf(r2.y, wgt, cg[i]);
        cb[i] = fmaf(r2.z, wgt, cb[i]);
        T[i] = T[i] * (1.0f - alpha);
        if (kLast) last[i] = alpha != 0.0f ? k0 + j + 1 : last[i];
      }
    }
  }
  // Entries examined in this pass (the backward's work estimate; exact when
  // profiling, else to the batch) and, when profiling, the furthest last hit:
  // maxima over the items of the tile.  (Exact counting stays out of the
  // regular blend, which then runs with fewer registers and more waves.)  Then
  // which bands stay open: a whole-tile item is its tile's only writer; band
  // items of a split tile write one byte per band.
  const bool whole = t.first == 0 && t.last == kPix - 1;
  if (a.walk && lane == 0) {
    const uint32_t n = (kLast ? walked : stop) - t.start;
    if (whole) a.walk[t.gtile] = n;
    else atomicMax(a.walk + t.gtile, n);
  }
  if (kLast && a.last) {
    uint32_t m = t.start;
#pragma unroll
    for (int i = 0; i < kPix; ++i) m = max(m, last[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), off)));
    if (lane == 0 && m > t.start) atomicMax(a.last + t.gtile, m - t.start);
  }
  if (a.band_open) {
    if (whole) {
      const bool open = any_active(T);
      if (lane == 0) reinterpret_cast<uint32_t*>(a.band_open)[t.gtile] = open ? 0x01010101u : 0u;
    } else {
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        const bool open = __ballot(T[i] > kTStop) != 0;
        if (lane == 0 && i >= t.first && i <= t.last) a.band_open[size_t(t.gtile) * 4 + i] = open ? 1 : 0;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    uint32_t y = t.y0 + (lane >> 4) + 4 * i;
    if (x >= t.width || y >= t.height || i < t.first || i > t.last) continue;
    size_t pix = t.pixel_base + size_t(y) * t.width + x;
    a.image[pix * 3] = cr[i];
    a.image[pix * 3 + 1] = cg[i];
    a.image[pix * 3 + 2] = cb[i];
    if (a.tstate) a.tstate[pix] = T[i];
    if (kLast && a.state) {
      a.state[pix * 2] = T[i];
      a.state[pix * 2 + 1] = float(last[i]);
    }
  }
}

// kLast (the optional state output and the profiling counters) is a kernel of
// its own, so the regular blend keeps its smaller register budget.
template <bool kLast, bool kNoCull>
__global__ __launch_bounds__(64) void forward_kernel(RasterArgs a) {
  __shared__ float4 lds[64][4];
  forward_tile<kLast, kNoCull>(a, blockIdx.x, lds);
}

__device__ inline uint32_t pop_work(uint32_t* queue, uint32_t n) {
  uint32_t w0 = 0;
  if (__lane_id() == 0) w0 = atomicAdd(queue, n);
  return __builtin_amdgcn_readfirstlane(w0);
}

// Work distribution of the persistent kernels.  Workgroups are dealt to the
// chip's kXcds dies in turn (workgroup i on die i mod kXcds), each die with its
// own L2, and neighbouring tiles of a view share most of their splat records.
// So the whole tiles are cut into kXcds contiguous parts, part d served first to
// the workgroups of die d from its own counter; a die that runs out helps the
// others in turn.  The split tiles' items come first, from a shared counter.
// queue: [0] the split items' counter, [kQueueStride (1 + d)] part d's.
constexpr uint32_t kXcds = 8;

template <class Blend>
__device__ __forceinline__ void for_each_item(const RasterArgs& a, Blend&& blend) {
  const uint32_t nh = a.split ? a.heavy_count[0] : 0u;
  const uint32_t num_work = a.work_count ? *a.work_count : a.num_work;
  for (;;) {
    const uint32_t w0 = pop_work(a.queue, a.tile_batch);
    if (w0 >= nh) break;
    const uint32_t w1 = min(nh, w0 + a.tile_batch);
    for (uint32_t w = w0; w < w1; ++w) blend(w);
  }
  const uint32_t home = blockIdx.x % kXcds;
  for (uint32_t k = 0; k < kXcds; ++k) {
    const uint32_t part = (home + k) % kXcds;
    const uint32_t lo = nh + uint32_t(uint64_t(num_work) * part / kXcds);
    const uint32_t hi = nh + uint32_t(uint64_t(num_work) * (part + 1) / kXcds);
    for (;;) {
      const uint32_t w0 = lo + pop_work(a.queue + kQueueStride * (1 + part), a.tile_batch);
      if (w0 >= hi) break;
      const uint32_t w1 = min(hi, w0 + a.tile_batch);
      for (uint32_t w = w0; w < w1; ++w) blend(w);
    }
  }
}

// Six waves per SIMD (80 VGPRs, four spilled) for the regular blend: its walk waits
// on LDS and memory a third of its time (forward 9.34 -> 8.98 ms on the 3M-Gaussian
// training scene, 2.62 -> 2.59 ms on the 32-view bench scene); the profiling
// variant would spill 44 and keeps the compiler's choice.
template <bool kLast, bool kNoCull>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kLast ? 1 : 6))) void forward_persistent_kernel(
    RasterArgs a) {
  __shared__ float4 lds[64][4];
  for_each_item(a, [&](uint32_t w) __attribute__((always_inline)) { forward_tile<kLast, kNoCull>(a, w, lds); });
}

// ---- backward ------------------------------------------------------------------

template <int CTRL>
__device__ inline float dpp_src(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over each row of 16 lanes, left in every lane of the row.
__device__ inline float row_sum(float v) {
  v += dpp_src<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_src<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_src<0x124>(v);  // row_ror:4
  v += dpp_src<0x128>(v);  // row_ror:8
  return v;
}

// Lanes l and l + 32 added: lanes 0-31 receive a's sums, lanes 32-63 b's.
__device__ inline float fold32(float a, float b) {
  const auto s = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}

// Rows 0+1 and 2+3 added: rows 0 and 2 receive a's sums, rows 1 and 3 b's.
__device__ inline float fold16(float a, float b) {
  const auto s = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(s[0]) + __uint_as_float(s[1]);
}

// Wave sums of the nine gradient terms by transposition: two folds halve the
// lanes while pairing values, then rows of 16 are summed.  Lane (16 row + col)
// ends with the total of term reduce9_slot(lane) (col 0: terms 0 2 1 3 by row,
// col 1: 4 6 5 7, lane 2: term 8); the result is that lane's value.
__device__ inline float reduce9(const float g[9], uint32_t lane) {
  const float h0 = fold32(g[0], g[1]), h1 = fold32(g[2], g[3]);
  const float h2 = fold32(g[4], g[5]), h3 = fold32(g[6], g[7]);
  const float h4 = fold32(g[8], 0.0f);
  const float k0 = row_sum(fold16(h0, h1));
  const float k1 = row_sum(fold16(h2, h3));
  const float k2 = row_sum(fold16(h4, 0.0f));
  const uint32_t col = lane & 15;
  return col == 0 ? k0 : (col == 1 ? k1 : k2);
}

// Gradient slot summed into lane `lane` by reduce9, or -1 for none.
__device__ inline int reduce9_slot(uint32_t lane) {
  const uint32_t row = lane >> 4, col = lane & 15;
  const int by_row = int((row & 1) * 2 + (row >> 1));  // rows 0 1 2 3 -> 0 2 1 3
  if (col == 0) return by_row;
  if (col == 1) return 4 + by_row;
  return lane == 2 ? 8 : -1;
}

// Wave sums of the nine pixel moments (kMomentSlots) from each lane's six sums
// over its pixels (colour terms, S0 = sum dP, S1 = sum dP dy, S2 = sum dP dy^2).
// The lanes of a column share dx, so the x-weighted moments are formed once the
// first fold has paired rows r and r + 2: one fold32 per two raw sums, two
// multiplies for dx S0, dx S1 and dx^2 S0, then the fold16s and row sums as in
// reduce9.  dx^2 S0 skips the fold16 (it has no partner): its rows 0 and 1 are
// summed, and row 1 adds row 0's total by row_bcast:15, a DPP move at half the
// cost of a swap.  Lane (16 row + col) ends with moment moment_slot(lane).
__device__ inline float reduce_moments(float gc0, float gc1, float gc2, float s0, float s1, float s2, float dx,
                                       uint32_t lane) {
  const float h0 = fold32(gc0, gc1), h1 = fold32(gc2, s2), h2 = fold32(s0, s1);
  const float h3 = dx * h2;  // lanes 0-31: dx S0, lanes 32-63: dx S1
  const float h4 = dx * h3;  // lanes 0-31: dx^2 S0 (lanes 32-63 unused)
  const float k0 = row_sum(fold16(h0, h1));  // rows: colour 0, colour 2, colour 1, S2
  const float k1 = row_sum(fold16(h2, h3));  // rows: S0, dx S0, S1, dx S1
  const float r2 = row_sum(h4);
  // row_bcast:15 (0x142) into row 1 only (row mask 0x2): lane 15's value.
  const float k2 = r2 + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(r2), 0x142, 0x2, 0xF, false));
  const uint32_t col = lane & 15;
  return col == 0 ? k0 : (col == 1 ? k1 : k2);  // k2: dx^2 S0 in row 1
}

// Moment slot summed into lane `lane` by reduce_moments, or -1 for none.
__device__ inline int moment_slot(uint32_t lane) {
  const uint32_t row = lane >> 4, col = lane & 15;
  constexpr int kCol0[4] = {0, 2, 1, 8}, kCol1[4] = {3, 4, 5, 7};
  if (col == 0) return kCol0[row];
  if (col == 1) return kCol1[row];
  return lane == 18 ? 6 : -1;
}

__device__ inline float sign_f(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }

template <bool kNoCull, bool kMoments>
__device__ __forceinline__ void backward_tile(const RasterArgs& a, uint32_t w, float4 (*lds)[4], uint64_t* rec_lds) {
  const uint32_t lane = __lane_id();
  Tile t;
  if (!load_tile(a, w, t)) return;
  const uint32_t x = t.x0 + (lane & 15);
  const float px = float(x) + 0.5f;
  const int slot = kMoments ? moment_slot(lane) : reduce9_slot(lane);
  // Per pixel: dL/dC (d) and bd = d . (the colour behind the entries walked so
  // far): d . image less d . c times each walked entry's weight.  The behind
  // term of gpu.js:266-268 needs the behind colour only dotted with d, and d . c
  // is needed for the gradient anyway, so one running scalar replaces the
  // image and the accumulated colour (6 values per pixel): five waves per SIMD
  // instead of four, and four fewer instructions per evaluated band (backward
  // 5.35 -> 4.46 ms on the 32-view 1M-Gaussian step over the two changes).  bd
  // differs from d . (image - accumulated) by rounding only.
  const float py0 = float(t.y0 + (lane >> 4)) + 0.5f;  // row of pixel i: py0 + 4 i
  float T[kPix], bd[kPix], dr[kPix], dg[kPix], db[kPix];
#pragma unroll
  for (int i = 0; i < kPix; ++i) {
    uint32_t y = t.y0 + (lane >> 4) + 4 * i;
    bool inside = x < t.width && y < t.height && i >= t.first && i <= t.last;
    T[i] = inside ? 1.0f : 0.0f;
    bd[i] = dr[i] = dg[i] = db[i] = 0.0f;
    if (inside) {
      size_t pix = t.pixel_base + size_t(y) * t.width + x;
      const float ir = a.image[pix * 3], ig = a.image[pix * 3 + 1], ib = a.image[pix * 3 + 2];
      if (a.l1_target) {
        // The L1 gradient (lossGradient at lambda 0, gpu.js:389-392), as
        // loss_grad_kernel computes it, without the round trip through dl.
        const float n = float(t.width * t.height * 3u);
        dr[i] = sign_f(ir - a.l1_target[pix * 3]) / n;
        dg[i] = sign_f(ig - a.l1_target[pix * 3 + 1]) / n;
        db[i] = sign_f(ib - a.l1_target[pix * 3 + 2]) / n;
      } else {
        dr[i] = a.dl[pix * 3];
        dg[i] = a.dl[pix * 3 + 1];
        db[i] = a.dl[pix * 3 + 2];
      }
      bd[i] = fmaf(dr[i], ir, fmaf(dg[i], ig, db[i] * ib));
    }
  }
  const float band0 = float(t.y0) + 0.5f;
  const uint32_t items = item_bands(t);
  float* grads = a.grads + size_t(t.view) * a.grads_view_stride;
  const uint32_t record = a.grads_record ? a.grads_record : 16u;
  // The passes of a tile follow each other in depth order; pixel state carries over.
  for (int p = 0; p < a.num_passes; ++p) {
  tile_segment(a, p, t);
  Batch nb{};
  uint32_t ngi = 0;  // order indices of the batch after the next
  if (t.start < t.end) {
    nb = fetch_batch(t, fetch_index(t, t.start, lane), t.start, lane);
    ngi = fetch_index(t, t.start + 64, lane);
  }
  for (uint32_t k0 = t.start; k0 < t.end; k0 += 64) {
    if (!any_active(T)) break;
    const Box sbox = active_box(t.x0, t.y0, T);
    __syncthreads();
    const uint32_t sb = stage<kNoCull>(lds, lane, nb, band0, sbox) & items;
    rec_lds[lane] = uint64_t(nb.gi) * record;  // the record's offset, multiplied once per batch
    __syncthreads();
    // Unconditional: past the end the loads read the list's last entry and the
    // batch is flagged out (and never staged).
    nb = fetch_batch(t, ngi, k0 + 64, lane);
    ngi = fetch_index(t, k0 + 128, lane);
    uint64_t live = __ballot(sb != 0u);
    uint32_t seen = 0;
    while (live != 0) {
      const uint32_t j = uint32_t(__builtin_ctzll(live));
      live &= live - 1;
      if ((++seen & 15u) == 0u && !any_active(T)) break;
      const uint32_t bands = __builtin_amdgcn_readlane(sb, j);
      const float4 r0 = lds[j][0], r1 = lds[j][1], r2 = lds[j][2], r3 = lds[j][3];
      const float u = r0.x, v = r0.y, A = r0.z, B = r0.w, C = r1.x, op = r1.y, rad = r1.z;
      const Conic2 q{r3.x, r3.y, r3.z};
      const float dx = px - u;
      const Row row = row_terms(q, dx, rad);
      float gc0 = 0.0f, gc1 = 0.0f, gc2 = 0.0f, s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
      bool hit = false;
#pragma unroll
      for (int i = 0; i < kPix; ++i) {
        if (((bands >> i) & 1u) == 0) continue;
        const float dy = (py0 + 4.0f * i) - v;
        // Same alpha and transmittance as forward_tile, bit for bit.
        const float alpha = alpha_at(dy, q, row, op, rad, T[i]);
        const bool h = alpha != 0.0f;
        // A miss (alpha 0) leaves T and bd bit for bit as they are, so the
        // state is updated unconditionally (this form compiles
        // without copies of the pixel state); the gradient terms only where a
        // pixel of the band was hit.
        const float Tk = T[i];
        const float wgt = alpha * Tk;
        const float oma = 1.0f - alpha;
        T[i] = Tk * oma;
        const float dc = fmaf(dr[i], r2.x, fmaf(dg[i], r2.y, db[i] * r2.z));
        bd[i] = fmaf(-dc, wgt, bd[i]);
        if (__ballot(h) == 0) continue;
        hit = true;
        // gpu.js:266-268: colour gradient and dLoss/dalpha = d . (Tk c - behind / (1 - alpha)).
        gc0 = fmaf(dr[i], wgt, gc0);
        gc1 = fmaf(dg[i], wgt, gc1);
        gc2 = fmaf(db[i], wgt, gc2);
        float dA = fmaf(Tk, dc, -(bd[i] * __builtin_amdgcn_rcpf(oma)));
        // gpu.js:270-281: opacity logit and screen (u, v, conic) gradients, only
        // while the 0.99 clamp is not active.  A missed pixel of the band has
        // alpha 0 and a finite dA (1 - alpha = 1), so its dP is zero without a
        // test of its own.
        dA = alpha < kAlphaMax ? dA : 0.0f;
        const float dP = dA * alpha;
        const float dPdy = dP * dy;
        s0 += dP;
        s1 += dPdy;
        s2 = fmaf(dPdy, dy, s2);
      }
      if (!hit) continue;  // wave-uniform: set only for bands with a hit
      float val;
      if (kMoments) {
        val = reduce_moments(gc0, gc1, gc2, s0, s1, s2, dx, lane);
      } else {
        // gpu.js:270-281 summed over the lane's pixels: with S0 = sum dP, S1 =
        // sum dP dy and S2 = sum dP dy^2 (dP = dLoss/dpower), d opLogit = (1 -
        // op) S0, du = A dx S0 + B S1, dv = B dx S0 + C S1, dA = -dx^2/2 S0, dB =
        // -dx S1, dC = -S2/2 (dx is the lane's, shared by its pixels).
        const float g[9] = {gc0, gc1, gc2, s0 * (1.0f - op), fmaf(s1, B, s0 * (A * dx)), fmaf(s1, C, s0 * (B * dx)),
                            s0 * (-0.5f * dx * dx), -(s1 * dx), -0.5f * s2};
        val = reduce9(g, lane);
      }
      if (slot >= 0 && val != 0.0f) unsafeAtomicAdd(grads + rec_lds[j] + slot, val);
    }
  }
  }
}

template <bool kNoCull, bool kMoments>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void backward_kernel(RasterArgs a) {
  __shared__ float4 lds[64][4];
  __shared__ uint64_t rec_lds[64];
  backward_tile<kNoCull, kMoments>(a, blockIdx.x, lds, rec_lds);
}

// Six waves per SIMD (80 VGPRs, 13 spilled): on the 3M-Gaussian training scene
// the step went from 42.80 to 42.23 ms (interleaved runs); the 32-view bench
// step is unchanged (9.88 ms either way).
template <bool kNoCull, bool kMoments>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void backward_persistent_kernel(RasterArgs a) {
  __shared__ float4 lds[64][4];
  __shared__ uint64_t rec_lds[64];
  for_each_item(a, [&](uint32_t w)
                    __attribute__((always_inline)) { backward_tile<kNoCull, kMoments>(a, w, lds, rec_lds); });
}

// ---- loss gradient (gpu.js:336-409) ---------------------------------------------

// SSIM windows.  kWin == 7: the reference's 7x7 box, weights 1/49 (gpu.js:342-
// 372).  kWin == 11: an 11x11 Gaussian window, sigma 1.5, w(i, j) = g(i) g(j)
// with g normalised to sum 1 (the window of the usual 3DGS loss).  Windows are
// the valid ones (wholly inside the image); the mean SSIM is over windows and
// channels, as gpu.js:747 counts them.
template <int kWin>
__device__ inline void window_weights(float w[kWin]) {
  if (kWin == 7) {
#pragma unroll
    for (int k = 0; k < kWin; ++k) w[k] = 1.0f / 7.0f;
  } else {
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kWin; ++k) {
      const float d = float(k - kWin / 2);
      w[k] = expf(-d * d / (2.0f * 1.5f * 1.5f));
      sum += w[k];
    }
#pragma unroll
    for (int k = 0; k < kWin; ++k) w[k] /= sum;
  }
}

// Both SSIM kernels work on 16 x 16 tiles of one view with their kWin - 1 halo
// staged in LDS, and filter separably: the window weights are w(i) w(j), so a
// row pass then a column pass replaces the kWin^2 taps of the direct sum by 2 kWin.
constexpr int kSsimTile = 16;

// Loads a (16 + 2R)^2 halo of `ch` interleaved channels starting at pixel
// (x0 - R, y0 - R) into lds[ch][hy][hx]; outside the image: 0.
template <int kR, int kCh>
__device__ inline void load_halo(const float* __restrict__ src, int W, int H, int x0, int y0,
                                 float (*lds)[kSsimTile + 2 * kR][kSsimTile + 2 * kR]) {
  constexpr int kHS = kSsimTile + 2 * kR;
  for (int k = threadIdx.x; k < kHS * kHS; k += blockDim.x) {
    const int hy = k / kHS, hx = k % kHS;
    const int gx = x0 + hx - kR, gy = y0 + hy - kR;
    const bool in = gx >= 0 && gy >= 0 && gx < W && gy < H;
    const size_t q = in ? (size_t(gy) * W + gx) * kCh : 0;
#pragma unroll
    for (int c = 0; c < kCh; ++c) lds[c][hy][hx] = in ? src[q + c] : 0.0f;
  }
}

// SSIM_COEF (gpu.js:342-372) for window kWin: per valid window centre and
// channel, the coefficients (a, b, d) with dSSIM_window/dx_p = w_p (a + 2 b x_p +
// d y_p), from the window's weighted moments of x, y, x^2, y^2 and x y.
template <int kWin>
__global__ __launch_bounds__(256) void ssim_coef_kernel(const ViewDesc* __restrict__ views,
                                                        const float* __restrict__ image,
                                                        const float* __restrict__ target, float* __restrict__ coef) {
  constexpr int kR = kWin / 2, kHS = kSsimTile + 2 * kR;
  __shared__ float sx[3][kHS][kHS], sy[3][kHS][kHS];
  __shared__ float rows[5][kHS][kSsimTile];  // row-filtered moments of one channel
  const ViewDesc vd = views[blockIdx.z];
  const int W = int(vd.width), H = int(vd.height);
  const int x0 = int(blockIdx.x) * kSsimTile, y0 = int(blockIdx.y) * kSsimTile;
  if (x0 >= W || y0 >= H) return;  // the whole block: grid sized for the largest view
  float wt[kWin];
  window_weights<kWin>(wt);
  load_halo<kR, 3>(image + vd.pixel_base * 3, W, H, x0, y0, sx);
  load_halo<kR, 3>(target + vd.pixel_base * 3, W, H, x0, y0, sy);
  __syncthreads();
  const int ox = int(threadIdx.x) % kSsimTile, oy = int(threadIdx.x) / kSsimTile;
  const int x = x0 + ox, y = y0 + oy;
  const bool pixel = x < W && y < H;
  const bool centre = x >= kR && y >= kR && x + kR < W && y + kR < H;
  float* out = coef + (vd.pixel_base + size_t(y) * W + x) * 9;
  for (int c = 0; c < 3; ++c) {
    for (int k = threadIdx.x; k < kHS * kSsimTile; k += blockDim.x) {
      const int hy = k / kSsimTile, hx = k % kSsimTile;
      float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int d = 0; d < kWin; ++d) {
        const float xv = sx[c][hy][hx + d], yv = sy[c][hy][hx + d], w = wt[d];
        m[0] += w * xv; m[1] += w * yv; m[2] += w * xv * xv; m[3] += w * yv * yv; m[4] += w * xv * yv;
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) rows[q][hy][hx] = m[q];
    }
    __syncthreads();
    float mx = 0.0f, my = 0.0f, sxx = 0.0f, syy = 0.0f, sxy = 0.0f;
#pragma unroll
    for (int d = 0; d < kWin; ++d) {
      const float w = wt[d];
      mx += w * rows[0][oy + d][ox]; my += w * rows[1][oy + d][ox];
      sxx += w * rows[2][oy + d][ox]; syy += w * rows[3][oy + d][ox]; sxy += w * rows[4][oy + d][ox];
    }
    float ca = 0.0f, cb = 0.0f, cd = 0.0f;
    if (centre) {
      const float c1 = 0.0001f, c2 = 0.0009f;
      float a1 = 2.0f * mx * my + c1;
      float a2 = 2.0f * (sxy - mx * my) + c2;
      float b1 = mx * mx + my * my + c1;
      float b2 = (sxx - mx * mx) + (syy - my * my) + c2;
      float ssim = a1 * a2 / (b1 * b2);
      ca = ssim * (2.0f * my / a1 - 2.0f * my / a2 - 2.0f * mx / b1 + 2.0f * mx / b2);
      cb = -ssim / b2;
      cd = 2.0f * ssim / a2;
    }
    if (pixel) {
      out[c * 3] = ca;
      out[c * 3 + 1] = cb;
      out[c * 3 + 2] = cd;
    }
    __syncthreads();  // rows is refilled for the next channel
  }
}

// LOSS_GRAD (gpu.js:383-409) without the SSIM term: (1 - lambda) sign(x - y) /
// (3P), one thread per pixel.
__global__ __launch_bounds__(256) void loss_l1_kernel(const ViewDesc* __restrict__ views,
                                                      const float* __restrict__ image,
                                                      const float* __restrict__ target, float lambda,
                                                      float* __restrict__ grad) {
  const ViewDesc vd = views[blockIdx.y];
  const uint32_t P = vd.width * vd.height;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const float n = float(P * 3u);
  const size_t pix = vd.pixel_base + p;
  for (int c = 0; c < 3; ++c)
    grad[pix * 3 + c] = (1.0f - lambda) * sign_f(image[pix * 3 + c] - target[pix * 3 + c]) / n;
}

// LOSS_GRAD (gpu.js:383-409): dLoss/dpixel of (1 - lambda) L1 + lambda (1 - SSIM)
// = (1 - lambda) sign(x - y) / (3P) - lambda / windows * sum over the windows
// holding the pixel of w (a + 2 b x + d y).  The window weights are symmetric,
// so the sum over the windows holding a pixel is the same separable filter over
// the coefficient maps (zero where no window is centred).
template <int kWin>
__global__ __launch_bounds__(256) void loss_grad_kernel(const ViewDesc* __restrict__ views,
                                                        const float* __restrict__ image,
                                                        const float* __restrict__ target,
                                                        const float* __restrict__ coef, float lambda,
                                                        float* __restrict__ grad) {
  constexpr int kR = kWin / 2, kHS = kSsimTile + 2 * kR;
  __shared__ float sc[9][kHS][kHS];
  __shared__ float rows[9][kHS][kSsimTile];
  const ViewDesc vd = views[blockIdx.z];
  const int W = int(vd.width), H = int(vd.height);
  const int x0 = int(blockIdx.x) * kSsimTile, y0 = int(blockIdx.y) * kSsimTile;
  if (x0 >= W || y0 >= H) return;  // the whole block
  const float windows = float(max(0, W - 2 * kR) * max(0, H - 2 * kR) * 3);
  float wt[kWin];
  window_weights<kWin>(wt);
  if (windows > 0.0f) {
    load_halo<kR, 9>(coef + vd.pixel_base * 9, W, H, x0, y0, sc);
    __syncthreads();
    for (int k = threadIdx.x; k < kHS * kSsimTile; k += blockDim.x) {
      const int hy = k / kSsimTile, hx = k % kSsimTile;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        float acc = 0.0f;
#pragma unroll
        for (int d = 0; d < kWin; ++d) acc += wt[d] * sc[q][hy][hx + d];
        rows[q][hy][hx] = acc;
      }
    }
    __syncthreads();
  }
  const int ox = int(threadIdx.x) % kSsimTile, oy = int(threadIdx.x) / kSsimTile;
  const int x = x0 + ox, y = y0 + oy;
  if (x >= W || y >= H) return;
  const float n = float(uint32_t(W * H) * 3u);
  const size_t pix = vd.pixel_base + size_t(y) * W + x;
  for (int c = 0; c < 3; ++c) {
    const float xv = image[pix * 3 + c], yv = target[pix * 3 + c];
    float g = (1.0f - lambda) * sign_f(xv - yv) / n;
    if (windows > 0.0f) {
      float sa = 0.0f, sb = 0.0f, sd = 0.0f;
#pragma unroll
      for (int d = 0; d < kWin; ++d) {
        const float w = wt[d];
        sa += w * rows[c * 3][oy + d][ox];
        sb += w * rows[c * 3 + 1][oy + d][ox];
        sd += w * rows[c * 3 + 2][oy + d][ox];
      }
      const float dS = (sa + 2.0f * xv * sb + yv * sd) / windows;
      g -= lambda * dS;
    }
    grad[pix * 3 + c] = g;
  }
}

// Copies src to dst (or zeroes dst when src is null), `words` 32-bit words.  The
// words [head, head + 4 * n16) move as 16-byte vectors (dst and src share their
// 16-byte phase there); the words outside that body move one by one.
__global__ __launch_bounds__(256) void copy_words_kernel(uint32_t* __restrict__ dst,
                                                         const uint32_t* __restrict__ src, size_t words,
                                                         size_t head, size_t n16) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  uint4* d4 = reinterpret_cast<uint4*>(dst + head);
  if (src) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src + head);
    for (size_t k = i; k < n16; k += stride) d4[k] = s4[k];
  } else {
    for (size_t k = i; k < n16; k += stride) d4[k] = make_uint4(0u, 0u, 0u, 0u);
  }
  const size_t body_end = head + 4 * n16;
  for (size_t k = i; k < head + (words - body_end); k += stride) {
    const size_t w = k < head ? k : body_end + (k - head);
    dst[w] = src ? src[w] : 0u;
  }
}

__global__ __launch_bounds__(256) void fill_work_kernel(const ViewDesc* __restrict__ views, uint32_t num_views,
                                                        uint32_t total_tiles, TileWork* __restrict__ work) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_tiles) return;
  uint32_t v = 0;
  while (v + 1 < num_views && views[v + 1].tile_base <= g) ++v;
  const ViewDesc vd = views[v];
  const uint32_t t = g - vd.tile_base;
  work[g] = make_work(v, t % vd.tiles_x, t / vd.tiles_x, 0, 3);
}

// One thread per tile; the items of split tiles (two per tile: bands 0-1 and
// 2-3) are appended with one atomic per wave.  Each item walks the tile's whole
// list for its bands, so the entry staging is repeated per item: two halves
// shorten the launch's tail nearly as well as four bands at half that cost
// (measured on uniform and on clustered scenes).
constexpr uint32_t kSplitItems = 2;

__global__ __launch_bounds__(256) void split_list_kernel(SplitArgs a) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  bool split = false;
  if (g < a.total_tiles) {
    uint32_t cost = 0;
    if (a.walk) {
      for (int p = 0; p < a.num_passes; ++p) cost += a.walk[size_t(p) * a.total_tiles + g];
    } else {
      for (int p = 0; p < a.num_passes; ++p) cost += a.ranges[p][size_t(g) * 2 + 1] - a.ranges[p][size_t(g) * 2];
    }
    split = cost > a.split_cost;
    a.split[g] = split ? 1 : 0;
  }
  const uint64_t m = __ballot(split);
  if (m == 0) return;
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(a.heavy_count, kSplitItems * uint32_t(__popcll(m)));
  base = __shfl(base, 0);
  if (!split) return;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint32_t at = base + kSplitItems * uint32_t(__popcll(m & below));
  const TileWork w = a.tiles[g];
  const uint32_t v = work_view(w), tx = w.xy & 0xFFFFu, ty = w.xy >> 16;
  constexpr uint32_t kBandsPerItem = uint32_t(kPix) / kSplitItems;
  for (uint32_t b = 0; b < kSplitItems; ++b)
    a.heavy[at + b] = make_work(v, tx, ty, b * kBandsPerItem, b * kBandsPerItem + kBandsPerItem - 1);
}

// Live lists.  A tile is live when one of the passes gives it entries and it is
// neither done nor split (its half-tile items are listed by split_list_kernel).
// Later depth slices find most tiles saturated (two thirds on the 32-view bench
// scene) and some tiles have no entries at all; without the list, each such
// item still costs a wave the queue pop and the dependent loads of its work
// item, view, flags and range before it returns.
__device__ inline bool live_tile(const LiveArgs& a, uint32_t g) {
  if (g >= a.total_tiles) return false;
  if ((a.done && a.done[g]) || (a.split && a.split[g])) return false;
  uint32_t any = 0;
  for (int p = 0; p < a.num_passes; ++p) any |= a.ranges[p][size_t(g) * 2 + 1] ^ a.ranges[p][size_t(g) * 2];
  return any != 0;
}

// Live tiles per block of 256.
__global__ __launch_bounds__(256) void live_count_kernel(LiveArgs a) {
  __shared__ uint32_t wave_live[4];
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  const uint32_t c = uint32_t(__popcll(__ballot(live_tile(a, g))));
  if (__lane_id() == 0) wave_live[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) a.counts[blockIdx.x] = wave_live[0] + wave_live[1] + wave_live[2] + wave_live[3];
}

// Each block writes its live tiles, in order, after those of the blocks before
// it; the last block writes the total.
__global__ __launch_bounds__(256) void live_scatter_kernel(LiveArgs a) {
  __shared__ uint32_t sums[4], wave_live[4];
  const uint32_t wave = threadIdx.x / 64, lane = __lane_id();
  uint32_t before = 0;
  for (uint32_t k = threadIdx.x; k < blockIdx.x; k += 256) before += a.counts[k];
  for (int off = 32; off > 0; off >>= 1) before += __shfl_xor(before, off);
  if (lane == 0) sums[wave] = before;
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  const bool live = live_tile(a, g);
  const uint64_t m = __ballot(live);
  if (lane == 0) wave_live[wave] = uint32_t(__popcll(m));
  __syncthreads();
  const uint32_t base = sums[0] + sums[1] + sums[2] + sums[3];
  uint32_t at = base;
  for (uint32_t w = 0; w < wave; ++w) at += wave_live[w];
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (live) a.out[at + uint32_t(__popcll(m & below))] = a.tiles[g];
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *a.out_count = base + wave_live[0] + wave_live[1] + wave_live[2] + wave_live[3];
}

// After the first pass (which blends every tile) done is assigned; after later
// passes it only gains tiles (tiles a pass skipped keep their flags).
__global__ __launch_bounds__(256) void done_update_kernel(uint32_t total_tiles,
                                                          const uint32_t* __restrict__ band_open,
                                                          uint8_t* __restrict__ done, int assign) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_tiles) return;
  const bool closed = band_open[g] == 0;
  if (assign) done[g] = closed ? 1 : 0;
  else if (closed) done[g] = 1;
}

}  // namespace

void launch_fill_work(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t total_tiles,
                      TileWork* work) {
  if (total_tiles == 0) return;
  hipLaunchKernelGGL(fill_work_kernel, dim3((total_tiles + 255) / 256), dim3(256), 0, s, views, num_views,
                     total_tiles, work);
}

void launch_split_list(hipStream_t s, const SplitArgs& a) {
  if (a.total_tiles == 0) return;
  hipLaunchKernelGGL(split_list_kernel, dim3((a.total_tiles + 255) / 256), dim3(256), 0, s, a);
}

void launch_live_list(hipStream_t s, const LiveArgs& a) {
  if (a.total_tiles == 0) return;
  const dim3 grid((a.total_tiles + 255) / 256);
  hipLaunchKernelGGL(live_count_kernel, grid, dim3(256), 0, s, a);
  hipLaunchKernelGGL(live_scatter_kernel, grid, dim3(256), 0, s, a);
}

void launch_done_update(hipStream_t s, uint32_t total_tiles, const uint8_t* band_open, uint8_t* done, int assign) {
  if (total_tiles == 0) return;
  hipLaunchKernelGGL(done_update_kernel, dim3((total_tiles + 255) / 256), dim3(256), 0, s, total_tiles,
                     reinterpret_cast<const uint32_t*>(band_open), done, assign);
}

template <bool kLast, bool kNoCull>
void launch_forward_variant(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid) {
  if (persistent)
    hipLaunchKernelGGL((forward_persistent_kernel<kLast, kNoCull>), dim3(grid), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((forward_kernel<kLast, kNoCull>), dim3(a.num_work), dim3(64), 0, s, a);
}

void launch_forward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid) {
  if (a.num_work == 0) return;
  const bool track = a.state || a.last;
  if (a.no_cull) {
    if (track) launch_forward_variant<true, true>(s, a, persistent, grid);
    else launch_forward_variant<false, true>(s, a, persistent, grid);
  } else {
    if (track) launch_forward_variant<true, false>(s, a, persistent, grid);
    else launch_forward_variant<false, false>(s, a, persistent, grid);
  }
}

template <bool kNoCull, bool kMoments>
void launch_backward_variant(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid) {
  if (persistent)
    hipLaunchKernelGGL((backward_persistent_kernel<kNoCull, kMoments>), dim3(grid), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((backward_kernel<kNoCull, kMoments>), dim3(a.num_work), dim3(64), 0, s, a);
}

void launch_backward(hipStream_t s, const RasterArgs& a, int persistent, uint32_t grid) {
  if (a.num_work == 0) return;
  if (a.moments) {
    if (a.no_cull) launch_backward_variant<true, true>(s, a, persistent, grid);
    else launch_backward_variant<false, true>(s, a, persistent, grid);
  } else {
    if (a.no_cull) launch_backward_variant<true, false>(s, a, persistent, grid);
    else launch_backward_variant<false, false>(s, a, persistent, grid);
  }
}

void raster_occupancy(int* forward_per_cu, int* backward_per_cu) {
  int f = 0, b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&f, forward_persistent_kernel<false, false>, 64, 0) != hipSuccess)
    f = 8;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, backward_persistent_kernel<false, true>, 64, 0) != hipSuccess) b = 8;
  *forward_per_cu = f;
  *backward_per_cu = b;
}

void launch_loss_grad(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                      const float* image, const float* target, float lambda, float* coef, float* grad,
                      int window) {
  uint32_t maxp = 0;
  for (uint32_t v = 0; v < num_views; ++v) {
    uint32_t p = host_views[v].width * host_views[v].height;
    maxp = p > maxp ? p : maxp;
  }
  if (maxp == 0) return;
  if (!(lambda > 0.0f)) {
    hipLaunchKernelGGL(loss_l1_kernel, dim3((maxp + 255) / 256, num_views), dim3(256), 0, s, views, image, target,
                       lambda, grad);
    return;
  }
  uint32_t maxw = 0, maxh = 0;
  for (uint32_t v = 0; v < num_views; ++v) {
    maxw = std::max(maxw, host_views[v].width);
    maxh = std::max(maxh, host_views[v].height);
  }
  const dim3 tiles((maxw + kSsimTile - 1) / kSsimTile, (maxh + kSsimTile - 1) / kSsimTile, num_views);
  if (window == 11) {
    hipLaunchKernelGGL(ssim_coef_kernel<11>, tiles, dim3(256), 0, s, views, image, target, coef);
    hipLaunchKernelGGL(loss_grad_kernel<11>, tiles, dim3(256), 0, s, views, image, target, coef, lambda, grad);
  } else {
    hipLaunchKernelGGL(ssim_coef_kernel<7>, tiles, dim3(256), 0, s, views, image, target, coef);
    hipLaunchKernelGGL(loss_grad_kernel<7>, tiles, dim3(256), 0, s, views, image, target, coef, lambda, grad);
  }
}

void launch_copy(hipStream_t s, void* dst, const void* src, size_t bytes) {
  // dst and src are 4-byte aligned and bytes is a multiple of 4.
  const size_t words = bytes / 4;
  if (words == 0) return;
  const uintptr_t d = reinterpret_cast<uintptr_t>(dst);
  size_t head = ((16 - d % 16) % 16) / 4;
  if (src && (d - reinterpret_cast<uintptr_t>(src)) % 16 != 0) head = words;  // no common 16-byte phase
  if (head > words) head = words;
  const size_t n16 = (words - head) / 4;
  const size_t scalar = words - 4 * n16;
  size_t blocks = ((n16 > scalar ? n16 : scalar) + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(copy_words_kernel, dim3(uint32_t(blocks)), dim3(256), 0, s, static_cast<uint32_t*>(dst),
                     static_cast<const uint32_t*>(src), words, head, n16);
}

void launch_zero(hipStream_t s, void* p, size_t bytes) { launch_copy(s, p, nullptr, bytes); }

}  // namespace gsr
