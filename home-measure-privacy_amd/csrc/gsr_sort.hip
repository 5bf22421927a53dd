// Tile binning with the ordering contract of tileLists (trainer/gpu.js:450-481):
// per tile, splats in (depth, Gaussian index) order, tiles in row-major order.
//
// The order is a stable LSD radix sort of the 64-bit keys (tile << 32 | depth
// bits) over the entries, split where it pays on this chip:
//   * the 32 depth bits are sorted once per Gaussian (N keys, stable, so equal
//     depths keep index order), because every entry of a Gaussian shares its
//     depth: four 8-bit passes, or (the batched step) three, the last on 9 bits
//     of the key's top 16 less the view's least (see launch_depth_sort);
//   * the tile digit is one counting pass over the entries: blocks walk a chunk of
//     the depth-sorted Gaussians in order and keep one counter per tile in LDS,
//     so each entry is written exactly once (no multi-pass radix over E).
// Both sorts are segmented by view (one segment per camera of a batch).
#include "gsr_internal.h"

namespace gsr {

namespace {

constexpr int kRadixBits = 8;
constexpr int kBins = 1 << kRadixBits;
constexpr int kWideBits = 9;  // the offset pass of the three-pass sort
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;                            // per thread
constexpr int kSortChunk = kSortThreads * kSortItems;     // 4096 keys per block
constexpr int kWaveItems = kSortChunk / (kSortThreads / 64);  // 1024 per wave

// The empty tile rectangle in both formats (x0 = 1 > x1 = 0).
__device__ inline uint2 empty_rect() { return make_uint2(1u, 0u); }

// Workgroup L runs on die L mod 8 (eight dies, each with its own L2).  Groups of
// 8 kDieRun workgroups give each die kDieRun consecutive logical blocks, so the
// partial-line stores of neighbouring blocks meet in one L2 and leave it as whole
// lines; the groups interleave the dies, which keeps their loads even, and the
// tail maps to itself.
constexpr uint32_t kDieRun = 8;
__device__ inline uint32_t die_run_block(uint32_t L, uint32_t total) {
  constexpr uint32_t group = 8 * kDieRun;
  if (L >= total / group * group) return L;
  return L / group * group + (L % 8) * kDieRun + (L / 8) % kDieRun;
}

__device__ inline uint64_t lanemask_lt() {
  uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of the wave whose kBits-bit digit equals this lane's (active lanes only).
template <int kBits = kRadixBits>
__device__ inline uint64_t match_digit(uint32_t d, bool active) {
  uint64_t m = __ballot(active);
#pragma unroll
  for (int b = 0; b < kBits; ++b) {
    bool bit = (d >> b) & 1u;
    uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// Digit of a key in one pass: bits [shift, shift + kBits) of it, or (kOffset, the
// last pass of the three-pass sort) its top 16 bits less lo, the view's least
// over its valid keys, capped at the last digit.  Valid keys are the bits of a
// depth >= 0.2 (their top 16 bits at most 0x7F80, +inf), so with the view's
// range within 511 values the offset digit orders them as the top 16 bits do,
// and invalid keys (0xFFFFFFFF) all take the last digit, after them.
template <int kBits, bool kOffset>
__device__ inline uint32_t key_digit(uint32_t key, int shift, uint32_t lo) {
  constexpr uint32_t kMask = (1u << kBits) - 1u;
  if (kOffset) return min((key >> 16) - lo, kMask);
  return (key >> shift) & kMask;
}

// range[2 v] = 0xFFFF - the least top 16 bits of view v's valid keys, range[2 v +
// 1] = the greatest (range_reduce_kernel; 0 and 0 for a view without valid keys).
// From range + 2 V on: the same per (view, chunk), written by the first pass.
__device__ inline uint32_t range_lo(const uint32_t* range, uint32_t v) { return 0xFFFFu - range[2 * v]; }

// Digit histogram of one 4096-key chunk of one view: hist[v][digit][chunk].
// Mantissa digits of the depth keys are nearly uniform, so one LDS atomic per
// key rarely collides; the top digit (sign and high exponent bits) takes a
// handful of values, so there lanes with equal digits are matched by ballots
// first and add their count once (kMatch).
// kRange (the first pass of the three-pass sort) also stores the chunk's range
// of its valid keys' top 16 bits (range_reduce_kernel then folds the chunks of
// each view); kOffset (its last pass) digits by the view's range.
template <int kBits, bool kMatch, bool kOffset, bool kRange>
__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* __restrict__ keys, uint32_t n,
                                                                   uint32_t chunks, int shift,
                                                                   uint32_t* __restrict__ hist,
                                                                   uint32_t* __restrict__ range,
                                                                   uint32_t* __restrict__ range_flag) {
  constexpr uint32_t kB = 1u << kBits;
  __shared__ uint32_t h[kB];
  __shared__ uint32_t rmin, rmax;
  // A chunk's counts sit next to its neighbouring chunks' (hist[v][digit][chunk]):
  // consecutive chunks on one die, so they leave its L2 as whole lines (the
  // counts were written ten times over before).
  const uint32_t L = die_run_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const uint32_t v = L / gridDim.x, c = L % gridDim.x;
  for (uint32_t b = threadIdx.x; b < kB; b += kSortThreads) h[b] = 0;
  if (kRange && threadIdx.x == 0) {
    rmin = 0xFFFFu;
    rmax = 0u;
  }
  __syncthreads();
  const uint32_t lo = kOffset ? range_lo(range, v) : 0u;
  const size_t base = size_t(v) * n + size_t(c) * kSortChunk;
  const uint32_t len = min(uint32_t(kSortChunk), n - c * kSortChunk);
  // All of the thread's loads are issued before the first is used.
  uint32_t kr[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint32_t j = uint32_t(r) * kSortThreads + threadIdx.x;
    kr[r] = j < len ? keys[base + j] : 0xFFFFFFFFu;
  }
  uint32_t mn = 0xFFFFu, mx = 0u;
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool active = uint32_t(r) * kSortThreads + threadIdx.x < len;
    const uint32_t d = key_digit<kBits, kOffset>(kr[r], shift, lo);
    if (kRange && kr[r] != kInvalidKey) {
      mn = min(mn, kr[r] >> 16);
      mx = max(mx, kr[r] >> 16);
    }
    if constexpr (kMatch) {
      const uint64_t peers = match_digit<kBits>(d, active);
      if (active && (peers & lanemask_lt()) == 0) atomicAdd(&h[d], uint32_t(__popcll(peers)));
    } else if constexpr (kOffset) {
      // Invalid keys (and none other but the view's deepest) take the last
      // digit: one atomic per wave for them, as same-address LDS atomics
      // serialise (a sixth of the keys on the 3M-Gaussian training scene).
      const bool top = active && d == kB - 1u;
      const uint64_t tops = __ballot(top);
      if (top && (tops & lanemask_lt()) == 0) atomicAdd(&h[kB - 1u], uint32_t(__popcll(tops)));
      else if (active && !top) atomicAdd(&h[d], 1u);
    } else if (active) {
      atomicAdd(&h[d], 1u);
    }
  }
  if (kRange) {
    // Over the wave first: every thread's LDS atomic on two words serialises.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn = min(mn, uint32_t(__shfl_xor(int(mn), off)));
      mx = max(mx, uint32_t(__shfl_xor(int(mx), off)));
    }
    if (__lane_id() == 0 && mn <= mx) {
      atomicMin(&rmin, mn);
      atomicMax(&rmax, mx);
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kB; b += kSortThreads) hist[(size_t(v) * kB + b) * chunks + c] = h[b];
  if (kRange && threadIdx.x == 0) {
    // Plain stores (no atomics: 245 blocks of a 1M-Gaussian view queued on one
    // word took the pass from 53 to 122 us).
    uint32_t* part = range + 2 * (size_t(gridDim.y) + size_t(v) * chunks + c);
    part[0] = 0xFFFFu - rmin;
    part[1] = rmax;
  }
}

// One block per view folds the chunks' ranges into the view's, and flags a view
// whose valid keys' top 16 bits span more than 511 values: range_flag |= 2, its
// three-pass order is not valid and the step is run again with four passes.
__global__ __launch_bounds__(kSortThreads) void range_reduce_kernel(uint32_t* __restrict__ range, uint32_t chunks,
                                                                   uint32_t* __restrict__ range_flag) {
  __shared__ uint32_t wlo[kSortThreads / 64], whi[kSortThreads / 64];
  const uint32_t v = blockIdx.x;
  const uint32_t* part = range + 2 * (size_t(gridDim.x) + size_t(v) * chunks);
  uint32_t ilo = 0u, hi = 0u;  // ilo: 0xFFFF - the least (the larger, the less)
  for (uint32_t c = threadIdx.x; c < chunks; c += kSortThreads) {
    ilo = max(ilo, part[2 * c]);
    hi = max(hi, part[2 * c + 1]);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ilo = max(ilo, uint32_t(__shfl_xor(int(ilo), off)));
    hi = max(hi, uint32_t(__shfl_xor(int(hi), off)));
  }
  if (__lane_id() == 0) {
    wlo[threadIdx.x / 64] = ilo;
    whi[threadIdx.x / 64] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kSortThreads / 64; ++w) {
      ilo = max(ilo, wlo[w]);
      hi = max(hi, whi[w]);
    }
    range[2 * v] = ilo;
    range[2 * v + 1] = hi;
    const uint32_t lo = 0xFFFFu - ilo;
    if (hi >= lo && hi - lo > 510u) atomicOr(range_flag, 2u);
  }
}

// Stable scatter of one chunk.  Wave w owns keys [w*1024, (w+1)*1024) of the
// chunk and ranks them in 16 rounds of 64 with ballot peer masks; the chunk is
// then reordered by digit in LDS (stable: wave, round, lane order is index
// order) and written out digit run by digit run, so consecutive threads store
// to consecutive addresses.  In the last pass (rect_sorted set) the sorted keys
// are not needed: each slot gets its Gaussian's tile rectangle instead,
// rect_sorted[v][p] = rects[v][vals_out[v][p]] (empty rectangles mark rejected
// Gaussians), so the binning blocks read rectangles coalesced in depth order.
__device__ inline uint32_t block_exclusive_scan256(uint32_t x, uint32_t* tmp);

template <int kBits, bool kOffset>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, uint32_t n, uint32_t chunks,
    int shift, const uint32_t* __restrict__ hist, uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    const uint2* __restrict__ rects, uint2* __restrict__ rect_sorted, bool packed, const uint32_t* __restrict__ range) {
  constexpr uint32_t kB = 1u << kBits;
  constexpr uint32_t kPer = kB / kSortThreads;  // digits per thread in the digit scan
  __shared__ uint32_t wcnt[4][kB];  // per-wave digit counts, then the waves' local digit bases
  __shared__ uint32_t lstart[kB];   // first local slot of each digit
  __shared__ uint32_t gpos[kB];     // first output slot of each digit for this chunk
  __shared__ uint32_t tmp[kSortThreads];
  __shared__ uint32_t lkeys[kSortChunk];
  __shared__ uint32_t lvals[kSortChunk];
  // The gathering pass reads rects[v] at random: when the views split evenly
  // over the eight dies (workgroup L runs on die L mod 8, each die with its own
  // L2), die d takes views d, d + 8, ... one after another, so its L2 holds one
  // view's rectangles rather than a slice of every view's.
  uint32_t v = blockIdx.y, c = blockIdx.x;
  if (rect_sorted && gridDim.y % 8 == 0) {
    const uint32_t L = blockIdx.x + gridDim.x * blockIdx.y, k = L / 8;
    v = L % 8 + 8 * (k / gridDim.x);
    c = k % gridDim.x;
  } else {
    // Each block writes its chunk as one short run per digit, next to the runs
    // of the neighbouring chunks: consecutive chunks on one die.
    const uint32_t L = die_run_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    v = L / gridDim.x;
    c = L % gridDim.x;
  }
  const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
  const uint32_t lo = kOffset ? range_lo(range, v) : 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) wcnt[k][threadIdx.x * kPer + q] = 0;
  __syncthreads();
  const size_t seg = size_t(v) * n;
  const uint32_t first = c * kSortChunk + w * kWaveItems;
  uint32_t kr[kSortItems], vr[kSortItems], dl[kSortItems];
  const uint64_t lt = lanemask_lt();
  // All loads first, so a wave has its 16 rounds in flight at once.
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint32_t j = first + r * 64 + lane;
    const bool active = j < n;
    kr[r] = active ? keys_in[seg + j] : 0u;
    vr[r] = active ? (vals_in ? vals_in[seg + j] : j) : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint32_t j = first + r * 64 + lane;
    const bool active = j < n;
    const uint32_t key = kr[r];
    uint32_t d = key_digit<kBits, kOffset>(key, shift, lo);
    uint64_t peers = match_digit<kBits>(d, active);
    uint32_t rank = __popcll(peers & lt);
    uint32_t before = active ? wcnt[w][d] : 0u;
    if (active && rank == 0) wcnt[w][d] = before + uint32_t(__popcll(peers));
    dl[r] = active ? (d << 16) | (before + rank) : 0xFFFFFFFFu;
  }
  __syncthreads();
  {
    // Thread t scans digits t kPer .. t kPer + kPer - 1.
    uint32_t tot[kPer], sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint32_t d = threadIdx.x * kPer + q;
      tot[q] = wcnt[0][d] + wcnt[1][d] + wcnt[2][d] + wcnt[3][d];
      sum += tot[q];
    }
    uint32_t ls = block_exclusive_scan256(sum, tmp);
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint32_t d = threadIdx.x * kPer + q;
      const uint32_t c0 = wcnt[0][d], c1 = wcnt[1][d], c2 = wcnt[2][d];
      lstart[d] = ls;
      gpos[d] = hist[(size_t(v) * kB + d) * chunks + c];
      wcnt[0][d] = ls;
      wcnt[1][d] = ls + c0;
      wcnt[2][d] = ls + c0 + c1;
      wcnt[3][d] = ls + c0 + c1 + c2;
      ls += tot[q];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (dl[r] == 0xFFFFFFFFu) continue;
    const uint32_t local = wcnt[w][dl[r] >> 16] + (dl[r] & 0xFFFFu);
    lkeys[local] = kr[r];
    lvals[local] = vr[r];
  }
  __syncthreads();
  const uint32_t len = min(uint32_t(kSortChunk), n - c * kSortChunk);
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const uint32_t li = uint32_t(r) * kSortThreads + threadIdx.x;
    if (li >= len) continue;
    const uint32_t key = lkeys[li];
    const uint32_t d = key_digit<kBits, kOffset>(key, shift, lo);
    const uint32_t dst = gpos[d] + (li - lstart[d]);
    const uint32_t val = lvals[li];
    vals_out[seg + dst] = val;
    // A Gaussian without tiles has the invalid key and the empty rectangle, so
    // its rectangle is written rather than gathered (16% of the gathers on the
    // 3M-Gaussian 64-view training scene).  The gathers are bound by the rate of
    // random accesses, not by their latency: issuing all sixteen before the
    // first store left the pass as it was.
    const bool none = key == kInvalidKey;
    if (rect_sorted && packed)
      reinterpret_cast<uint32_t*>(rect_sorted)[seg + dst] =
          none ? empty_rect().x : reinterpret_cast<const uint32_t*>(rects)[seg + val];
    else if (rect_sorted)
      rect_sorted[seg + dst] = none ? empty_rect() : rects[seg + val];
    else
      keys_out[seg + dst] = key;
  }
}

// Inclusive scans over the wave with row shifts and row broadcasts (DPP):
// rows of 16 first, then row 0's total into row 1 (and row 2's into row 3),
// then lane 31's running total into rows 2 and 3.
template <bool kMax>
__device__ inline uint32_t dpp_step(uint32_t x, uint32_t y) {
  return kMax ? (x > y ? x : y) : x + y;
}
template <bool kMax>
__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xF, 0xF, true)));  // row_shr:1
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xF, 0xF, true)));  // row_shr:2
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x114, 0xF, 0xF, true)));  // row_shr:4
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x118, 0xF, 0xF, true)));  // row_shr:8
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x142, 0xA, 0xF, false)));  // row_bcast:15
  x = dpp_step<kMax>(x, uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x143, 0xC, 0xF, false)));  // row_bcast:31
  return x;
}

// ---- segmented exclusive scan over equal-length segments (u32) ------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanPiece = kScanThreads * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t x, uint32_t* tmp, uint32_t& total) {
  // Inclusive scan of x over the block through LDS (Hillis-Steele), returns exclusive.
  const int t = threadIdx.x;
  tmp[t] = x;
  __syncthreads();
  for (int off = 1; off < int(blockDim.x); off <<= 1) {
    uint32_t y = t >= off ? tmp[t - off] : 0u;
    __syncthreads();
    tmp[t] += y;
    __syncthreads();
  }
  uint32_t incl = tmp[t];
  total = tmp[blockDim.x - 1];
  __syncthreads();
  return incl - x;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ inline uint32_t block_exclusive_scan256(uint32_t x, uint32_t* tmp) {
  uint32_t total;
  return block_exclusive_scan(x, tmp, total);
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ data, size_t seg_len,
                                                                   uint32_t pieces, uint32_t* __restrict__ piece) {
  __shared__ uint32_t tmp[kScanThreads];
  const uint32_t s = blockIdx.y, p = blockIdx.x;
  const size_t base = size_t(s) * seg_len + size_t(p) * kScanPiece;
  const size_t lim = seg_len - size_t(p) * kScanPiece;
  uint32_t sum = 0;
  for (int k = 0; k < kScanItems; ++k) {
    size_t j = size_t(k) * kScanThreads + threadIdx.x;
    if (j < lim) sum += data[base + j];
  }
  uint32_t total;
  block_exclusive_scan(sum, tmp, total);
  if (threadIdx.x == 0) piece[size_t(s) * pieces + p] = total;
}

__global__ __launch_bounds__(1024) void scan_pieces_kernel(uint32_t* __restrict__ piece, uint32_t pieces) {
  __shared__ uint32_t tmp[1024];
  const uint32_t s = blockIdx.x;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < pieces; b += 1024) {
    uint32_t j = b + threadIdx.x;
    uint32_t x = j < pieces ? piece[size_t(s) * pieces + j] : 0u;
    uint32_t total;
    uint32_t ex = block_exclusive_scan(x, tmp, total);
    if (j < pieces) piece[size_t(s) * pieces + j] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(uint32_t* __restrict__ data, size_t seg_len,
                                                                  uint32_t pieces,
                                                                  const uint32_t* __restrict__ piece) {
  // The piece passes through LDS (one pad word per 32 against bank conflicts),
  // so the global loads and stores are coalesced while each thread scans
  // kScanItems consecutive elements: consecutive words per thread straight
  // from memory left each wave load spread over 64 lines (the pass ran at
  // under 1 TB/s on the 3M-Gaussian training scene's histograms).
  __shared__ uint32_t tile[kScanPiece + kScanPiece / 32];
  __shared__ uint32_t wave_total[kScanThreads / 64];
  const uint32_t s = blockIdx.y, p = blockIdx.x, t = threadIdx.x, w = t >> 6;
  const size_t base = size_t(s) * seg_len + size_t(p) * kScanPiece;
  const size_t lim = seg_len - size_t(p) * kScanPiece;
  auto pad = [](uint32_t i) { return i + (i >> 5); };
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t j = uint32_t(k) * kScanThreads + t;
    tile[pad(j)] = j < lim ? data[base + j] : 0u;
  }
  __syncthreads();
  uint32_t x[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    x[k] = tile[pad(t * kScanItems + k)];
    sum += x[k];
  }
  const uint32_t incl = wave_inclusive_scan<false>(sum);
  if (__lane_id() == 63) wave_total[w] = incl;
  __syncthreads();
  uint32_t run = incl - sum + piece[size_t(s) * pieces + p];
  for (uint32_t k = 0; k < w; ++k) run += wave_total[k];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    tile[pad(t * kScanItems + k)] = run;
    run += x[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint32_t j = uint32_t(k) * kScanThreads + t;
    if (j < lim) data[base + j] = tile[pad(j)];
  }
}

void scan_segments(hipStream_t st, uint32_t* data, size_t seg_len, uint32_t segs, uint32_t* piece) {
  uint32_t pieces = uint32_t((seg_len + kScanPiece - 1) / kScanPiece);
  dim3 grid(pieces, segs);
  hipLaunchKernelGGL(scan_reduce_kernel, grid, dim3(kScanThreads), 0, st, data, seg_len, pieces, piece);
  hipLaunchKernelGGL(scan_pieces_kernel, dim3(segs), dim3(1024), 0, st, piece, pieces);
  hipLaunchKernelGGL(scan_apply_kernel, grid, dim3(kScanThreads), 0, st, data, seg_len, pieces, piece);
}

// ---- binning -------------------------------------------------------------------
// Blocks work on (chunk of depth-sorted Gaussians, band of tile rows).  A band
// holds about kBandTiles tiles, so its counters fit a few KB of LDS and many
// blocks stay resident per CU.  A pass bins one depth slice [pbeg, pend) of every
// view's sorted Gaussians; tiles flagged done (already saturated by the slices
// before) receive no entries.  The count matrix of a pass is laid out
// [view][chunk][tile] at offset chunks * tile_base of the view.
constexpr uint32_t kBandTiles = 2048;      // count kernel: 4 B of LDS per tile
// The emit kernel's band (BinScratch::emit_band: 256 tiles for the single-view
// path, the batch's knob otherwise) costs
// 12 B of LDS per tile: smaller bands keep more waves resident, larger ones
// rescan each chunk's rectangles fewer times.

__host__ __device__ inline uint32_t band_rows(uint32_t tiles_x, uint32_t tiles_y, uint32_t budget) {
  uint32_t r = budget / (tiles_x ? tiles_x : 1);
  r = r < 1 ? 1 : r;
  return r < tiles_y ? r : (tiles_y ? tiles_y : 1);
}

struct Rect {
  uint32_t x0, x1, y0, y1;  // inclusive; empty when x0 > x1 or y0 > y1
};

// A rectangle as loaded (raw.x holds the packed word when kPacked).
template <bool kPacked>
__device__ inline Rect unpack_rect(uint2 raw) {
  if (kPacked) return Rect{raw.x & 0xFFu, (raw.x >> 8) & 0xFFu, (raw.x >> 16) & 0xFFu, raw.x >> 24};
  return Rect{raw.x & 0xFFFFu, raw.x >> 16, raw.y & 0xFFFFu, raw.y >> 16};
}
template <bool kPacked>
__device__ inline uint2 load_rect(const uint2* __restrict__ rects, size_t i) {
  if (kPacked) return make_uint2(reinterpret_cast<const uint32_t*>(rects)[i], 0u);
  return rects[i];
}

struct Band {
  uint32_t tx, r0, r1;  // tiles per row, band rows [r0, r1)
  bool valid;
};

__device__ inline Band band_of(const ViewDesc& vd, uint32_t band, uint32_t budget) {
  const uint32_t R = band_rows(vd.tiles_x, vd.tiles_y, budget);
  Band b;
  b.tx = vd.tiles_x;
  b.r0 = band * R;
  b.valid = b.r0 < vd.tiles_y;
  b.r1 = min(vd.tiles_y, b.r0 + R);
  return b;
}

// Chunks of view v that hold positions of the pass: all of them, or for a
// filtered pass (kept set) those below the view's kept count.  Counts of the
// chunks past it are neither written nor read.
__device__ inline uint32_t live_chunks(uint32_t v, uint32_t chunks, uint32_t chunk, const uint32_t* kept) {
  return kept ? min(chunks, (kept[v] + chunk - 1) / chunk) : chunks;
}

// Pass 1: per (chunk, tile) entry counts.  Each Gaussian adds its rectangle
// (clipped to the band) to a 2D difference array with four LDS atomics; prefix
// sums along x then y turn it into counts.  Done tiles count zero.
template <bool kPacked>
__global__ __launch_bounds__(64) void bin_count_kernel(const ViewDesc* __restrict__ views, uint32_t n, uint32_t chunk,
                                                       uint32_t pbeg, uint32_t pend, uint32_t chunks,
                                                       const uint2* __restrict__ sorted,
                                                       const uint8_t* __restrict__ done,
                                                       uint32_t* __restrict__ matrix,
                                                       const uint32_t* __restrict__ kept) {
  extern __shared__ int diff[];
  const uint32_t v = blockIdx.z, c = blockIdx.x, lane = threadIdx.x;
  const ViewDesc vd = views[v];
  const Band bd = band_of(vd, blockIdx.y, kBandTiles);
  if (!bd.valid || c >= live_chunks(v, chunks, chunk, kept)) return;
  const uint32_t rows = bd.r1 - bd.r0, tx = bd.tx, W = tx + 1;
  for (uint32_t k = lane; k < (rows + 1) * W; k += 64) diff[k] = 0;
  __syncthreads();
  const size_t seg = size_t(v) * n;
  const uint32_t begin = pbeg + c * chunk, end = min(kept ? kept[v] : pend, begin + chunk);
  // Eight rectangles per lane are loaded before any is used (past the end: the
  // empty rectangle).
  constexpr uint32_t kAhead = 8;
  for (uint32_t p0 = begin; p0 < end; p0 += 64 * kAhead) {
    uint2 rc[kAhead];
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const uint32_t p = p0 + 64 * u + lane;
      rc[u] = p < end ? load_rect<kPacked>(sorted, seg + p) : empty_rect();
    }
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const Rect r = unpack_rect<kPacked>(rc[u]);
      uint32_t y0 = max(r.y0, bd.r0), y1 = min(r.y1, bd.r1 - 1);
      if (r.x0 > r.x1 || y0 > y1) continue;
      y0 -= bd.r0;
      y1 -= bd.r0;
      atomicAdd(&diff[y0 * W + r.x0], 1);
      atomicAdd(&diff[y0 * W + r.x1 + 1], -1);
      atomicAdd(&diff[(y1 + 1) * W + r.x0], -1);
      atomicAdd(&diff[(y1 + 1) * W + r.x1 + 1], 1);
    }
  }
  __syncthreads();
  for (uint32_t y = lane; y < rows; y += 64) {
    int run = 0;
    for (uint32_t x = 0; x < tx; ++x) {
      run += diff[y * W + x];
      diff[y * W + x] = run;
    }
  }
  __syncthreads();
  for (uint32_t x = lane; x < tx; x += 64) {
    int run = 0;
    for (uint32_t y = 0; y < rows; ++y) {
      run += diff[y * W + x];
      diff[y * W + x] = run;
    }
  }
  __syncthreads();
  const uint32_t T = tx * vd.tiles_y;
  uint32_t* out = matrix + size_t(chunks) * vd.tile_base + size_t(c) * T + size_t(bd.r0) * tx;
  const uint8_t* dn = done ? done + vd.tile_base + size_t(bd.r0) * tx : nullptr;
  const uint32_t cells = rows * tx;
  for (uint32_t k0 = 0; k0 < cells; k0 += 64 * kAhead) {
    uint32_t dk[kAhead] = {};  // done flags, loaded before any count is written
    if (dn) {
#pragma unroll
      for (uint32_t u = 0; u < kAhead; ++u) dk[u] = dn[min(k0 + 64 * u + lane, cells - 1)];
    }
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const uint32_t k = k0 + 64 * u + lane;
      if (k < cells) out[k] = dk[u] ? 0u : uint32_t(diff[(k / tx) * W + k % tx]);
    }
  }
}

// tile_tot[tile_base + t] = sum over chunks of the pass's counts.
__global__ __launch_bounds__(256) void bin_colsum_kernel(const ViewDesc* __restrict__ views, uint32_t chunks,
                                                         const uint32_t* __restrict__ matrix,
                                                         uint32_t* __restrict__ tile_tot, uint32_t chunk,
                                                         const uint32_t* __restrict__ kept) {
  const uint32_t v = blockIdx.y;
  const ViewDesc vd = views[v];
  const uint32_t T = vd.tiles_x * vd.tiles_y;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const uint32_t* m = matrix + size_t(chunks) * vd.tile_base + t;
  const uint32_t live = live_chunks(v, chunks, chunk, kept);
  // Eight rows in flight per thread: one at a time leaves the walk latency-bound.
  uint32_t s = 0, c = 0;
  for (; c + 8 <= live; c += 8) {
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = m[size_t(c + k) * T];
#pragma unroll
    for (int k = 0; k < 8; ++k) s += x[k];
  }
  for (; c < live; ++c) s += m[size_t(c) * T];
  tile_tot[vd.tile_base + t] = s;
}

// Exclusive scan over a view's tiles: ranges (gpu.js:468-474) and the view's
// entry count for the pass.
__global__ __launch_bounds__(1024) void bin_tilescan_kernel(const ViewDesc* __restrict__ views, uint32_t chunks,
                                                            const uint32_t* __restrict__ tile_tot,
                                                            uint32_t* __restrict__ ranges,
                                                            uint64_t* __restrict__ entries) {
  // Thread i takes tiles [i K, i K + K) in turn, after one scan of the threads'
  // sums (wave scans, then the 16 waves' totals): one block-wide step instead of
  // a barrier-bound scan per 1024 tiles.
  __shared__ uint32_t wave_total[16];
  const uint32_t v = blockIdx.x;
  const ViewDesc vd = views[v];
  const uint32_t T = vd.tiles_x * vd.tiles_y, K = (T + 1023) / 1024;
  const uint32_t t0 = min(T, threadIdx.x * K), t1 = min(T, t0 + K);
  const uint32_t* tot = tile_tot + vd.tile_base;
  uint32_t sum = 0;
  if (chunks > 0)
    for (uint32_t t = t0; t < t1; ++t) sum += tot[t];
  const uint32_t w = threadIdx.x >> 6;
  const uint32_t incl = wave_inclusive_scan<false>(sum);
  if (__lane_id() == 63) wave_total[w] = incl;
  __syncthreads();
  uint32_t run = incl - sum;
  for (uint32_t k = 0; k < w; ++k) run += wave_total[k];
  if (threadIdx.x == 1023) entries[v] = run + sum;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint32_t x = chunks > 0 ? tot[t] : 0u;
    ranges[size_t(vd.tile_base + t) * 2] = run;
    ranges[size_t(vd.tile_base + t) * 2 + 1] = run + x;
    run += x;
  }
}

// First order slot of each view's entries for the pass: a running total over
// passes and views (*run), reset by the first pass.
__global__ void bin_entry_base_kernel(uint32_t num_views, const uint64_t* __restrict__ entries,
                                      uint64_t* __restrict__ ebase, uint64_t* __restrict__ run, int first) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t r = first ? 0 : *run;
  for (uint32_t v = 0; v < num_views; ++v) {
    ebase[v] = r;
    r += entries[v];
  }
  *run = r;
}

// matrix[v][c][t] <- ranges start of t + sum_{c' < c} matrix[v][c'][t].
__global__ __launch_bounds__(256) void bin_colscan_kernel(const ViewDesc* __restrict__ views, uint32_t chunks,
                                                          const uint32_t* __restrict__ ranges,
                                                          uint32_t* __restrict__ matrix, uint32_t chunk,
                                                          const uint32_t* __restrict__ kept) {
  const uint32_t v = blockIdx.y;
  const ViewDesc vd = views[v];
  const uint32_t T = vd.tiles_x * vd.tiles_y;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  uint32_t* m = matrix + size_t(chunks) * vd.tile_base + t;
  uint32_t run = ranges[size_t(vd.tile_base + t) * 2];
  const uint32_t live = live_chunks(v, chunks, chunk, kept);
  uint32_t c = 0;
  for (; c + 8 <= live; c += 8) {  // eight rows loaded before any is rewritten
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = m[size_t(c + k) * T];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      m[size_t(c + k) * T] = run;
      run += x[k];
    }
  }
  for (; c < live; ++c) {
    const uint32_t x = m[size_t(c) * T];
    m[size_t(c) * T] = run;
    run += x;
  }
}

// Pass 2: the order itself.  The block holds per-tile slot cursors of its band,
// seeded with the scanned matrix, and walks its chunk's Gaussians in depth order,
// 64 at a time.  Gaussians whose rectangle meets the band join an LDS queue, in
// order; every 64 queued Gaussians (and the rest at the end) are flushed.  A
// flush flattens their (Gaussian, tile) pairs Gaussian-major (queue entry j's
// pairs occupy [start_j, start_j + cnt_j)) and hands them out 64 per batch, one
// per lane.  Within a batch the pairs of one tile come from distinct Gaussians
// in queue (= depth) order, so a pair's slot is the tile's cursor plus the
// number of lower lanes holding the same tile: every lane sets its bit in a
// 64-bit LDS mask of its tile, reads the mask back and counts the bits below its
// own; the highest lane of each mask then advances the cursor and clears the
// mask.  Slots of a tile therefore follow (depth, index).  Done tiles carry
// cursor 0xFFFFFFFF and are skipped; bands whose tiles are all done return at
// once.

constexpr uint32_t kEmitQueue = 128;
constexpr uint32_t kOpenWords = 8;  // open-column mask of a band: up to 256 tiles per row
constexpr uint32_t kLaterBandScale = 4;  // emit band of later depth slices, in first-slice bands
constexpr uint32_t kMaxEmitBand = 4096;  // tiles; 12 B of LDS each (gsr_batch_set_emit_band's limit)
constexpr uint32_t kEmitFixedLds = kEmitQueue * 16 + 64 * 4 + kOpenWords * 4;  // queue, owner marks, open columns
// A tile's slots from neighbouring chunks are neighbours in the order, so each
// die takes kDieRun consecutive chunks of a band (die_run_block): their short
// runs then meet in one L2 and leave it as whole lines.

template <bool kPacked>
__global__ __launch_bounds__(64) void bin_emit_kernel(const ViewDesc* __restrict__ views, uint32_t n, uint32_t chunk,
                                                      uint32_t pbeg, uint32_t pend, uint32_t chunks,
                                                      const uint32_t* __restrict__ vals,
                                                      const uint2* __restrict__ sorted,
                                                      const uint8_t* __restrict__ done,
                                                      const uint32_t* __restrict__ matrix,
                                                      const uint64_t* __restrict__ ebase_v,
                                                      uint32_t* __restrict__ order, uint64_t capacity,
                                                      uint32_t* __restrict__ overflow, uint32_t emit_band,
                                                      const uint32_t* __restrict__ kept) {
  extern __shared__ uint4 lds_quads[];
  const uint32_t X = gridDim.x, XY = gridDim.x * gridDim.y;
  const uint32_t logical = die_run_block(blockIdx.x + X * blockIdx.y + XY * blockIdx.z, XY * gridDim.z);
  const uint32_t v = logical / XY, c = logical % X, lane = threadIdx.x;
  const ViewDesc vd = views[v];
  const Band bd = band_of(vd, (logical % XY) / X, emit_band);
  if (!bd.valid || c >= live_chunks(v, chunks, chunk, kept)) return;
  const uint32_t tx = bd.tx, ntile = (bd.r1 - bd.r0) * tx;
  // queue record: band-local tile of the rectangle's corner, width, pair count
  // (replaced by the first pair's position during a flush), Gaussian index
  uint4* queue = lds_quads;                                                  // [kEmitQueue]
  uint32_t* owner_mark = reinterpret_cast<uint32_t*>(lds_quads + kEmitQueue);  // [64]
  uint64_t* mask = reinterpret_cast<uint64_t*>(owner_mark + 64);             // [ntile]
  uint32_t* cursor = reinterpret_cast<uint32_t*>(mask + ntile);              // [ntile]
  // Later depth slices: columns of the band with an open (unsaturated) tile in
  // some row.  Each rectangle is clipped to the open columns before it is
  // queued, so the pairs of saturated tiles (most of a later slice's pairs on
  // the bench scene) are not ranked just to be dropped at the store.
  uint32_t* open_cols = cursor + ntile;  // [kOpenWords]
  const bool clip = done != nullptr && tx <= 32 * kOpenWords;
  if (clip && lane < kOpenWords) open_cols[lane] = 0;
  __syncthreads();
  const uint32_t T = tx * vd.tiles_y;
  const uint32_t* base = matrix + size_t(chunks) * vd.tile_base + size_t(c) * T + size_t(bd.r0) * tx;
  const uint8_t* dn = done ? done + vd.tile_base + size_t(bd.r0) * tx : nullptr;
  bool open = false;
  for (uint32_t k0 = 0; k0 < ntile; k0 += 256) {  // four tiles per lane loaded at once
    uint32_t bk[4], dk[4] = {};
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = min(k0 + 64 * u + lane, ntile - 1);
      bk[u] = base[k];
      if (dn) dk[u] = dn[k];
    }
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) {
      const uint32_t k = k0 + 64 * u + lane;
      if (k >= ntile) continue;
      cursor[k] = dk[u] ? 0xFFFFFFFFu : bk[u];
      mask[k] = 0;
      open |= dk[u] == 0;
      if (clip && dk[u] == 0) {
        const uint32_t col = k % tx;
        atomicOr(&open_cols[col >> 5], 1u << (col & 31u));
      }
    }
  }
  if (__ballot(open) == 0) return;  // every tile of the band is saturated
  __syncthreads();
  const size_t seg = size_t(v) * n;
  const uint32_t begin = pbeg + c * chunk, end = min(kept ? kept[v] : pend, begin + chunk);
  const uint64_t ebase = ebase_v[v];
  const uint64_t below = (uint64_t(1) << lane) - 1;
  bool over = false;
  uint32_t queued = 0;  // wave-uniform

  // Emits the pairs of queue entries [0, count) and drops them from the queue.
  auto flush = [&](uint32_t count) {
    __syncthreads();
    uint4 rec = make_uint4(0u, 1u, 0u, 0u);
    if (lane < count) rec = queue[lane];
    const uint32_t cnt = rec.z;
    const uint32_t incl = wave_inclusive_scan<false>(cnt);
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    const uint32_t start = incl - cnt;
    if (lane < count) queue[lane].z = start;
    for (uint32_t b0 = 0; b0 < total; b0 += 64) {
      // Mark where each entry's pairs begin in this batch (lane + 1), then a
      // running max hands every pair its entry.
      owner_mark[lane] = 0;
      if (cnt > 0 && start < b0 + 64 && incl > b0) owner_mark[(start > b0 ? start : b0) - b0] = lane + 1;
      __syncthreads();
      // Every position of the batch has an owner (position 0 always has its
      // mark), so the lanes past the last pair read valid queue slots, tiles and
      // cursors too, and only the stores are conditional: fewer exec-mask
      // regions (scalar instructions) per batch than branches around each step
      // (emit 1.42 -> 1.33 ms at 256-tile bands on the 32-view bench scene,
      // 3.96 -> 3.62 ms on the 3M-Gaussian training scene).
      // Marked positions belong to consecutive entries, position 0's first: a
      // pair's entry is position 0's plus the marks up to its position (a
      // ballot and a lane bit count instead of a running max over the wave).
      const uint32_t mk = owner_mark[lane];
      const uint64_t marks = __ballot(mk != 0u);
      const uint32_t below_marks =
          __builtin_amdgcn_mbcnt_hi(uint32_t(marks >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(marks), 0u));
      const uint32_t own = __builtin_amdgcn_readfirstlane(mk) - 2u + below_marks + (mk != 0u ? 1u : 0u);
      const uint32_t e = b0 + lane;
      const bool act = e < total;
      const uint4 o = queue[own];
      const uint32_t k = act ? e - o.z : 0u, ow = o.y;
      // k < the rectangle's tiles in the band (< 2^16): an approximate
      // reciprocal, off by at most one, then one exact correction each way.
      uint32_t row = uint32_t(float(k) * __builtin_amdgcn_rcpf(float(ow)));
      row = __umul24(row, ow) > k ? row - 1u : row;
      row = __umul24(row + 1u, ow) <= k ? row + 1u : row;
      const uint32_t t = o.x + __umul24(row, tx) + (k - __umul24(row, ow));
      const uint32_t pos = cursor[t];
      atomicOr(reinterpret_cast<unsigned long long*>(&mask[t]),
               act ? static_cast<unsigned long long>(1) << lane : 0ull);
      __syncthreads();
      const uint64_t m = mask[t];
      const bool put = act && pos != 0xFFFFFFFFu;
      const uint64_t slot = ebase + pos + uint32_t(__popcll(m & below));
      over |= put && slot >= capacity;
      if (put && slot < capacity) order[slot] = o.w;
      if (act && (m >> lane) == 1) {  // highest lane of this tile in the batch
        cursor[t] = pos != 0xFFFFFFFFu ? pos + uint32_t(__popcll(m)) : pos;
        mask[t] = 0;
      }
      __syncthreads();
    }
    const uint32_t rest = queued - count;  // < 64
    uint4 moved = make_uint4(0u, 0u, 0u, 0u);
    if (lane < rest) moved = queue[count + lane];
    __syncthreads();
    if (lane < rest) queue[lane] = moved;
    queued = rest;
    __syncthreads();
  };

  // Rectangles and indices are loaded kAhead groups at a time, one load round
  // ahead: stores count in vmcnt too, so the wait for a round also waits for the
  // flushes' stores, and rounds make that wait rarer.
  constexpr int kAhead = 4;
  uint2 nrc[kAhead];
  uint32_t ngi[kAhead];
  auto load_round = [&](uint32_t q0) {
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const uint32_t p = q0 + 64 * u + lane;
      nrc[u] = empty_rect();
      ngi[u] = 0;
      if (p < end) {
        nrc[u] = load_rect<kPacked>(sorted, seg + p);
        ngi[u] = vals[seg + p];
      }
    }
  };
  load_round(begin);
  for (uint32_t q0 = begin; q0 < end; q0 += 64 * kAhead) {
    uint2 crc[kAhead];
    uint32_t cgi[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      crc[u] = nrc[u];
      cgi[u] = ngi[u];
    }
    if (q0 + 64 * kAhead < end) load_round(q0 + 64 * kAhead);
#pragma unroll 1
    for (int u = 0; u < kAhead; ++u) {
      const Rect r = unpack_rect<kPacked>(crc[0]);
      const uint32_t gi = cgi[0];
#pragma unroll
      for (int k = 0; k + 1 < kAhead; ++k) {
        crc[k] = crc[k + 1];
        cgi[k] = cgi[k + 1];
      }
      const uint32_t y0 = max(r.y0, bd.r0), y1 = min(r.y1, bd.r1 - 1);
      uint32_t x0 = r.x0, x1 = r.x1;
      bool hit = x0 <= x1 && y0 <= y1;
      if (clip && hit && (x1 >> 5) <= (x0 >> 5) + 1) {  // wider rectangles keep their columns
        const uint32_t w0 = x0 >> 5, b0 = w0 << 5;
        uint64_t m = open_cols[w0];
        if ((x1 >> 5) > w0) m |= uint64_t(open_cols[w0 + 1]) << 32;
        m &= (~0ull << (x0 - b0)) & (~0ull >> (63u - (x1 - b0)));
        hit = m != 0;
        if (hit) {
          x0 = b0 + uint32_t(__builtin_ctzll(m));
          x1 = b0 + uint32_t(63 - __builtin_clzll(m));
        }
      }
      const uint64_t hits = __ballot(hit);
      if (hits == 0) continue;
      if (hit) {
        const uint32_t rank =
            __builtin_amdgcn_mbcnt_hi(uint32_t(hits >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(hits), 0u));
        const uint32_t w = x1 - x0 + 1;
        queue[queued + rank] = make_uint4(__umul24(y0 - bd.r0, tx) + x0, w, __umul24(w, y1 - y0 + 1), gi);
      }
      queued += uint32_t(__popcll(hits));
      if (queued >= 64) flush(64);
    }
  }
  if (queued > 0) flush(queued);
  if (over) atomicOr(overflow, 1u);
}

// ---- later depth slices: the Gaussians that reach an open tile -------------------
// A later slice is binned only for the tiles the slices before it left open
// (unsaturated), and most of its Gaussians reach none of them: the count and emit
// kernels would scan every one of them for every band with an open tile.  A
// stable filter first keeps, in depth order, the Gaussians whose rectangle meets
// an open tile.  The lists do not change: a Gaussian whose rectangle meets no
// open tile adds no entry to the pass.
constexpr uint32_t kFilterPiece = 4096;  // sorted positions per wave of the filter
constexpr uint32_t kFilterWaves = 4;     // waves (pieces) per block
constexpr uint32_t kMaxFilterTiles = 32768;  // views filtered: 64 KB of LDS table

// sat[tile_base + y tx + x] = open tiles (done[.] == 0) in columns 0..x of rows
// 0..y of the view, modulo 2^16: a summed-area table, so a rectangle's open tiles
// take four reads.  The four-corner sum is exact modulo 2^16, and a view of at
// most kMaxFilterTiles tiles holds fewer than 2^16 tiles in any rectangle, so
// the sum is zero only when the rectangle holds no open tile.  One block per
// view: rows by wave scans, then columns down.
__global__ __launch_bounds__(1024) void open_sat_kernel(const ViewDesc* __restrict__ views,
                                                        const uint8_t* __restrict__ done, uint16_t* __restrict__ sat) {
  const ViewDesc vd = views[blockIdx.x];
  const uint32_t tx = vd.tiles_x, ty = vd.tiles_y, lane = __lane_id(), w = threadIdx.x >> 6;
  uint16_t* st = sat + vd.tile_base;
  const uint8_t* dn = done + vd.tile_base;
  for (uint32_t y = w; y < ty; y += 16) {  // wave-uniform loops: every lane takes part in the scans
    uint32_t carry = 0;
    for (uint32_t x0 = 0; x0 < tx; x0 += 64) {
      const uint32_t x = x0 + lane;
      const uint32_t open = x < tx && dn[size_t(y) * tx + x] == 0 ? 1u : 0u;
      const uint32_t inc = wave_inclusive_scan<false>(open) + carry;
      if (x < tx) st[size_t(y) * tx + x] = uint16_t(inc);
      carry = __builtin_amdgcn_readlane(inc, 63);
    }
  }
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < tx; x += 1024) {
    uint32_t run = 0, y = 0;
    for (; y + 8 <= ty; y += 8) {  // eight rows loaded before any is rewritten
      uint32_t r[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) r[k] = st[size_t(y + k) * tx + x];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        run += r[k];
        st[size_t(y + k) * tx + x] = uint16_t(run);
      }
    }
    for (; y < ty; ++y) {
      run += st[size_t(y) * tx + x];
      st[size_t(y) * tx + x] = uint16_t(run);
    }
  }
}

// True when the rectangle (inclusive, within the view's tiles, or empty) holds an
// open tile; st is the view's table in LDS.
__device__ inline bool meets_open(const uint16_t* st, uint32_t tx, const Rect& r) {
  if (r.x0 > r.x1 || r.y0 > r.y1) return false;
  const uint32_t a = st[r.y1 * tx + r.x1];
  const uint32_t b = r.x0 ? st[r.y1 * tx + r.x0 - 1] : 0u;
  const uint32_t c = r.y0 ? st[(r.y0 - 1) * tx + r.x1] : 0u;
  const uint32_t d = r.x0 && r.y0 ? st[(r.y0 - 1) * tx + r.x0 - 1] : 0u;
  return ((a - b - c + d) & 0xFFFFu) != 0u;
}

// Wave pc of a view's slice [pbeg, pend) tests positions [pbeg + pc kFilterPiece,
// +kFilterPiece): mask[v][(p - pbeg) / 64] gets the kept bits of each 64 of them,
// piece[v][pc] the wave's kept count.
template <bool kPacked>
__global__ __launch_bounds__(64 * kFilterWaves) void slice_mark_kernel(
    const ViewDesc* __restrict__ views, uint32_t n, uint32_t pbeg, uint32_t pend, uint32_t pieces,
    const uint2* __restrict__ sorted, const uint16_t* __restrict__ sat, uint64_t* __restrict__ mask,
    uint32_t mask_words, uint32_t* __restrict__ piece) {
  extern __shared__ uint16_t lsat[];
  const uint32_t v = blockIdx.y, lane = __lane_id(), pc = blockIdx.x * kFilterWaves + (threadIdx.x >> 6);
  const ViewDesc vd = views[v];
  const uint32_t T = vd.tiles_x * vd.tiles_y;
  for (uint32_t k = threadIdx.x; k < T; k += 64 * kFilterWaves) lsat[k] = sat[vd.tile_base + k];
  __syncthreads();
  if (pc >= pieces) return;
  const size_t seg = size_t(v) * n;
  const uint32_t begin = pbeg + pc * kFilterPiece, end = min(pend, begin + kFilterPiece);
  uint64_t* mk = mask + size_t(v) * mask_words;
  uint32_t count = 0;
  constexpr uint32_t kAhead = 8;
  for (uint32_t p0 = begin; p0 < end; p0 += 64 * kAhead) {
    uint2 rc[kAhead];
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const uint32_t p = p0 + 64 * u + lane;
      rc[u] = p < end ? load_rect<kPacked>(sorted, seg + p) : empty_rect();
    }
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const uint64_t m = __ballot(meets_open(lsat, vd.tiles_x, unpack_rect<kPacked>(rc[u])));
      const uint32_t g = p0 + 64 * u;
      if (g < end && lane == 0) mk[(g - pbeg) / 64] = m;
      count += uint32_t(__popcll(m));
    }
  }
  if (lane == 0) piece[size_t(v) * pieces + pc] = count;
}

// After the pieces are scanned per view: wave pc copies its kept Gaussians'
// rectangles and indices to out_rects / out_vals[v][piece base + rank], in
// position order (only the kept ones are loaded), and the view's last piece
// writes kept[v].
template <bool kPacked>
__global__ __launch_bounds__(64 * kFilterWaves) void slice_keep_kernel(
    uint32_t n, uint32_t pbeg, uint32_t pend, uint32_t pieces, const uint2* __restrict__ sorted,
    const uint32_t* __restrict__ vals, const uint64_t* __restrict__ mask, uint32_t mask_words,
    const uint32_t* __restrict__ piece, uint2* __restrict__ out_rects, uint32_t* __restrict__ out_vals,
    uint32_t* __restrict__ kept) {
  const uint32_t v = blockIdx.y, lane = __lane_id(), pc = blockIdx.x * kFilterWaves + (threadIdx.x >> 6);
  if (pc >= pieces) return;
  const size_t seg = size_t(v) * n;
  const uint32_t begin = pbeg + pc * kFilterPiece, end = min(pend, begin + kFilterPiece);
  const uint64_t* mk = mask + size_t(v) * mask_words;
  const uint64_t below = (uint64_t(1) << lane) - 1;
  uint32_t run = piece[size_t(v) * pieces + pc];  // wave-uniform
  constexpr uint32_t kAhead = 8;
  for (uint32_t p0 = begin; p0 < end; p0 += 64 * kAhead) {
    uint64_t m[kAhead];
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      const uint32_t g = p0 + 64 * u;
      m[u] = g < end ? mk[(g - pbeg) / 64] : 0ull;
    }
#pragma unroll
    for (uint32_t u = 0; u < kAhead; ++u) {
      if ((m[u] >> lane) & 1u) {  // kept bits lie below end
        const size_t p = seg + p0 + 64 * u + lane;
        const size_t dst = seg + run + uint32_t(__popcll(m[u] & below));
        out_vals[dst] = vals[p];
        if (kPacked)
          reinterpret_cast<uint32_t*>(out_rects)[dst] = reinterpret_cast<const uint32_t*>(sorted)[p];
        else
          out_rects[dst] = sorted[p];
      }
      run += uint32_t(__popcll(m[u]));
    }
  }
  if (lane == 0 && pc == pieces - 1) kept[v] = run;
}

}  // namespace

size_t depth_sort_range_elems(uint32_t n, uint32_t num_views) {
  const size_t chunks = (size_t(n) + kSortChunk - 1) / kSortChunk;
  return 2 * size_t(num_views) * (1 + chunks);
}

size_t radix_hist_elems(uint32_t n, uint32_t num_views, bool wide) {
  size_t chunks = (size_t(n) + kSortChunk - 1) / kSortChunk;
  return size_t(num_views) * (wide ? (1u << kWideBits) : kBins) * chunks;
}

size_t scan_piece_elems(size_t seg_len, uint32_t num_segs) {
  return size_t(num_segs) * ((seg_len + kScanPiece - 1) / kScanPiece);
}

void launch_depth_sort(hipStream_t s, uint32_t* keys, uint32_t n, uint32_t num_views, const SortScratch& sc,
                       const uint2* rects, uint2* rect_sorted, bool packed, uint32_t* range, uint32_t* range_flag) {
  if (n == 0 || num_views == 0) return;
  const uint32_t chunks = (n + kSortChunk - 1) / kSortChunk;
  dim3 grid(chunks, num_views);
  // With range ([V][2] then [V][chunks][2]): three passes, bits 0-7, bits 8-15, then
  // the 9-bit offset of the top 16 bits, for views whose valid keys' top 16 bits
  // span at most 511 values (a depth range within a factor of about 16; 292 on
  // the 32-view bench scene).  A wider view sets range_flag bit 1 and its order
  // is not valid; the caller runs the step again without range.  Without range:
  // four 8-bit passes.  The values end in sc.vals either way.  (With the 9-bit
  // digit in the middle pass, over bits 8-16 of offset << 16 | low 16 bits, the
  // sort took 1.12 instead of 1.07 ms on the bench scene.)
  const int passes = range ? 3 : 32 / kRadixBits;
  uint32_t* kin = keys;
  uint32_t* kout = sc.keys_alt;
  const uint32_t* vin = nullptr;
  uint32_t* vout = (passes & 1) ? sc.vals : sc.vals_alt;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * kRadixBits;
    const bool last = pass == passes - 1;
    const bool offset = range && last;
    const size_t bins = offset ? (1u << kWideBits) : kBins;
    uint2* rd = last ? rect_sorted : nullptr;
    if (offset)
      hipLaunchKernelGGL((radix_hist_kernel<kWideBits, false, true, false>), grid, dim3(kSortThreads), 0, s, kin, n,
                         chunks, shift, sc.hist, range, range_flag);
    else if (range && pass == 0)
      hipLaunchKernelGGL((radix_hist_kernel<kRadixBits, false, false, true>), grid, dim3(kSortThreads), 0, s, kin, n,
                         chunks, shift, sc.hist, range, range_flag);
    else if (shift == 32 - kRadixBits)
      hipLaunchKernelGGL((radix_hist_kernel<kRadixBits, true, false, false>), grid, dim3(kSortThreads), 0, s, kin, n,
                         chunks, shift, sc.hist, range, range_flag);
    else
      hipLaunchKernelGGL((radix_hist_kernel<kRadixBits, false, false, false>), grid, dim3(kSortThreads), 0, s, kin, n,
                         chunks, shift, sc.hist, range, range_flag);
    scan_segments(s, sc.hist, bins * chunks, num_views, sc.piece);
    if (range && pass == 0)
      hipLaunchKernelGGL(range_reduce_kernel, dim3(num_views), dim3(kSortThreads), 0, s, range, chunks, range_flag);
    if (offset)
      hipLaunchKernelGGL((radix_scatter_kernel<kWideBits, true>), grid, dim3(kSortThreads), 0, s, kin, vin, n, chunks,
                         shift, sc.hist, kout, vout, rects, rd, packed, range);
    else
      hipLaunchKernelGGL((radix_scatter_kernel<kRadixBits, false>), grid, dim3(kSortThreads), 0, s, kin, vin, n,
                         chunks, shift, sc.hist, kout, vout, rects, rd, packed, range);
    uint32_t* kt = kin; kin = kout; kout = kt;
    vin = vout;
    vout = (vout == sc.vals_alt) ? sc.vals : sc.vals_alt;
  }
}

size_t bin_matrix_elems(const ViewDesc* host_views, uint32_t num_views, uint32_t n, uint32_t chunk) {
  size_t chunks = (size_t(n) + chunk - 1) / chunk + 1;  // +1: a slice may start mid-chunk
  size_t total = 0;
  for (uint32_t v = 0; v < num_views; ++v) total += chunks * host_views[v].tiles_x * host_views[v].tiles_y;
  return total;
}

namespace {
struct BandGeometry {
  uint32_t max_bands = 1, max_emit_bands = 1, max_tiles = 1;
  size_t count_lds = 4, emit_lds = 4;
};

BandGeometry band_geometry(const ViewDesc* host_views, uint32_t num_views, uint32_t emit_band) {
  BandGeometry g;
  for (uint32_t v = 0; v < num_views; ++v) {
    const uint32_t tx = host_views[v].tiles_x, ty = host_views[v].tiles_y;
    const uint32_t R = band_rows(tx, ty, kBandTiles), Re = band_rows(tx, ty, emit_band);
    g.max_bands = std::max(g.max_bands, (ty + R - 1) / R);
    g.max_emit_bands = std::max(g.max_emit_bands, (ty + Re - 1) / Re);
    g.max_tiles = std::max(g.max_tiles, tx * ty);
    g.count_lds = std::max(g.count_lds, size_t(R + 1) * (tx + 1) * sizeof(int));
    g.emit_lds = std::max(g.emit_lds, kEmitFixedLds + size_t(Re) * tx * (sizeof(uint64_t) + sizeof(uint32_t)));
  }
  return g;
}

uint32_t slice_chunks(const BinPass& p, uint32_t chunk) {
  return p.pend > p.pbeg ? (p.pend - p.pbeg + chunk - 1) / chunk : 0;
}
}  // namespace

void launch_bin_count(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                      uint32_t n, uint32_t chunk, const BinScratch& bs, const BinPass& pass) {
  if (num_views == 0) return;
  const uint32_t chunks = slice_chunks(pass, chunk);
  const BandGeometry g = band_geometry(host_views, num_views, bs.emit_band);
  if (chunks > 0) {
    const dim3 grid(chunks, g.max_bands, num_views);
    // A filtered pass reads its kept Gaussians from position 0.
    const uint2* rs = pass.kept ? pass.rects : bs.rect_sorted;
    const uint32_t pb = pass.kept ? 0u : pass.pbeg, pe = pass.kept ? pass.pend - pass.pbeg : pass.pend;
    if (bs.packed)
      hipLaunchKernelGGL(bin_count_kernel<true>, grid, dim3(64), g.count_lds, s, views, n, chunk, pb, pe, chunks, rs,
                         pass.done, bs.matrix, pass.kept);
    else
      hipLaunchKernelGGL(bin_count_kernel<false>, grid, dim3(64), g.count_lds, s, views, n, chunk, pb, pe, chunks,
                         rs, pass.done, bs.matrix, pass.kept);
    dim3 tgrid((g.max_tiles + 255) / 256, num_views);
    hipLaunchKernelGGL(bin_colsum_kernel, tgrid, dim3(256), 0, s, views, chunks, bs.matrix, bs.tile_tot, chunk,
                       pass.kept);
  }
  hipLaunchKernelGGL(bin_tilescan_kernel, dim3(num_views), dim3(1024), 0, s, views, chunks, bs.tile_tot, pass.ranges,
                     pass.entries);
  hipLaunchKernelGGL(bin_entry_base_kernel, dim3(1), dim3(64), 0, s, num_views, pass.entries, pass.ebase, bs.run,
                     pass.first ? 1 : 0);
}

void launch_bin_emit(hipStream_t s, const ViewDesc* views, const ViewDesc* host_views, uint32_t num_views,
                     uint32_t n, uint32_t chunk, const uint32_t* vals, const BinScratch& bs, const BinPass& pass,
                     uint32_t* order, uint64_t order_capacity, uint32_t* overflow) {
  const uint32_t chunks = slice_chunks(pass, chunk);
  if (num_views == 0 || chunks == 0) return;
  // Later slices bin only unsaturated tiles (a third of them on the 32-view
  // bench scene).  Unfiltered, their bands are four times taller, so each
  // chunk's rectangles are scanned by a quarter as many blocks for about the
  // same pairs (emit 1.90 -> 1.53 ms there; twice and eight times were slower).
  // Filtered, a chunk holds only Gaussians with pairs, so the scans are cheap
  // and the regular bands give four times the blocks (later-slice emit 0.62 ->
  // 0.38 ms on the bench scene, 2.46 -> 1.89 ms on the 3M-Gaussian training
  // scene; twice as tall: 0.45 ms).
  const uint32_t band =
      pass.done && !pass.kept ? std::min(bs.emit_band * kLaterBandScale, kMaxEmitBand) : bs.emit_band;
  const BandGeometry g = band_geometry(host_views, num_views, band);
  dim3 tgrid((g.max_tiles + 255) / 256, num_views);
  hipLaunchKernelGGL(bin_colscan_kernel, tgrid, dim3(256), 0, s, views, chunks, pass.ranges, bs.matrix, chunk,
                     pass.kept);
  const dim3 grid(chunks, g.max_emit_bands, num_views);
  const uint2* rs = pass.kept ? pass.rects : bs.rect_sorted;
  const uint32_t* vs = pass.kept ? pass.vals : vals;
  const uint32_t pb = pass.kept ? 0u : pass.pbeg, pe = pass.kept ? pass.pend - pass.pbeg : pass.pend;
  if (bs.packed)
    hipLaunchKernelGGL(bin_emit_kernel<true>, grid, dim3(64), g.emit_lds, s, views, n, chunk, pb, pe, chunks, vs, rs,
                       pass.done, bs.matrix, pass.ebase, order, order_capacity, overflow, band, pass.kept);
  else
    hipLaunchKernelGGL(bin_emit_kernel<false>, grid, dim3(64), g.emit_lds, s, views, n, chunk, pb, pe, chunks, vs,
                       rs, pass.done, bs.matrix, pass.ebase, order, order_capacity, overflow, band, pass.kept);
}

bool slice_filter_fits(uint32_t max_tiles) { return max_tiles <= kMaxFilterTiles; }

size_t slice_filter_pieces(uint32_t n) { return (size_t(n) + kFilterPiece - 1) / kFilterPiece; }

void launch_slice_filter(hipStream_t s, const ViewDesc* views, uint32_t num_views, uint32_t max_tiles, uint32_t n,
                         uint32_t pbeg, uint32_t pend, const uint2* rect_sorted, const uint32_t* sorted_vals,
                         bool packed, const uint8_t* done, const SliceFilter& f, uint2* rects, uint32_t* vals,
                         uint32_t* kept) {
  if (num_views == 0) return;
  if (pend <= pbeg) {
    launch_zero(s, kept, size_t(num_views) * sizeof(uint32_t));
    return;
  }
  const uint32_t pieces = (pend - pbeg + kFilterPiece - 1) / kFilterPiece;
  const uint32_t mask_words = (n + 63) / 64;
  const dim3 grid((pieces + kFilterWaves - 1) / kFilterWaves, num_views), block(64 * kFilterWaves);
  const size_t lds = (size_t(max_tiles) * sizeof(uint16_t) + 15) / 16 * 16;
  hipLaunchKernelGGL(open_sat_kernel, dim3(num_views), dim3(1024), 0, s, views, done, f.sat);
  if (packed)
    hipLaunchKernelGGL(slice_mark_kernel<true>, grid, block, lds, s, views, n, pbeg, pend, pieces, rect_sorted, f.sat,
                       f.mask, mask_words, f.piece);
  else
    hipLaunchKernelGGL(slice_mark_kernel<false>, grid, block, lds, s, views, n, pbeg, pend, pieces, rect_sorted,
                       f.sat, f.mask, mask_words, f.piece);
  hipLaunchKernelGGL(scan_pieces_kernel, dim3(num_views), dim3(1024), 0, s, f.piece, pieces);
  if (packed)
    hipLaunchKernelGGL(slice_keep_kernel<true>, grid, block, 0, s, n, pbeg, pend, pieces, rect_sorted, sorted_vals,
                       f.mask, mask_words, f.piece, rects, vals, kept);
  else
    hipLaunchKernelGGL(slice_keep_kernel<false>, grid, block, 0, s, n, pbeg, pend, pieces, rect_sorted, sorted_vals,
                       f.mask, mask_words, f.piece, rects, vals, kept);
}

}  // namespace gsr
