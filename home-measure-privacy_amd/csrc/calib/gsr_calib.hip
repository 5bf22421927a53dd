// Measurement kernels for the bench's rooflines (not part of the rasterizer):
// the chip's sustained VALU issue rate, so the blend kernels' VALU instruction
// counts (rocprofv3 SQ_INSTS_VALU) can be set against a measured peak, and the
// issue rate of each instruction class the blends use.
//
// gfx950 can issue two VALU instructions in one quad-cycle (four cycles of a
// 16-lane SIMD; counter SQ_ACTIVE_INST_VALU2 counts such quad-cycles), so plain
// FP32 arithmetic on VGPR operands issues at up to two wave64 instructions per
// four cycles.  Which instructions pair is measured here, not assumed: each
// kernel runs eight independent chains of one instruction form per thread, 16
// deep per loop trip, on 8 blocks of 4 waves per CU.  The forms are written as
// inline assembly so that the instruction and its operand kinds are exactly the
// ones named (an SGPR operand, for one, changes the rate), reading a register of
// their own chain besides the shared one, as the blends' instructions do (eight
// chains reading one register twice, v_fma_f32 x, x, v, v, pair far less).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <utility>

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 16;

// The peak: v_fma_f32 with all three operands in VGPRs.
__global__ __launch_bounds__(256) void valu_peak_kernel(float* __restrict__ out, uint32_t trips, float a, float b) {
  float x[kChains];
  float av = a, bv = b;
  asm volatile("" : "+v"(av), "+v"(bv));  // the operands live in VGPRs
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = float(threadIdx.x + c);
  for (uint32_t t = 0; t < trips; ++t) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) x[c] = fmaf(x[c], av, bv);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;  // keeps every chain live
}

// Instruction forms of the issue table (names: gsr_calib_issue_name).
constexpr int kIssueForms = 10;

// One step of a chain of form K (kIssuePer[K] instructions).
template <int K>
__device__ __forceinline__ void issue_step(float& x, float& y, float v, float s, uint64_t m, uint64_t& sc) {
  if constexpr (K == 0) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(x) : "v"(v), "v"(y));
  if constexpr (K == 1) asm volatile("v_fma_f32 %0, %1, %0, %2" : "+v"(x) : "s"(s), "v"(y));
  if constexpr (K == 2) asm volatile("v_cmp_gt_f32_e64 %0, %1, %2" : "=s"(sc) : "v"(x), "v"(y));
  if constexpr (K == 3) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(x) : "v"(y), "s"(m));
  if constexpr (K == 4) asm volatile("v_min_f32_e32 %0, %0, %1" : "+v"(x) : "v"(y));
  if constexpr (K == 5) asm volatile("v_add_f32_dpp %0, %1, %0 row_shr:1 row_mask:0xf bank_mask:0xf" : "+v"(x) : "v"(y));
  if constexpr (K == 6) asm volatile("v_exp_f32 %0, %0" : "+v"(x));
  if constexpr (K == 7) asm volatile("v_rcp_f32 %0, %0" : "+v"(x));
  if constexpr (K == 8) asm volatile("v_permlane32_swap_b32 %0, %1" : "+v"(x), "+v"(y));
  if constexpr (K == 9) asm volatile("v_fma_f32 %0, %0, %2, %3\n\tv_cmp_gt_f32_e64 %1, %0, %3" : "+v"(x), "=s"(sc) : "v"(v), "v"(y));
}

constexpr uint32_t kIssuePer[kIssueForms] = {1, 1, 1, 1, 1, 1, 1, 1, 1, 2};

template <int K>
__global__ __launch_bounds__(256) void issue_kernel(float* __restrict__ out, uint32_t trips, float s) {
  float x[kChains], y[kChains];
  float v = 1.0f + float(threadIdx.x) * 1e-6f;
  asm volatile("" : "+v"(v));
  const uint64_t m = 0x5555555555555555ull ^ uint64_t(blockIdx.x);  // uniform: an SGPR pair
  uint64_t sc = 0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    x[c] = 0.5f + float(c) * 1e-3f + float(threadIdx.x) * 1e-6f;
    y[c] = 1.0f + float(c) * 1e-4f;
  }
  for (uint32_t t = 0; t < trips; ++t) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) issue_step<K>(x[c], y[c], v, s, m, sc);
    }
  }
  float r = float(uint32_t(sc) ^ uint32_t(sc >> 32));
#pragma unroll
  for (int c = 0; c < kChains; ++c) r += x[c] + y[c];
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = r;
}

using IssueKernel = void (*)(float*, uint32_t, float);
template <int... Ks>
constexpr auto issue_table(std::integer_sequence<int, Ks...>) {
  return std::array<IssueKernel, sizeof...(Ks)>{issue_kernel<Ks>...};
}

// Write-combining probe: the same buffer of n floats (n a multiple of 16384)
// written in full four ways, to see when the L2 merges partial stores into one
// memory write (rocprofv3 WRITE_SIZE per kernel).  A sector is 8 floats.
//   0: coalesced, thread i writes float i;
//   1: each lane fills its own sector with 8 separate stores (8 instructions);
//   2: word k of every sector of a group is written by block 8 g + k: the eight
//      writers of a sector are consecutive workgroups, on eight different dies
//      (workgroup L runs on die L mod 8);
//   3: as 2, but the eight writers are workgroups L, L + 8, ..., L + 56 of a
//      group of 64, all on die L mod 8.
__global__ __launch_bounds__(256) void wcomb_full(float* __restrict__ b, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = float(i);
}
__global__ __launch_bounds__(256) void wcomb_same_wave(float* __restrict__ b, uint32_t n) {
  const uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s >= n / 8) return;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    b[size_t(s) * 8 + k] = float(k);
    __builtin_amdgcn_s_sleep(1);  // separate instructions, apart in time
  }
}
__global__ __launch_bounds__(256) void wcomb_cross_die(float* __restrict__ b, uint32_t n) {
  const uint32_t k = blockIdx.x % 8, g = blockIdx.x / 8;
  const uint32_t s = g * 256 + threadIdx.x;
  if (s < n / 8) b[size_t(s) * 8 + k] = float(k);
}
__global__ __launch_bounds__(256) void wcomb_same_die(float* __restrict__ b, uint32_t n) {
  const uint32_t L = blockIdx.x, r = L % 64;
  const uint32_t g = (L / 64) * 8 + r % 8, k = r / 8;  // group of die r % 8, word r / 8
  const uint32_t s = g * 256 + threadIdx.x;
  if (s < n / 8) b[size_t(s) * 8 + k] = float(k);
}

// Times `launch` (a warm-up launch first) with HIP events on the default stream.
template <class F>
hipError_t time_launch(F&& launch, float* ms) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) return e;
  launch();
  (void)hipEventRecord(e0, 0);
  launch();
  (void)hipEventRecord(e1, 0);
  if ((e = hipEventSynchronize(e1)) == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace

extern "C" {

// Compute units and peak engine clock of device 0.
int gsr_calib_device(int* cus, int* clock_khz) {
  hipError_t e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, 0);
  if (e == hipSuccess) e = hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, 0);
  return int(e);
}

// Runs the peak kernel on the current device's default stream: `blocks` blocks
// of 256 threads, `trips` loop trips each.  *ms receives the kernel time (HIP
// events), *fma_per_thread the v_fma_f32 each thread issued.  out: device
// buffer of blocks * 256 floats.  Returns 0 or the HIP error.
int gsr_calib_valu(float* out, uint32_t blocks, uint32_t trips, float* ms, uint64_t* fma_per_thread) {
  const hipError_t e = time_launch(
      [&] { hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, 0, out, trips, 0.999f, 0.001f); }, ms);
  *fma_per_thread = uint64_t(trips) * kChains * kUnroll;
  return int(e);
}

// Number of instruction forms of the issue table.
int gsr_calib_issue_forms(void) { return kIssueForms; }

// Name of form k: the instruction and its operand kinds.
const char* gsr_calib_issue_name(int k) {
  static const char* const kNames[kIssueForms] = {
      "v_fma_f32, VGPR operands",     "v_fma_f32, one SGPR operand", "v_cmp_gt_f32_e64 (to an SGPR pair)",
      "v_cndmask_b32_e64 (SGPR mask)", "v_min_f32",                   "v_add_f32 with DPP row_shr",
      "v_exp_f32",                    "v_rcp_f32",                   "v_permlane32_swap_b32",
      "v_fma_f32 + v_cmp_gt_f32_e64 pairs"};
  return k >= 0 && k < kIssueForms ? kNames[k] : nullptr;
}

// Runs form k like gsr_calib_valu; *instr_per_thread receives the
// instructions each thread issued.
int gsr_calib_issue(int k, float* out, uint32_t blocks, uint32_t trips, float* ms, uint64_t* instr_per_thread) {
  if (k < 0 || k >= kIssueForms) return int(hipErrorInvalidValue);
  static const auto table = issue_table(std::make_integer_sequence<int, kIssueForms>{});
  const IssueKernel kern = table[k];
  const hipError_t e =
      time_launch([&] { hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, out, trips, 0.999f); }, ms);
  *instr_per_thread = uint64_t(trips) * kChains * kUnroll * kIssuePer[k];
  return int(e);
}

// Runs write-combining probe `mode` (0-3, see above) on b[0, n), n a multiple
// of 16384 floats; *ms receives the kernel time.  Returns 0 or the HIP error.
int gsr_calib_wcomb(int mode, float* b, uint32_t n, float* ms) {
  if (mode < 0 || mode > 3 || n == 0 || n % 16384 != 0) return int(hipErrorInvalidValue);
  const uint32_t sectors = n / 8;
  hipError_t e;
  switch (mode) {
    case 0: e = time_launch([&] { hipLaunchKernelGGL(wcomb_full, dim3(n / 256), dim3(256), 0, 0, b, n); }, ms); break;
    case 1: e = time_launch([&] { hipLaunchKernelGGL(wcomb_same_wave, dim3(sectors / 256), dim3(256), 0, 0, b, n); }, ms); break;
    case 2: e = time_launch([&] { hipLaunchKernelGGL(wcomb_cross_die, dim3(sectors / 256 * 8), dim3(256), 0, 0, b, n); }, ms); break;
    default: e = time_launch([&] { hipLaunchKernelGGL(wcomb_same_die, dim3(sectors / 256 * 8), dim3(256), 0, 0, b, n); }, ms); break;
  }
  return int(e);
}

}  // extern "C"
