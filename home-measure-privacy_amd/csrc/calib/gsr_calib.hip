// Measurement kernels for the bench's rooflines (not part of the rasterizer):
// the chip's sustained VALU issue rate, so the blend kernels' VALU instruction
// counts (rocprofv3 SQ_INSTS_VALU) can be set against a measured peak.
//
// Each thread runs kChains independent fma chains, kUnroll deep per loop trip:
// kChains * kUnroll v_fma_f32 per trip (packed FP32 is disabled in this build,
// so each is one non-packed VALU instruction for the wave), plus scalar loop
// control, which issues on the scalar unit.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 16;

__global__ __launch_bounds__(256) void valu_peak_kernel(float* __restrict__ out, uint32_t trips, float a, float b) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = float(threadIdx.x + c);
  for (uint32_t t = 0; t < trips; ++t) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
      for (int c = 0; c < kChains; ++c) x[c] = fmaf(x[c], a, b);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = s;  // keeps every chain live
}

}  // namespace

extern "C" {

// Compute units and peak engine clock of device 0.
int gsr_calib_device(int* cus, int* clock_khz) {
  hipError_t e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, 0);
  if (e == hipSuccess) e = hipDeviceGetAttribute(clock_khz, hipDeviceAttributeClockRate, 0);
  return int(e);
}

// Runs the kernel on the current device's default stream: `blocks` blocks of
// 256 threads, `trips` loop trips each.  *ms receives the kernel time (HIP
// events), *fma_per_thread the v_fma_f32 each thread issued.  out: device
// buffer of blocks * 256 floats.  Returns 0 or the HIP error.
int gsr_calib_valu(float* out, uint32_t blocks, uint32_t trips, float* ms, uint64_t* fma_per_thread) {
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return int(e);
  if ((e = hipEventCreate(&e1)) != hipSuccess) return int(e);
  hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, 0, out, trips, 0.999f, 0.001f);  // warm
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, 0, out, trips, 0.999f, 0.001f);
  (void)hipEventRecord(e1, 0);
  if ((e = hipEventSynchronize(e1)) != hipSuccess) return int(e);
  if ((e = hipEventElapsedTime(ms, e0, e1)) != hipSuccess) return int(e);
  *fma_per_thread = uint64_t(trips) * kChains * kUnroll;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return int(hipGetLastError());
}

}  // extern "C"
