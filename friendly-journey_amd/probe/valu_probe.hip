// VALU issue-rate probe: measures the non-packed, non-FMA FP32 and 32-bit integer lane-op
// rate of the device, the roof the voxelize SAT kernel and the DDA walk are measured against.
// Each lane runs kChains independent dependency chains of a binary32 multiply then add
// (two VALU instructions, never fused: built with -ffp-contract=off) -- or an integer add
// then xor -- so the count of VALU instructions per lane is known exactly:
// iters * kUnroll * kChains * 2. Measurement infrastructure, not part of the C ABI in vxl.h.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 4;

__global__ __launch_bounds__(256) void k_valu_f32(float *__restrict__ out, int iters, float a,
                                                  float b) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++) x[c] = float(threadIdx.x + c) * 1e-3f;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++) {
        x[c] = x[c] * a;
        x[c] = x[c] + b;
      }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; c++) s += x[c];
  if (s == 12345.678f) out[blockIdx.x] = s;  // keeps the chains live; never true here
}

__global__ __launch_bounds__(256) void k_valu_u32(uint32_t *__restrict__ out, int iters,
                                                  uint32_t a, uint32_t b) {
  uint32_t x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++) x[c] = threadIdx.x * 2654435761u + c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++) {
        x[c] = x[c] + a;
        x[c] = x[c] ^ b;
      }
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < kChains; c++) s ^= x[c];
  if (s == 0x9E3779B9u) out[blockIdx.x] = s;
}

}  // namespace

extern "C" {

// Runs the probe on the current device with `blocks` x 256 threads, `iters` loop trips;
// kind 0 = binary32 mul + add, 1 = u32 add + xor. *ms: kernel time (HIP events), *lane_ops:
// the VALU lane-instructions executed. Returns 0 on success, else the HIP error code.
int vxl_probe_valu(int kind, int blocks, int iters, double *ms, uint64_t *lane_ops) {
  void *buf = nullptr;
  hipError_t e = hipMalloc(&buf, size_t(blocks) * 4 + 256);
  if (e != hipSuccess) return int(e);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int rep = 0; rep < 2; rep++) {  // the first launch loads the code object
    (void)hipEventRecord(e0, nullptr);
    if (kind == 0)
      hipLaunchKernelGGL(k_valu_f32, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<float *>(buf), iters, 0.9999999f, 1e-7f);
    else
      hipLaunchKernelGGL(k_valu_u32, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<uint32_t *>(buf), iters, 0x9E3779B9u, 0x7F4A7C15u);
    (void)hipEventRecord(e1, nullptr);
  }
  e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  *ms = t;
  *lane_ops = uint64_t(blocks) * 256ull * uint64_t(iters) * kUnroll * kChains * 2ull;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  return int(e);
}

}  // extern "C"
