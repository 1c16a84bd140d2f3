// VALU issue-rate probe: measures the rate at which the device issues non-packed FP32 and
// 32-bit integer VALU instructions, the roof the voxelize SAT kernel and the DDA walk are
// measured against, and how that rate depends on the instruction form and on the number of
// active lanes. Each lane runs kChains independent dependency chains, two instructions per
// link, so the count of VALU instructions per wave is known exactly:
// iters * kUnroll * kChains * 2. Measurement infrastructure, not part of the C ABI in vxl.h.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kChains = 8;
constexpr int kUnroll = 4;

// A wave-uniform value in a vector register: the compiler would keep it in a scalar
// register and use that as the instruction's operand, which is a different issue class.
__device__ __forceinline__ float in_vgpr(float s) {
  float v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}
__device__ __forceinline__ uint32_t in_vgpr(uint32_t s) {
  uint32_t v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// Binary32 multiply then add (never fused: built with -ffp-contract=off) by the first
// `active` lanes of each wave; operands in vector registers, or (kSgpr) the two constants
// in scalar registers as the compiler places wave-uniform values.
template <bool kSgpr>
__global__ __launch_bounds__(256) void k_valu_f32(float *__restrict__ out, int iters, float a0,
                                                  float b0, int active) {
  if (int(threadIdx.x & 63u) >= active) return;
  const float a = kSgpr ? a0 : in_vgpr(a0), b = kSgpr ? b0 : in_vgpr(b0);
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

// 32-bit integer add then xor, operands in vector registers, written in inline asm: from C
// the compiler fuses each xor with the next add into one three-source v_xad_u32, a
// different instruction form.
__global__ __launch_bounds__(256) void k_valu_u32(uint32_t *__restrict__ out, int iters,
                                                  uint32_t a0, uint32_t b0) {
  const uint32_t a = in_vgpr(a0), b = in_vgpr(b0);
  uint32_t x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; c++) x[c] = threadIdx.x * 2654435761u + c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++)
        asm volatile("v_add_u32 %0, %0, %1\n\tv_xor_b32 %0, %0, %2" : "+v"(x[c]) : "v"(a), "v"(b));
  }
  uint32_t s = 0;
#pragma unroll
  for (int c = 0; c < kChains; c++) s ^= x[c];
  if (s == 0x9E3779B9u) out[blockIdx.x] = s;
}

// Chains of one instruction form written out in inline asm, two per link, so the compiler
// neither merges nor drops them. kForms lists them; the compare writes VCC only.
constexpr int kForms = 8;
template <int kOp>
__global__ __launch_bounds__(256) void k_valu_form(float *__restrict__ out, int iters, float s0,
                                                   uint64_t mask) {
  float x[kChains];
  const float y = float(threadIdx.x) * 1e-9f + 0.5f, z = float(threadIdx.x) * 1e-9f + 0.25f;
#pragma unroll
  for (int c = 0; c < kChains; c++) x[c] = float(threadIdx.x + c);
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
#pragma unroll
      for (int c = 0; c < kChains; c++) {
        if (kOp == 0)
          asm volatile("v_add_f32 %0, %0, %1\n\tv_add_f32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
        else if (kOp == 1)
          asm volatile("v_add_f32 %0, %1, %0\n\tv_add_f32 %0, %1, %0" : "+v"(x[c]) : "s"(s0));
        else if (kOp == 2)
          asm volatile("v_fma_f32 %0, %0, %1, %2\n\tv_fma_f32 %0, %0, %1, %2"
                       : "+v"(x[c])
                       : "v"(y), "v"(z));
        else if (kOp == 3)
          asm volatile("v_max_f32 %0, %0, %1\n\tv_max_f32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
        else if (kOp == 4)
          asm volatile("v_min3_f32 %0, %0, %1, %2\n\tv_min3_f32 %0, %0, %1, %2"
                       : "+v"(x[c])
                       : "v"(y), "v"(z));
        else if (kOp == 5)
          asm volatile("v_cmp_gt_f32 vcc, %0, %1\n\tv_cmp_gt_f32 vcc, %1, %0"
                       :
                       : "v"(x[c]), "v"(y)
                       : "vcc");
        else if (kOp == 6)
          asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2\n\tv_cndmask_b32_e64 %0, %0, %1, %2"
                       : "+v"(x[c])
                       : "v"(y), "s"(mask));
        else
          asm volatile("v_and_b32 %0, %0, %1\n\tv_and_b32 %0, %0, %1" : "+v"(x[c]) : "v"(y));
      }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; c++) s += x[c];
  if (s == 12345.678f) out[blockIdx.x] = s;
}

template <int kOp>
void launch_form(int blocks, float *o, int iters) {
  hipLaunchKernelGGL(k_valu_form<kOp>, dim3(blocks), dim3(256), 0, nullptr, o, iters, 0.5f,
                     0x5555555555555555ull);
}

using Launch = void (*)(int, float *, int);
constexpr Launch kFormLaunch[kForms] = {launch_form<0>, launch_form<1>, launch_form<2>,
                                        launch_form<3>, launch_form<4>, launch_form<5>,
                                        launch_form<6>, launch_form<7>};

// Times launch(buf) twice (the first launch loads the code object) with HIP events.
template <typename F>
int timed(int blocks, double *ms, F launch) {
  void *buf = nullptr;
  hipError_t e = hipMalloc(&buf, size_t(blocks) * 4 + 256);
  if (e != hipSuccess) return int(e);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int rep = 0; rep < 2; rep++) {
    (void)hipEventRecord(e0, nullptr);
    launch(buf);
    (void)hipEventRecord(e1, nullptr);
  }
  e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  *ms = t;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  return int(e);
}

}  // namespace

extern "C" {

// Runs the probe on the current device with `blocks` x 256 threads and `iters` loop trips:
// kind 0 = binary32 mul + add with vector-register operands, 1 = u32 add + xor (vector
// registers), 2 = binary32 mul + add with the constants in scalar registers. *ms: kernel time
// (HIP events), *lane_ops: the VALU lane-instructions executed. Returns 0 on success, else
// the HIP error code.
int vxl_probe_valu(int kind, int blocks, int iters, double *ms, uint64_t *lane_ops) {
  if (kind < 0 || kind > 2) return int(hipErrorInvalidValue);
  *lane_ops = uint64_t(blocks) * 256ull * uint64_t(iters) * kUnroll * kChains * 2ull;
  return timed(blocks, ms, [&](void *buf) {
    if (kind == 0)
      hipLaunchKernelGGL(k_valu_f32<false>, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<float *>(buf), iters, 0.9999999f, 1e-7f, 64);
    else if (kind == 1)
      hipLaunchKernelGGL(k_valu_u32, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<uint32_t *>(buf), iters, 0x9E3779B9u, 0x7F4A7C15u);
    else
      hipLaunchKernelGGL(k_valu_f32<true>, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<float *>(buf), iters, 0.9999999f, 1e-7f, 64);
  });
}

// The binary32 chains of kind 0 (sgpr = 0) or kind 2 (sgpr = 1) run by the first `active`
// lanes (1..64) of each wave: *wave_instr the VALU wave-instructions of the chains (every
// wave issues them whatever its lane count).
int vxl_probe_valu_lanes(int sgpr, int active, int blocks, int iters, double *ms,
                         uint64_t *wave_instr) {
  if (active < 1 || active > 64) return int(hipErrorInvalidValue);
  *wave_instr = uint64_t(blocks) * 4ull * uint64_t(iters) * kUnroll * kChains * 2ull;
  return timed(blocks, ms, [&](void *buf) {
    if (sgpr)
      hipLaunchKernelGGL(k_valu_f32<true>, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<float *>(buf), iters, 0.9999999f, 1e-7f, active);
    else
      hipLaunchKernelGGL(k_valu_f32<false>, dim3(blocks), dim3(256), 0, nullptr,
                         static_cast<float *>(buf), iters, 0.9999999f, 1e-7f, active);
  });
}

// Chains of one instruction form, every lane active: op 0 = v_add_f32 (vector operands),
// 1 = v_add_f32 with a scalar-register operand, 2 = v_fma_f32, 3 = v_max_f32,
// 4 = v_min3_f32, 5 = v_cmp_gt_f32 (to VCC), 6 = v_cndmask_b32 (VOP3, mask in a scalar
// register pair), 7 = v_and_b32. *wave_instr: the VALU wave-instructions of the chains.
int vxl_probe_valu_form(int op, int blocks, int iters, double *ms, uint64_t *wave_instr) {
  if (op < 0 || op >= kForms) return int(hipErrorInvalidValue);
  *wave_instr = uint64_t(blocks) * 4ull * uint64_t(iters) * kUnroll * kChains * 2ull;
  return timed(blocks, ms,
               [&](void *buf) { kFormLaunch[op](blocks, static_cast<float *>(buf), iters); });
}

}  // extern "C"
