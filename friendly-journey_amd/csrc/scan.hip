// Device-wide exclusive prefix sums (reduce -> scan block sums -> downsweep).
// Blocks of 256 threads x 8 items; wave64 inclusive scans via DPP-friendly shuffles.
#include "vxl_internal.hpp"

namespace vxl {
namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr uint64_t kTile = uint64_t(kThreads) * kItems;

template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across the block; returns the block total.
template <class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *excl) {
  __shared__ T wave_tot[kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T inc = wave_inclusive_scan(v);
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  T before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; w++) {
    T t = wave_tot[w];
    if (w < wave) before += t;
    total += t;
  }
  *excl = before + inc - v;
  __syncthreads();
  return total;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_reduce(const T *__restrict__ in, uint64_t n,
                                                       T *__restrict__ sums) {
  const uint64_t base = uint64_t(blockIdx.x) * kTile + uint64_t(threadIdx.x) * kItems;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t j = base + i;
    if (j < n) s += in[j];
  }
  T excl;
  T tot = block_exclusive_scan(s, &excl);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_scan(const T *__restrict__ in, T *__restrict__ out,
                                                     uint64_t n, const T *__restrict__ offsets,
                                                     T *__restrict__ total) {
  const uint64_t base = uint64_t(blockIdx.x) * kTile + uint64_t(threadIdx.x) * kItems;
  T v[kItems];
  T s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t j = base + i;
    v[i] = j < n ? in[j] : T(0);
    s += v[i];
  }
  T excl;
  block_exclusive_scan(s, &excl);
  T run = excl + (offsets ? offsets[blockIdx.x] : T(0));
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t j = base + i;
    if (j < n) {
      out[j] = run;
      if (j == n - 1 && total) *total = run + v[i];
    }
    run += v[i];
  }
}

template <class T>
__global__ void k_set_total_zero(T *total) { *total = 0; }

template <class T>
void scan_rec(const T *in, T *out, uint64_t n, T *total, T *tmp, hipStream_t s) {
  if (n == 0) {
    if (total) hipLaunchKernelGGL(k_set_total_zero<T>, dim3(1), dim3(1), 0, s, total);
    return;
  }
  uint64_t nb = (n + kTile - 1) / kTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan<T>, dim3(1), dim3(kThreads), 0, s, in, out, n, nullptr, total);
    return;
  }
  T *sums = tmp;
  T *rest = tmp + nb;
  hipLaunchKernelGGL(k_reduce<T>, dim3(nb), dim3(kThreads), 0, s, in, n, sums);
  scan_rec<T>(sums, sums, nb, nullptr, rest, s);  // in-place exclusive scan of block sums
  hipLaunchKernelGGL(k_scan<T>, dim3(nb), dim3(kThreads), 0, s, in, out, n, sums, total);
}

template <class T>
uint64_t tmp_elems(uint64_t n) {
  uint64_t t = 0;
  while (n > kTile) {
    n = (n + kTile - 1) / kTile;
    t += n;
  }
  return t + 1;
}

}  // namespace

void scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total,
                        DevBuf &tmp, hipStream_t s) {
  tmp.reserve(tmp_elems<uint64_t>(n) * sizeof(uint64_t));
  scan_rec<uint64_t>(in, out, n, total, tmp.as<uint64_t>(), s);
  VXL_HIP(hipGetLastError());
}

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total,
                        DevBuf &tmp, hipStream_t s) {
  tmp.reserve(tmp_elems<uint32_t>(n) * sizeof(uint32_t));
  scan_rec<uint32_t>(in, out, n, total, tmp.as<uint32_t>(), s);
  VXL_HIP(hipGetLastError());
}

}  // namespace vxl
