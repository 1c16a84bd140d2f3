// Procedural noise fixtures straight into a device tree: generate_noise_tree
// (voxel-gpu/src/generate.rs:99) over NoiseField (voxel-core/src/fixtures.rs:259-346,
// noise.rs). The field is evaluated in f64 with the reference CPU field's operations in
// its order (no contraction), so the occupancy is the CPU NoiseField's bit for bit -- the
// reference's f32 GPU port disagrees with it at threshold-grazing voxels.
//
// One wave per 4^3 sub-brick word: lane l is the voxel with intra-word Morton index l, so
// the wave's ballot of "occupied" is the leaf word itself, written once into the dense
// leaf grid (no atomics on words; the brick's bitmap bit by one atomic OR). The usual
// tree build follows on the device; nothing is read back.
#include "vxl_internal.hpp"

namespace vxl {
namespace {

__device__ __forceinline__ uint32_t noise_hash(uint32_t seed, int64_t ix, int64_t iy, int64_t iz) {
  const uint32_t x = uint32_t(ix), y = uint32_t(iy), z = uint32_t(iz);
  uint32_t h = seed ^ (x * 0x8da6b343u) ^ (y * 0xd8163841u) ^ (z * 0xcb1ab31fu);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h;
}

__device__ __forceinline__ double noise_grad(uint32_t hash, double x, double y, double z) {
  const uint32_t h = hash & 15u;
  double u = h < 8 ? x : y;
  const double v0 = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
  double v = v0;
  if (h & 1u) u = -u;
  if (h & 2u) v = -v;
  return u + v;
}

__device__ __forceinline__ double noise_fade(double t) {
  return t * t * t * (t * (t * 6.0 - 15.0) + 10.0);
}

__device__ __forceinline__ double noise_lerp(double a, double b, double t) { return a + t * (b - a); }

__device__ double perlin3(uint32_t seed, double x, double y, double z) {
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int64_t ix = int64_t(fx), iy = int64_t(fy), iz = int64_t(fz);
  const double xf = x - fx, yf = y - fy, zf = z - fz;
  const double u = noise_fade(xf), v = noise_fade(yf), w = noise_fade(zf);
  const double g000 = noise_grad(noise_hash(seed, ix, iy, iz), xf, yf, zf);
  const double g100 = noise_grad(noise_hash(seed, ix + 1, iy, iz), xf - 1.0, yf, zf);
  const double g010 = noise_grad(noise_hash(seed, ix, iy + 1, iz), xf, yf - 1.0, zf);
  const double g110 = noise_grad(noise_hash(seed, ix + 1, iy + 1, iz), xf - 1.0, yf - 1.0, zf);
  const double g001 = noise_grad(noise_hash(seed, ix, iy, iz + 1), xf, yf, zf - 1.0);
  const double g101 = noise_grad(noise_hash(seed, ix + 1, iy, iz + 1), xf - 1.0, yf, zf - 1.0);
  const double g011 = noise_grad(noise_hash(seed, ix, iy + 1, iz + 1), xf, yf - 1.0, zf - 1.0);
  const double g111 =
      noise_grad(noise_hash(seed, ix + 1, iy + 1, iz + 1), xf - 1.0, yf - 1.0, zf - 1.0);
  const double x00 = noise_lerp(g000, g100, u);
  const double x10 = noise_lerp(g010, g110, u);
  const double x01 = noise_lerp(g001, g101, u);
  const double x11 = noise_lerp(g011, g111, u);
  const double y0 = noise_lerp(x00, x10, v), y1 = noise_lerp(x01, x11, v);
  return noise_lerp(y0, y1, w);
}

__device__ double fractal(uint32_t seed, double px, double py, double pz, uint32_t octaves,
                          double lac, double gain, bool ridged) {
  double amp = 1.0, freq = 1.0, sum = 0.0, norm = 0.0;
  for (uint32_t o = 0; o < octaves; o++) {
    const double off = double(o);
    const double n = perlin3(seed + o * 0x9e3779b9u, px * freq + off * 17.13,
                             py * freq - off * 9.71, pz * freq + off * 5.37);
    double contrib = n;
    if (ridged) {
      const double r = 1.0 - fabs(n);
      contrib = r * r;
    }
    sum += amp * contrib;
    norm += amp;
    amp *= gain;
    freq *= lac;
  }
  const double mean = norm > 0.0 ? sum / norm : 0.0;
  return ridged ? mean * 2.0 - 1.0 : mean;
}

// NoiseField::value (fixtures.rs:316-346).
__device__ double noise_value(const vxl_noise &f, uint32_t n, uint32_t x, uint32_t y, uint32_t z) {
  const double s = f.frequency / double(n);
  double px = (double(x) + 0.5) * s, py = (double(y) + 0.5) * s, pz = (double(z) + 0.5) * s;
  const bool ridged = f.ridged != 0;
  if (f.warp > 0.0) {
    const uint32_t wo = f.octaves < 3 ? f.octaves : 3;
    const double wx = fractal(f.seed ^ 0x11111111u, px + 1.7, py + 9.2, pz + 3.3, wo, f.lacunarity,
                              f.gain, ridged);
    const double wy = fractal(f.seed ^ 0x22222222u, px - 5.1, py + 2.8, pz - 7.4, wo, f.lacunarity,
                              f.gain, ridged);
    const double wz = fractal(f.seed ^ 0x33333333u, px + 8.6, py - 4.5, pz + 1.9, wo, f.lacunarity,
                              f.gain, ridged);
    px = px + f.warp * wx;
    py = py + f.warp * wy;
    pz = pz + f.warp * wz;
  }
  return fractal(f.seed, px, py, pz, f.octaves, f.lacunarity, f.gain, ridged);
}

__global__ __launch_bounds__(256) void k_noise_words(vxl_noise f, uint32_t n, uint64_t nwords,
                                                     unsigned long long *__restrict__ leafgrid,
                                                     unsigned long long *__restrict__ b1) {
  const uint32_t lane = threadIdx.x & 63;
  // grid-stride over words, one word per wave at a time (wave-uniform bounds)
  for (uint64_t gw = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); gw < nwords;
       gw += uint64_t(gridDim.x) * 4) {
  const uint32_t code = uint32_t(gw >> 3);
  const uint32_t m = uint32_t(gw & 7) * 64 + lane;  // intra-brick Morton index
  const uint32_t lx = (m & 1u) | ((m >> 2) & 2u) | ((m >> 4) & 4u);
  const uint32_t ly = ((m >> 1) & 1u) | ((m >> 3) & 2u) | ((m >> 5) & 4u);
  const uint32_t lz = ((m >> 2) & 1u) | ((m >> 4) & 2u) | ((m >> 6) & 4u);
  const uint32_t x = dev::compact3_10(code) * 8 + lx, y = dev::compact3_10(code >> 1) * 8 + ly,
                 z = dev::compact3_10(code >> 2) * 8 + lz;
  const bool occ = noise_value(f, n, x, y, z) > f.threshold;
  const unsigned long long word = __ballot(occ);
  if (lane == 0 && word) {
    leafgrid[gw] = word;
    atomicOr(&b1[code >> 6], 1ull << (code & 63));
  }
  }
}

}  // namespace

vxl_tree *generate_noise_tree(vxl_ctx *ctx, uint32_t n, uint32_t k, const vxl_noise &f) {
  prepare_grid(ctx, n);
  const uint64_t nb = uint64_t(n / 8);
  const uint64_t nwords = nb * nb * nb * 8;
  // a launch holds at most 2^32 - 1 threads per dimension: cap the grid, the kernel strides
  const uint64_t want = (nwords + 3) / 4;
  const uint64_t blocks = want < (1ull << 22) ? want : (1ull << 22);
  hipLaunchKernelGGL(k_noise_words, dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, f, n, nwords,
                     ctx->leafgrid.as<unsigned long long>(), ctx->bitmap.as<unsigned long long>());
  VXL_HIP(hipGetLastError());
  return build_tree_from_grid(ctx, n, k, nullptr, nullptr);
}

}  // namespace vxl
