// Conservative triangle/box SAT voxelization into the dense leaf grid.
//
// Stage 1 (k_tri_setup): world->grid transform (voxelizer/src/core.rs:101-107) and the
//   epsilon-padded, clamped candidate box of reference_cpu.rs:135-150; count of
//   candidate voxels per triangle.
// Stage 2 (scans + k_compact): drop triangles with no candidates; exclusive scan of
//   candidate counts gives every (triangle, voxel) pair a global index.
// Stage 3 (k_bstart): for each block-sized chunk of pairs, the triangle holding its
//   first pair.
// Stage 4 (k_sat): one lane per pair. The lane's triangle is found by a binary search
//   in LDS over the chunk's window of triangle offsets; the pair decodes to its voxel
//   in the reference's scan order (x fastest, reference_cpu.rs:152-154); the lane runs
//   the 13-axis SAT of reference_cpu.rs:27-85 with the same binary32 operations in
//   the same order (no FMA contraction: built -ffp-contract=off); an overlapping voxel
//   sets its bit in the dense leaf grid with a 64-bit atomic OR, and the first bit of
//   a 4^3 sub-brick word sets the brick's bit in the brick bitmap.
//
// Bit-exactness notes:
//  * predicate order is free: the result is a pure conjunction without side effects;
//  * an axis component that is identically 0 contributes v*0 = +-0 to a dot product;
//    adding +-0 is exact, so those products are dropped when the triangle's grid
//    vertices are finite (changes only the sign of a zero, which no comparison sees).
//    Triangles with a non-finite grid vertex take the literal path with every product.
//  * normal.dot(ZERO) + d (reference_cpu.rs:79) equals d up to the sign of zero when the
//    normal is finite; with a non-finite normal it is NaN and the plane test passes.
#include "vxl_internal.hpp"

namespace vxl {
namespace {

struct __align__(16) TriRec {
  float g[9];         // grid-space vertices
  int32_t mn[3];      // candidate box min (inclusive)
  uint32_t ex;        // box extent along x
  uint32_t exy;       // ex * ey
  uint32_t flags;     // bit 0: literal path (non-finite grid vertex); bit 1: count >= 2^32
  uint32_t pad;
};
static_assert(sizeof(TriRec) == 64, "TriRec is one 64-byte record");

__device__ __forceinline__ int32_t sat_i32(float v) {
  // Rust `f32 as i32`: saturating, NaN -> 0
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

__device__ __forceinline__ void tri_grid(const float *__restrict__ tri, const VoxelizeParams &p,
                                         float g[9]) {
#pragma unroll
  for (int v = 0; v < 3; v++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float s = p.inv * tri[3 * v + a];
      g[3 * v + a] = s + p.tx[a];
    }
}

__global__ void k_tri_setup(const float *__restrict__ tris, uint64_t T, VoxelizeParams p,
                            uint64_t *__restrict__ count, uint32_t *__restrict__ flag) {
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= T) return;
  float g[9];
  tri_grid(tris + 9 * t, p, g);
  uint64_t c = 1;
  const float hi_clamp = (float)(p.n - 1);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float lo = fminf(fminf(g[a], g[3 + a]), g[6 + a]) - p.eps;
    float hi = fmaxf(fmaxf(g[a], g[3 + a]), g[6 + a]) + p.eps;
    int32_t mn = sat_i32(fmaxf(floorf(lo), 0.0f));
    int32_t mx = sat_i32(fminf(floorf(hi), hi_clamp));
    c = mx < mn ? 0 : c * uint64_t(int64_t(mx) - int64_t(mn) + 1);
  }
  count[t] = c;
  flag[t] = c != 0;
}

__global__ void k_compact(const float *__restrict__ tris, uint64_t T, VoxelizeParams p,
                          const uint64_t *__restrict__ count, const uint32_t *__restrict__ flag,
                          const uint32_t *__restrict__ cidx, TriRec *__restrict__ rec,
                          uint64_t *__restrict__ ccount) {
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= T || !flag[t]) return;
  TriRec r;
  tri_grid(tris + 9 * t, p, r.g);
  bool finite = true;
#pragma unroll
  for (int i = 0; i < 9; i++) finite = finite && isfinite(r.g[i]);
  uint32_t ext[3];
  const float hi_clamp = (float)(p.n - 1);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    float lo = fminf(fminf(r.g[a], r.g[3 + a]), r.g[6 + a]) - p.eps;
    float hi = fmaxf(fmaxf(r.g[a], r.g[3 + a]), r.g[6 + a]) + p.eps;
    int32_t mn = sat_i32(fmaxf(floorf(lo), 0.0f));
    int32_t mx = sat_i32(fminf(floorf(hi), hi_clamp));
    r.mn[a] = mn;
    ext[a] = uint32_t(mx - mn + 1);
  }
  uint64_t c = count[t];
  r.ex = ext[0];
  r.exy = ext[0] * ext[1];
  r.flags = (finite ? 0u : 1u) | ((c >> 32) ? 2u : 0u);
  r.pad = 0;
  uint32_t ci = cidx[t];
  rec[ci] = r;
  ccount[ci] = c;
}

// bstart[b] = compacted triangle holding pair b * chunk.
__global__ void k_bstart(const uint64_t *__restrict__ coff, const uint64_t *__restrict__ ccount,
                         uint64_t Tc, uint64_t chunk, uint64_t nchunks,
                         uint32_t *__restrict__ bstart) {
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= Tc) return;
  uint64_t lo = coff[t], hi = lo + ccount[t];
  uint64_t b = (lo + chunk - 1) / chunk;
  for (; b * chunk < hi && b < nchunks; b++) bstart[b] = uint32_t(t);
}

// ---- the SAT predicate, reference_cpu.rs:27-85 -----------------------------------------
// Returns true when the unit box centred at (cx,cy,cz) overlaps the triangle g.
template <bool kLiteral>
__device__ __forceinline__ bool sat_overlap(float cx, float cy, float cz, const float *g) {
  const float h = 0.5f;
  const float v0x = g[0] - cx, v0y = g[1] - cy, v0z = g[2] - cz;
  const float v1x = g[3] - cx, v1y = g[4] - cy, v1z = g[5] - cz;
  const float v2x = g[6] - cx, v2y = g[7] - cy, v2z = g[8] - cz;
  const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
  const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
  const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;

  // plane test (:76-82): normal = e0 x e1, d = -normal.dot(v0)
  {
    const float nx = e0y * e1z - e1y * e0z;
    const float ny = e0z * e1x - e1z * e0x;
    const float nz = e0x * e1y - e1x * e0y;
    float dd = nx * v0x;
    dd = dd + ny * v0y;
    dd = dd + nz * v0z;
    const float d = -dd;
    float r = h * fabsf(nx);
    r = r + h * fabsf(ny);
    r = r + h * fabsf(nz);
    float s;
    if (kLiteral) {
      float z = nx * 0.0f;
      z = z + ny * 0.0f;
      z = z + nz * 0.0f;
      s = z + d;
    } else {
      // +-0 + d == d up to the sign of zero for a finite normal; NaN otherwise.
      const bool fin = isfinite(nx) && isfinite(ny) && isfinite(nz);
      s = fin ? d : __builtin_nanf("");
    }
    if (fabsf(s) > r) return false;
  }
  // box-face axes (:66-74)
  if (fminf(v0x, fminf(v1x, v2x)) > h || fmaxf(v0x, fmaxf(v1x, v2x)) < -h) return false;
  if (fminf(v0y, fminf(v1y, v2y)) > h || fmaxf(v0y, fmaxf(v1y, v2y)) < -h) return false;
  if (fminf(v0z, fminf(v1z, v2z)) > h || fmaxf(v0z, fmaxf(v1z, v2z)) < -h) return false;

  // the 9 edge x box-axis tests (:42-64); axis (ax, ay, az), p_i = v_i . axis
#define VXL_AXIS(AX, AY, AZ)                                                          \
  {                                                                                   \
    const float ax = (AX), ay = (AY), az = (AZ);                                      \
    float p0, p1, p2;                                                                 \
    if (kLiteral) {                                                                   \
      p0 = v0x * ax; p0 = p0 + v0y * ay; p0 = p0 + v0z * az;                          \
      p1 = v1x * ax; p1 = p1 + v1y * ay; p1 = p1 + v1z * az;                          \
      p2 = v2x * ax; p2 = p2 + v2y * ay; p2 = p2 + v2z * az;                          \
    } else {                                                                          \
      p0 = VXL_DOT(v0x, v0y, v0z); p1 = VXL_DOT(v1x, v1y, v1z); p2 = VXL_DOT(v2x, v2y, v2z); \
    }                                                                                 \
    const float mnp = fminf(p0, fminf(p1, p2));                                       \
    const float mxp = fmaxf(p0, fmaxf(p1, p2));                                       \
    float r = h * fabsf(ax);                                                          \
    r = r + h * fabsf(ay);                                                            \
    r = r + h * fabsf(az);                                                            \
    if (mnp > r || mxp < -r) return false;                                            \
  }
  // axes (0, -e.z, e.y): x term dropped
#define VXL_DOT(X, Y, Z) ((Y) * ay + (Z) * az)
  VXL_AXIS(0.0f, -e0z, e0y)
  VXL_AXIS(0.0f, -e1z, e1y)
  VXL_AXIS(0.0f, -e2z, e2y)
#undef VXL_DOT
  // axes (e.z, 0, -e.x): y term dropped
#define VXL_DOT(X, Y, Z) ((X) * ax + (Z) * az)
  VXL_AXIS(e0z, 0.0f, -e0x)
  VXL_AXIS(e1z, 0.0f, -e1x)
  VXL_AXIS(e2z, 0.0f, -e2x)
#undef VXL_DOT
  // axes (-e.y, e.x, 0): z term dropped
#define VXL_DOT(X, Y, Z) ((X) * ax + (Y) * ay)
  VXL_AXIS(-e0y, e0x, 0.0f)
  VXL_AXIS(-e1y, e1x, 0.0f)
  VXL_AXIS(-e2y, e2x, 0.0f)
#undef VXL_DOT
#undef VXL_AXIS
  return true;
}

constexpr int kMaxBlock = 1024;

// One lane per (triangle, candidate voxel) pair; `chunk` = blockDim.x * ppt pairs per block.
__global__ void k_sat(const TriRec *__restrict__ rec, const uint64_t *__restrict__ coff,
                      const uint32_t *__restrict__ bstart, uint64_t Tc, uint64_t P, int ppt,
                      unsigned long long *__restrict__ leafgrid,
                      unsigned long long *__restrict__ brickbits) {
  extern __shared__ uint32_t s_off[];  // window of chunk-relative triangle offsets
  __shared__ uint32_t s_meta[2];
  const uint64_t chunk = uint64_t(blockDim.x) * ppt;
  const uint64_t base = uint64_t(blockIdx.x) * chunk;
  if (threadIdx.x == 0) {
    uint32_t lo = bstart[blockIdx.x];
    uint64_t nb = (P + chunk - 1) / chunk;
    uint32_t hi = (blockIdx.x + 1 < nb) ? bstart[blockIdx.x + 1] : uint32_t(Tc - 1);
    s_meta[0] = lo;
    s_meta[1] = hi - lo + 1;
  }
  __syncthreads();
  const uint32_t t_lo = s_meta[0], win = s_meta[1];
  for (uint32_t i = threadIdx.x; i < win; i += blockDim.x) {
    uint64_t o = coff[t_lo + i];
    s_off[i] = o <= base ? 0u : uint32_t(o - base);
  }
  __syncthreads();

  for (int it = 0; it < ppt; it++) {
    const uint32_t rel = uint32_t(it) * blockDim.x + threadIdx.x;
    const uint64_t pidx = base + rel;
    if (pidx >= P) break;
    // largest w with s_off[w] <= rel
    uint32_t lo = 0, hi = win - 1;
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (s_off[mid] <= rel) lo = mid; else hi = mid - 1;
    }
    const uint32_t t = t_lo + lo;
    const TriRec r = rec[t];
    const uint64_t li = pidx - coff[t];
    uint32_t x, y, z;
    if (__builtin_expect(r.flags & 2u, 0)) {
      uint64_t zz = li / r.exy, rem = li - zz * r.exy;
      z = uint32_t(zz);
      y = uint32_t(rem / r.ex);
      x = uint32_t(rem - uint64_t(y) * r.ex);
    } else {
      uint32_t l32 = uint32_t(li);
      z = l32 / r.exy;
      uint32_t rem = l32 - z * r.exy;
      y = rem / r.ex;
      x = rem - y * r.ex;
    }
    x += uint32_t(r.mn[0]);
    y += uint32_t(r.mn[1]);
    z += uint32_t(r.mn[2]);
    const float cx = float(x) + 0.5f, cy = float(y) + 0.5f, cz = float(z) + 0.5f;
    const bool hit = (r.flags & 1u) ? sat_overlap<true>(cx, cy, cz, r.g)
                                    : sat_overlap<false>(cx, cy, cz, r.g);
    if (hit) {
      const uint32_t code = dev::brick_code(x >> 3, y >> 3, z >> 3);
      const uint32_t m = dev::morton9(x, y, z);
      unsigned long long old =
          atomicOr(&leafgrid[uint64_t(code) * 8 + (m >> 6)], 1ull << (m & 63));
      if (old == 0ull) atomicOr(&brickbits[code >> 6], 1ull << (code & 63));
    }
  }
}

}  // namespace

uint64_t run_sat(vxl_ctx *ctx, const float *d_tris, uint64_t T, const VoxelizeParams &p) {
  hipStream_t s = ctx->stream;
  const int tb = 256;
  ctx->tri_count.reserve(T * 8);
  ctx->tri_flag.reserve(T * 4);
  ctx->tri_cidx.reserve(T * 4 + 8);
  uint64_t *count = ctx->tri_count.as<uint64_t>();
  uint32_t *flag = ctx->tri_flag.as<uint32_t>();
  uint32_t *cidx = ctx->tri_cidx.as<uint32_t>();
  uint64_t *host = static_cast<uint64_t *>(ctx->host_pinned);
  uint32_t *total_dev = cidx + T;  // one extra slot (8-byte aligned region)

  hipLaunchKernelGGL(k_tri_setup, dim3((T + tb - 1) / tb), dim3(tb), 0, s, d_tris, T, p, count,
                     flag);
  scan_exclusive_u32(flag, cidx, T, total_dev, ctx->scan_tmp, s);
  VXL_HIP(hipMemcpyAsync(host, total_dev, 4, hipMemcpyDeviceToHost, s));
  VXL_HIP(hipStreamSynchronize(s));
  const uint64_t Tc = static_cast<uint32_t *>(ctx->host_pinned)[0];
  if (Tc == 0) return 0;

  ctx->rec.reserve(Tc * sizeof(TriRec));
  ctx->ccount.reserve(Tc * 8);
  ctx->coff.reserve(Tc * 8 + 8);
  TriRec *rec = ctx->rec.as<TriRec>();
  uint64_t *ccount = ctx->ccount.as<uint64_t>();
  uint64_t *coff = ctx->coff.as<uint64_t>();
  hipLaunchKernelGGL(k_compact, dim3((T + tb - 1) / tb), dim3(tb), 0, s, d_tris, T, p, count,
                     flag, cidx, rec, ccount);
  scan_exclusive_u64(ccount, coff, Tc, coff + Tc, ctx->scan_tmp, s);
  VXL_HIP(hipMemcpyAsync(host, coff + Tc, 8, hipMemcpyDeviceToHost, s));
  VXL_HIP(hipStreamSynchronize(s));
  const uint64_t P = host[0];
  if (P == 0) return 0;

  const int block = ctx->sat_block;
  const int ppt = ctx->sat_pairs_per_thread;
  const uint64_t chunk = uint64_t(block) * ppt;
  const uint64_t nchunks = (P + chunk - 1) / chunk;
  if (nchunks > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "voxelize: pair chunks exceed the 32-bit grid limit");
  ctx->bstart.reserve(nchunks * 4);
  uint32_t *bstart = ctx->bstart.as<uint32_t>();
  hipLaunchKernelGGL(k_bstart, dim3((Tc + tb - 1) / tb), dim3(tb), 0, s, coff, ccount, Tc, chunk,
                     nchunks, bstart);
  const size_t lds = (chunk + 1) * sizeof(uint32_t);
  mark(ctx, 1);
  hipLaunchKernelGGL(k_sat, dim3(uint32_t(nchunks)), dim3(block), lds, s, rec, coff, bstart, Tc,
                     P, ppt, ctx->leafgrid.as<unsigned long long>(),
                     ctx->bitmap.as<unsigned long long>());
  mark(ctx, 2);
  VXL_HIP(hipGetLastError());
  return P;
}

}  // namespace vxl
