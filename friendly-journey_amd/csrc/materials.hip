// Per-voxel materials: the mesh path compact_surface_sparse -> tree_from_compact
// (voxelizer/src/gpu/sparse.rs:61-111, shaders/compact_voxels.wgsl, materials.rs:153-192,
// voxel-core/src/sparse.rs:324-366), on the device.
//
// voxelize_tree_materials: the voxelize call builds the tree; a brick-code -> leaf map of it
//   lets a second SAT pass (owners only, k_sat's leaf mode) keep each leaf voxel's lowest
//   covering triangle in leaf_count x 512 slots, never an n^3 array; each occupied voxel's
//   material is then its owner's entry of the caller's table (two u16 per u32), 0 (MISSING)
//   for an owner past the table. The tree keeps the materials: 512 u16 per leaf in
//   intra-brick Morton order, as SparseTree keeps them.
// compact_tree: every occupied voxel of a tree as (g_origin + coord, material), in leaf order
//   then intra-brick Morton order (the reference's order is that of an atomic counter, so
//   any order is the same set).
// tree_from_compact: the in-range voxels OR'd into the dense leaf grid, the same tree build,
//   then per leaf voxel the material of the last list entry naming it (an atomic max of the
//   entry index), masked to 16 bits.
#include "vxl_internal.hpp"

namespace vxl {
namespace {

constexpr int kTB = 256;

// compact_voxels.wgsl:46-66: triangle t's material is bits (t & 1) * 16 of word t >> 1.
__device__ __forceinline__ uint32_t resolve_material(uint32_t tri, const uint32_t *__restrict__ table,
                                                     uint32_t len) {
  if (tri == 0xFFFFFFFFu || len == 0) return 0;
  const uint32_t w = tri >> 1;
  if (w >= len) return 0;
  return (table[w] >> ((tri & 1u) << 4)) & 0xFFFFu;
}

// (x, y, z) of bit m of an 8^3 leaf (the inverse of dev::morton9).
__device__ __forceinline__ void leaf_local(uint32_t m, uint32_t &x, uint32_t &y, uint32_t &z) {
  x = (m & 1u) | ((m >> 2) & 2u) | ((m >> 4) & 4u);
  y = ((m >> 1) & 1u) | ((m >> 3) & 2u) | ((m >> 5) & 4u);
  z = ((m >> 2) & 1u) | ((m >> 4) & 2u) | ((m >> 6) & 4u);
}

__global__ void k_leaf_map(const uint64_t *__restrict__ codes, uint64_t L,
                           uint32_t *__restrict__ map) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < L) map[codes[i]] = uint32_t(i);
}

// Owners start at UINT32_MAX (no owner) at occupied leaf voxels only: the owner pass hits
// no other voxel, and the materials pass reads no other owner. One thread per leaf word.
__global__ void k_owner_init(const uint64_t *__restrict__ leaves, uint64_t W,
                             uint32_t *__restrict__ owner) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < W;
       i += uint64_t(gridDim.x) * blockDim.x) {
    unsigned long long bits = leaves[i];
    while (bits) {
      const int b = __ffsll(bits) - 1;
      bits &= bits - 1ull;
      owner[i * 64 + uint64_t(b)] = 0xFFFFFFFFu;
    }
  }
}

// One thread per 8 consecutive leaf voxels (one 16-byte store of 8 materials): an occupied
// voxel's material is its owner's table entry, an empty voxel's 0. Owners are read for
// occupied voxels only.
__global__ void k_leaf_materials(const uint64_t *__restrict__ leaves,
                                 const uint32_t *__restrict__ owner, uint64_t N8,
                                 const uint32_t *__restrict__ table, uint32_t len,
                                 uint4 *__restrict__ mats8) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N8;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t occ = uint32_t(leaves[i >> 3] >> ((i & 7) * 8)) & 0xffu;
    uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (occ) {
#pragma unroll
      for (int k = 0; k < 8; k++)
        if ((occ >> k) & 1u) m[k] = resolve_material(owner[i * 8 + k], table, len);
    }
    mats8[i] = make_uint4(m[0] | (m[1] << 16), m[2] | (m[3] << 16), m[4] | (m[5] << 16),
                          m[6] | (m[7] << 16));
  }
}

__global__ void k_word_popc(const uint64_t *__restrict__ words, uint64_t W,
                            uint64_t *__restrict__ cnt) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < W) cnt[i] = uint64_t(__popcll(words[i]));
}

// One thread per leaf word: its set bits in order, from the word's scanned offset.
__global__ void k_compact_write(const uint64_t *__restrict__ leaves,
                                const uint64_t *__restrict__ codes,
                                const uint16_t *__restrict__ mats, const uint64_t *__restrict__ off,
                                uint64_t W, int32_t gx, int32_t gy, int32_t gz,
                                int4 *__restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= W) return;
  unsigned long long bits = leaves[i];
  if (!bits) return;
  const uint64_t leaf = i >> 3;
  const uint32_t code = uint32_t(codes[leaf]);
  const int32_t bx = int32_t(dev::compact3_10(code)) * 8, by = int32_t(dev::compact3_10(code >> 1)) * 8,
                bz = int32_t(dev::compact3_10(code >> 2)) * 8;
  uint64_t pos = off[i];
  while (bits) {
    const uint32_t b = uint32_t(__ffsll((long long)bits) - 1);
    bits &= bits - 1;
    const uint32_t m = uint32_t(i & 7) * 64 + b;
    uint32_t x, y, z;
    leaf_local(m, x, y, z);
    const uint32_t mat = mats ? uint32_t(mats[leaf * 512 + m]) : 0u;
    out[pos++] = make_int4(gx + bx + int32_t(x), gy + by + int32_t(y), gz + bz + int32_t(z),
                           int32_t(mat));
  }
}

// tree_from_compact's filter and occupancy: in-range voxels set their leaf-grid bit and
// their brick's bitmap bit; the rest are counted (one atomic per wave).
__global__ void k_compact_scatter(const int4 *__restrict__ vox, uint64_t count, uint32_t n,
                                  unsigned long long *__restrict__ leafgrid,
                                  unsigned long long *__restrict__ b1,
                                  unsigned long long *__restrict__ dropped) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool out = false;
  if (i < count) {
    const int4 v = vox[i];
    out = v.x < 0 || v.y < 0 || v.z < 0 || v.x >= int32_t(n) || v.y >= int32_t(n) ||
          v.z >= int32_t(n);
    if (!out) {
      const uint32_t x = uint32_t(v.x), y = uint32_t(v.y), z = uint32_t(v.z);
      const uint32_t code = dev::brick_code(x >> 3, y >> 3, z >> 3);
      const uint32_t m = dev::morton9(x, y, z);
      atomicOr(&leafgrid[uint64_t(code) * 8 + (m >> 6)], 1ull << (m & 63));
      atomicOr(&b1[code >> 6], 1ull << (code & 63));
    }
  }
  const unsigned long long ob = __ballot(out);
  if ((threadIdx.x & 63) == 0 && ob) atomicAdd(dropped, (unsigned long long)__popcll(ob));
}

// last[leaf * 512 + m] = 1 + the highest list index naming the voxel (0: none).
__global__ void k_compact_last(const int4 *__restrict__ vox, uint64_t count, uint32_t n,
                               const uint32_t *__restrict__ map, uint32_t *__restrict__ last) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int4 v = vox[i];
  if (v.x < 0 || v.y < 0 || v.z < 0 || v.x >= int32_t(n) || v.y >= int32_t(n) || v.z >= int32_t(n))
    return;
  const uint32_t x = uint32_t(v.x), y = uint32_t(v.y), z = uint32_t(v.z);
  const uint32_t leaf = map[dev::brick_code(x >> 3, y >> 3, z >> 3)];
  atomicMax(&last[uint64_t(leaf) * 512 + dev::morton9(x, y, z)], uint32_t(i + 1));
}

__global__ void k_mats_from_last(const uint32_t *__restrict__ last, uint64_t N,
                                 const int4 *__restrict__ vox, uint16_t *__restrict__ mats) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t l = last[i];
    mats[i] = l ? uint16_t(uint32_t(vox[l - 1].w) & 0xFFFFu) : uint16_t(0);
  }
}

// Blocks for N threads, for kernels that take one element per thread.
uint32_t grid_blocks(uint64_t N) {
  const uint64_t b = (N + kTB - 1) / kTB;
  if (b * kTB > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "materials: more than 2^32 threads in one launch");
  return uint32_t(b);
}

// Blocks for N elements of a grid-stride kernel: a launch holds at most 2^32 - 1 threads
// per dimension, so the grid is capped and the kernel strides.
uint32_t stride_blocks(uint64_t N) {
  const uint64_t b = (N + kTB - 1) / kTB;
  return uint32_t(b < (1ull << 22) ? b : (1ull << 22));
}

// The brick-code -> leaf index map of a tree, in ctx scratch ((n/8)^3 u32). Entries of
// bricks that are not leaves are left as they are: only leaf bricks are ever looked up.
uint32_t *leaf_map(vxl_ctx *ctx, const vxl_tree *t) {
  const uint64_t nb = uint64_t(t->n / 8);
  ctx->leaf_map.reserve(nb * nb * nb * 4);
  uint32_t *map = ctx->leaf_map.as<uint32_t>();
  hipLaunchKernelGGL(k_leaf_map, dim3(grid_blocks(t->leaf_count)), dim3(kTB), 0, ctx->stream,
                     t->codes, t->leaf_count, map);
  VXL_HIP(hipGetLastError());
  return map;
}

uint16_t *alloc_materials(uint64_t L) {
  void *p = nullptr;
  const hipError_t e = hipMalloc(&p, size_t(L) * 512 * 2);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(VXL_OOM, "materials: device allocation of " + std::to_string(L * 1024) +
                             " bytes failed");
  }
  return static_cast<uint16_t *>(p);
}

}  // namespace

vxl_tree *voxelize_tree_materials(vxl_ctx *ctx, const float *d_tris, uint64_t T, uint32_t n,
                                  uint32_t k, const VoxelizeParams &p, const uint32_t *d_table,
                                  uint32_t table_len, uint64_t *pairs) {
  if (T == 0) {
    if (pairs) *pairs = 0;
    return empty_tree_device(ctx, n, k);
  }
  prepare_grid(ctx, n);
  run_sat(ctx, d_tris, T, p);
  vxl_tree *t = build_tree_from_grid(ctx, n, k, sat_pairs_dev(ctx), pairs);
  if (t->leaf_count == 0) return t;
  try {
    const uint64_t N = t->leaf_count * 512;
    uint32_t *map = leaf_map(ctx, t);
    ctx->owner_scratch.reserve(N * 4);
    uint32_t *owner = ctx->owner_scratch.as<uint32_t>();
    const uint64_t W = t->leaf_count * 8;
    hipLaunchKernelGGL(k_owner_init, dim3(stride_blocks(W)), dim3(kTB), 0, ctx->stream,
                       reinterpret_cast<const uint64_t *>(t->leaves), W, owner);
    // the owner pass walks the same pairs: it reuses the records, offsets and plan above
    sat_owner_pass(ctx, T, owner, map, n);
    t->materials = alloc_materials(t->leaf_count);
    hipLaunchKernelGGL(k_leaf_materials, dim3(stride_blocks(N / 8)), dim3(kTB), 0, ctx->stream,
                       reinterpret_cast<const uint64_t *>(t->leaves), owner, N / 8, d_table,
                       table_len, reinterpret_cast<uint4 *>(t->materials));
    VXL_HIP(hipGetLastError());
  } catch (...) {
    tree_free_device(t);
    delete t;
    throw;
  }
  return t;
}

uint64_t compact_tree(vxl_ctx *ctx, const vxl_tree *t, const int32_t g[3], int32_t *d_out,
                      uint64_t capacity) {
  const uint64_t W = t->leaf_count * 8;
  if (W == 0) return 0;
  ctx->ccount.reserve(W * 8);
  ctx->coff.reserve(W * 8 + 8);
  uint64_t *cnt = ctx->ccount.as<uint64_t>(), *off = ctx->coff.as<uint64_t>();
  hipLaunchKernelGGL(k_word_popc, dim3(grid_blocks(W)), dim3(kTB), 0, ctx->stream, t->leaves, W,
                     cnt);
  scan_exclusive_u64(cnt, off, W, off + W, ctx->scan_tmp, ctx->stream);
  VXL_HIP(hipMemcpyAsync(ctx->host_pinned, off + W, 8, hipMemcpyDeviceToHost, ctx->stream));
  wait_stream(ctx, ctx->stream);
  const uint64_t total = *static_cast<uint64_t *>(ctx->host_pinned);
  if (d_out && capacity >= total) {
    hipLaunchKernelGGL(k_compact_write, dim3(grid_blocks(W)), dim3(kTB), 0, ctx->stream, t->leaves,
                       t->codes, t->materials, off, W, g[0], g[1], g[2],
                       reinterpret_cast<int4 *>(d_out));
    VXL_HIP(hipGetLastError());
  }
  return total;
}

vxl_tree *tree_from_compact(vxl_ctx *ctx, uint32_t n, uint32_t k, const int32_t *d_vox,
                            uint64_t count, uint64_t *dropped) {
  if (count > 0xFFFFFFFEull)
    throw Error(VXL_INVALID_ARG, "tree_from_compact: more than 2^32 - 2 voxels in one call");
  prepare_grid(ctx, n);
  const int4 *vox = reinterpret_cast<const int4 *>(d_vox);
  // the dropped count goes straight to the mapped readback area, past what the tree build
  // uses of it
  constexpr size_t kDropSlot = 3072;
  unsigned long long *dcount = reinterpret_cast<unsigned long long *>(
      static_cast<char *>(ctx->host_pinned_dev) + kDropSlot);
  volatile uint64_t *hdrop =
      reinterpret_cast<volatile uint64_t *>(static_cast<char *>(ctx->host_pinned) + kDropSlot);
  *hdrop = 0;
  if (count) {
    hipLaunchKernelGGL(k_compact_scatter, dim3(grid_blocks(count)), dim3(kTB), 0, ctx->stream,
                       vox, count, n, ctx->leafgrid.as<unsigned long long>(),
                       ctx->bitmap.as<unsigned long long>(), dcount);
    VXL_HIP(hipGetLastError());
  }
  vxl_tree *t = build_tree_from_grid(ctx, n, k, nullptr, nullptr);
  wait_stream(ctx, ctx->stream);
  if (dropped) *dropped = *hdrop;
  if (t->leaf_count == 0) return t;
  try {
    const uint64_t N = t->leaf_count * 512;
    uint32_t *map = leaf_map(ctx, t);
    ctx->owner_scratch.reserve(N * 4);
    uint32_t *last = ctx->owner_scratch.as<uint32_t>();
    VXL_HIP(hipMemsetAsync(last, 0, N * 4, ctx->stream));
    hipLaunchKernelGGL(k_compact_last, dim3(grid_blocks(count)), dim3(kTB), 0, ctx->stream, vox,
                       count, n, map, last);
    t->materials = alloc_materials(t->leaf_count);
    hipLaunchKernelGGL(k_mats_from_last, dim3(stride_blocks(N)), dim3(kTB), 0, ctx->stream, last, N,
                       vox, t->materials);
    VXL_HIP(hipGetLastError());
  } catch (...) {
    tree_free_device(t);
    delete t;
    throw;
  }
  return t;
}

}  // namespace vxl
