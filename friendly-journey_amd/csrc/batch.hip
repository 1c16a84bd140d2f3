// Many meshes per launch: the batch group driver and the pinned staging it uploads from.
//
// A group is a run of meshes whose dense leaf grids fit the context's batch budget. Each
// mesh gets a region of the pooled leaf grid (its (n/8)^3 bricks, rounded up to a whole
// upper-bitmap word of 64 x 64 bricks) and the matching words of the pooled brick bitmap.
// Per group: one setup launch over every triangle, one scan, one persistent SAT launch
// over every (triangle, voxel) pair, one batched tree build -- one host round trip (the
// per-mesh tree sizes), whatever the number of meshes. The pooled grid and bitmap are all-zero between groups
// (the leaf gather clears what it reads), like the single-mesh scratch.
#include "vxl_internal.hpp"

namespace vxl {

void *staging_begin(vxl_ctx *ctx, size_t bytes) {
  if (ctx->staging_free) VXL_HIP(hipEventSynchronize(ctx->staging_free));
  if (ctx->batch_host_bytes < bytes) {
    if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
    ctx->batch_host = nullptr;
    ctx->batch_host_bytes = 0;
    const size_t b = bytes < (64u << 10) ? (64u << 10) : (bytes + 4095) / 4096 * 4096;
    VXL_HIP(hipHostMalloc(&ctx->batch_host, b, hipHostMallocDefault));
    ctx->batch_host_bytes = b;
    VXL_HIP(hipHostGetDevicePointer(&ctx->batch_host_dev, ctx->batch_host, 0));
  }
  ctx->batch_host_used = 0;
  return ctx->batch_host;
}

void *staging_take(vxl_ctx *ctx, size_t bytes) {
  const size_t b = (bytes + 255) / 256 * 256;
  if (ctx->batch_host_used + b > ctx->batch_host_bytes)
    throw Error(VXL_DEVICE_ERROR, "internal: batch staging overflow");
  void *p = static_cast<char *>(ctx->batch_host) + ctx->batch_host_used;
  ctx->batch_host_used += b;
  return p;
}

void *staging_dev(vxl_ctx *ctx, void *host) {
  return static_cast<char *>(ctx->batch_host_dev) +
         (static_cast<char *>(host) - static_cast<char *>(ctx->batch_host));
}

void staging_end(vxl_ctx *ctx, hipStream_t s) {
  if (!ctx->staging_free)
    VXL_HIP(hipEventCreateWithFlags(&ctx->staging_free, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ctx->staging_free, s));
}

namespace {

uint64_t region_words(uint32_t n) {  // brick-bitmap words of a mesh region
  return (bitmap_level1_words(n) + 63) / 64 * 64;
}

void ensure_pool(vxl_ctx *ctx, uint64_t W1) {
  const size_t grid_bytes = size_t(W1) * 512 * 8;
  const size_t b1_bytes = size_t(W1) * 8 + 256;
  bool fresh = false;
  if (ctx->pool_grid.bytes < grid_bytes) {
    ctx->pool_grid.reserve(grid_bytes);
    fresh = true;
  }
  if (ctx->pool_b1.bytes < b1_bytes) {
    ctx->pool_b1.reserve(b1_bytes);
    fresh = true;
  }
  if (fresh || ctx->pool_dirty) {
    VXL_HIP(hipMemsetAsync(ctx->pool_grid.ptr, 0, ctx->pool_grid.bytes, ctx->stream));
    VXL_HIP(hipMemsetAsync(ctx->pool_b1.ptr, 0, ctx->pool_b1.bytes, ctx->stream));
    ctx->pool_dirty = false;
  }
}

void run_group(vxl_ctx *ctx, const BatchInput *in, uint32_t G, vxl_tree **out,
               uint64_t *pairs) {
  hipStream_t s = ctx->stream;
  const uint64_t per_block = setup_slots_per_block();
  staging_begin(ctx, batch_staging_bytes(G));
  BatchMesh *h_mesh = static_cast<BatchMesh *>(staging_take(ctx, G * sizeof(BatchMesh)));
  uint32_t *h_sb = static_cast<uint32_t *>(staging_take(ctx, (G + 1) * 4));
  uint64_t *h_ref = static_cast<uint64_t *>(staging_take(ctx, G * 8));  // zeroed counters
  std::memset(h_ref, 0, G * 8);
  std::vector<BatchMesh> meshes(G);
  uint64_t W1 = 0, blocks = 0;
  for (uint32_t g = 0; g < G; g++) {
    BatchMesh &m = meshes[g];
    m.tris = in[g].tris;
    m.T = in[g].T;
    m.b1_begin = W1;
    m.w1 = bitmap_level1_words(in[g].n);
    m.n = in[g].n;
    m.k = in[g].k;
    m.p = in[g].p;
    h_sb[g] = uint32_t(blocks);
    m.slot_begin = blocks * per_block;
    blocks += (m.T + per_block - 1) / per_block;
    m.slot_end = blocks * per_block;
    W1 += region_words(m.n);
    if (blocks > 0xffffffffull || W1 > 0xffffffffull)
      throw Error(VXL_WORKGROUPS_EXCEEDED, "voxelize batch: group exceeds the 32-bit limits");
  }
  h_sb[G] = uint32_t(blocks);
  std::memcpy(h_mesh, meshes.data(), G * sizeof(BatchMesh));
  ensure_pool(ctx, W1);
  const size_t mesh_bytes = (G * sizeof(BatchMesh) + 255) / 256 * 256;
  const size_t sb_bytes = ((G + 1) * 4 + 255) / 256 * 256;
  const size_t desc_bytes = mesh_bytes + sb_bytes + G * 8;
  ctx->batch_desc.reserve(desc_bytes + 256);
  char *d = static_cast<char *>(ctx->batch_desc.ptr);
  VXL_HIP(hipMemcpyAsync(d, h_mesh, desc_bytes, hipMemcpyHostToDevice, s));
  const BatchMesh *d_mesh = reinterpret_cast<const BatchMesh *>(d);
  unsigned long long *d_ref = reinterpret_cast<unsigned long long *>(d + mesh_bytes + sb_bytes);
  run_sat_batch(ctx, d_mesh, reinterpret_cast<const uint32_t *>(d + mesh_bytes), G,
                uint32_t(blocks), ctx->pool_grid.as<unsigned long long>(),
                ctx->pool_b1.as<unsigned long long>(), d_ref);  // no host round trip
  std::vector<vxl_tree *> trees(G, nullptr);
  std::vector<uint64_t> pr(G, 0);
  build_trees_batch(ctx, meshes, d_mesh, W1, trees.data(), pr.data(), d_ref);
  staging_end(ctx, s);
  for (uint32_t g = 0; g < G; g++) {
    out[in[g].index] = trees[g];
    ctx->trees.insert(trees[g]);
    if (pairs) pairs[in[g].index] = pr[g];
  }
}

}  // namespace

void voxelize_batch(vxl_ctx *ctx, const std::vector<BatchInput> &in, vxl_tree **out,
                    uint64_t *pairs) {
  size_t i = 0;
  while (i < in.size()) {
    uint64_t words = 0;
    size_t j = i;
    while (j < in.size()) {
      const uint64_t w = region_words(in[j].n);
      if (j > i && (words + w) * 4096 > ctx->pool_budget) break;
      words += w;
      j++;
    }
    run_group(ctx, in.data() + i, uint32_t(j - i), out, pairs);
    i = j;
  }
}

}  // namespace vxl
