// Internal declarations shared by the libvxl translation units (not installed).
//
// Layout in HBM (one tree):
//   nodes   : School-B node buffer, 12 B/node {mask_lo, mask_hi, child_base}
//             (voxel-core/src/node.rs:24-31, school_b.rs:29-64)
//   leaves  : Morton-sorted 8^3 bricks, 8 x u64 per leaf; word w = 4^3 sub-brick w
//             (voxel-core/src/leaf.rs:18-23)
//   bounds  : one packed LeafBounds u32 per leaf (leaf.rs:57-92)
//   codes   : brick Morton code per leaf (sparse.rs:38), kept for edits/batching
//
// Scratch owned by a context (grow-only, reused across calls):
//   leafgrid: dense (n/8)^3 x 8 u64 words addressed by brick Morton code; it is
//             ALL ZERO between calls -- the leaf gather clears what it reads.
//   bitmap  : brick-occupancy bitmap (1 bit per brick code) -- also all zero
//             between calls -- plus the per-level bitmaps of the node pyramid.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/vxl.h"

namespace vxl {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);

struct Status {
  vxl_status code;
  std::string msg;
};

#define VXL_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      throw ::vxl::Error(e_ == hipErrorOutOfMemory ? VXL_OOM : VXL_DEVICE_ERROR,     \
                         std::string(#expr) + ": " + hipGetErrorString(e_));           \
    }                                                                                  \
  } while (0)

struct Error {
  vxl_status code;
  std::string msg;
  Error(vxl_status c, std::string m) : code(c), msg(std::move(m)) {}
};

// ---------------------------------------------------------------------------
// device memory
// ---------------------------------------------------------------------------
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t want);  // grow-only, contents not preserved
  void release();
  template <class T>
  T *as() const { return static_cast<T *>(ptr); }
};

struct vxl_ctx_impl;

}  // namespace vxl

// Opaque handle types of the C ABI.
struct vxl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serialises calls on one context
  int cu_count = 0;
  // voxelize scratch
  vxl::DevBuf tris_in, tri_count, tri_flag, tri_cidx, rec, ccount, coff, bstart;
  vxl::DevBuf leafgrid;  // zero between calls
  uint32_t leafgrid_bricks = 0;  // (n/8)^3 the grid is sized for
  bool leafgrid_dirty = false;
  vxl::DevBuf bitmap;    // level bitmaps; level-1 part is zero between calls
  vxl::DevBuf levelscratch;
  vxl::DevBuf scan_tmp;
  vxl::DevBuf words_io;  // staging for dense words
  vxl::DevBuf rays_io, hits_io, t_io;
  void *host_pinned = nullptr;  // small readback area
  // tuning
  int sat_waves_per_cu = 28;  // resident SAT waves per CU (persistent, work-queue fed)
  int sat_units_per_wave = 1;  // work units per resident wave
  int dda_block = 256;
  // last-call timing (ms), filled when timing is enabled
  bool timing = false;
  hipEvent_t ev[8] = {};  // ev[0] call start, ev[1]/ev[2] around the SAT kernel, ev[3] end
  float last_ms[8] = {0};
  std::set<vxl_tree *> trees;  // live trees created on this context
};

struct vxl_tree {
  vxl_ctx *ctx = nullptr;
  uint32_t n = 0, k = 0;
  uint64_t node_count = 0, leaf_count = 0;
  uint32_t *nodes = nullptr;   // 3 x u32 per node (>= 1 entry allocated)
  uint64_t *leaves = nullptr;  // 8 x u64 per leaf (>= 1 entry allocated)
  uint32_t *bounds = nullptr;  // 1 x u32 per leaf (>= 1 entry allocated)
  uint64_t *codes = nullptr;   // 1 x u64 per leaf (>= 1 entry allocated)
  uint64_t level_nodes[16] = {0};
  size_t bytes = 0;
  bool pooled = false;  // one stream-ordered allocation starting at `nodes`
};

namespace vxl {

// Records timing event i on the context stream when timing is enabled.
inline void mark(vxl_ctx *ctx, int i) {
  if (ctx->timing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

// ---------------------------------------------------------------------------
// host-side helpers
// ---------------------------------------------------------------------------
bool resolution_k(uint32_t n, uint32_t *k);  // n = 8*4^k ?
bool is_device_ptr(const void *p);

// scan.hip: exclusive scans on `stream`; total (if non-null) is written to device memory.
void scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t count, uint64_t *total_dev,
                        DevBuf &tmp, hipStream_t s);
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t count, uint32_t *total_dev,
                        DevBuf &tmp, hipStream_t s);

// voxelize.hip
struct VoxelizeParams {
  uint32_t n;
  float inv;       // 1 / voxel_size
  float tx[3];     // inv * (-origin)
  float eps;
};
// Runs setup + SAT into ctx->leafgrid / ctx->bitmap (level-1). Returns pair count.
uint64_t run_sat(vxl_ctx *ctx, const float *d_tris, uint64_t T, const VoxelizeParams &p);

// tree.hip
// Builds a School-B tree from the brick bitmap + leafgrid in ctx (clearing both).
vxl_tree *build_tree_from_grid(vxl_ctx *ctx, uint32_t n, uint32_t k);
// Converts the leafgrid to x-major dense words (n^3/32 u32) and clears grid + bitmap.
void grid_to_words(vxl_ctx *ctx, uint32_t n, uint32_t *d_words);
// Loads x-major words into the leafgrid + brick bitmap.
void words_to_grid(vxl_ctx *ctx, uint32_t n, const uint32_t *d_words);
// Ensures leafgrid/bitmap capacity for n; clears them if dirty.
void prepare_grid(vxl_ctx *ctx, uint32_t n);
size_t bitmap_level1_words(uint32_t n);

// dda.hip
void launch_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                     uint32_t *d_hits, double *d_t, hipStream_t s);

void tree_free_device(vxl_tree *t);

}  // namespace vxl

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
namespace vxl {
namespace dev {

__host__ __device__ __forceinline__ uint32_t spread3_10(uint32_t v) {
  // spreads the low 10 bits of v so bit i lands at bit 3i (brick codes, n <= 8192)
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__host__ __device__ __forceinline__ uint32_t compact3_10(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030c30c3u;
  v = (v | (v >> 4)) & 0x0300f00fu;
  v = (v | (v >> 8)) & 0x030000ffu;
  v = (v | (v >> 16)) & 0x3ffu;
  return v;
}
// Morton code of a brick (bx,by,bz < 1024), x in the low bit of each triple.
__host__ __device__ __forceinline__ uint32_t brick_code(uint32_t bx, uint32_t by, uint32_t bz) {
  return spread3_10(bx) | (spread3_10(by) << 1) | (spread3_10(bz) << 2);
}
// 9-bit intra-brick Morton index (voxel-core/src/morton.rs:86-89).
__host__ __device__ __forceinline__ uint32_t morton9(uint32_t x, uint32_t y, uint32_t z) {
  return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) |
         ((z & 2) << 4) | ((x & 4) << 4) | ((y & 4) << 5) | ((z & 4) << 6);
}
// 6-bit child index of a 4^3 child (voxel-core/src/node.rs:78-83).
__host__ __device__ __forceinline__ uint32_t child_bit(uint32_t x, uint32_t y, uint32_t z) {
  return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) |
         ((z & 2) << 4);
}

}  // namespace dev
}  // namespace vxl
