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
#include <utility>
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

namespace vxl {
// One device allocation shared by the trees of one batch group; freed with the last tree.
struct SharedBlock {
  void *base = nullptr;
  uint32_t refs = 0;
  int slab = -1;  // arena slab holding base
};
// A device slab of the tree arena (tree.hip).
struct Slab {
  char *base = nullptr;
  size_t bytes = 0, used = 0;
  uint32_t live = 0;  // trees (or batch groups) allocated in it and not yet freed
};
}  // namespace vxl

// Opaque handle types of the C ABI.
struct vxl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;  // serialises calls on one context
  int cu_count = 0;
  // voxelize scratch
  vxl::DevBuf tris_in, tri_count, tri_flag, tri_cidx, rec, ccount, coff, bstart;
  vxl::DevBuf setup_sums;  // per-setup-block candidate sums, scanned in place
  vxl::DevBuf leafgrid;  // zero between calls
  uint32_t leafgrid_bricks = 0;  // (n/8)^3 the grid is sized for
  bool leafgrid_dirty = false;
  vxl::DevBuf bitmap;    // level bitmaps; level-1 part is zero between calls
  vxl::DevBuf levelscratch;
  vxl::DevBuf scan_tmp;
  vxl::DevBuf words_io;  // staging for dense words
  vxl::DevBuf rays_io, hits_io, t_io;
  void *host_pinned = nullptr;      // small readback area (mapped: kernels write it)
  void *host_pinned_dev = nullptr;  // its device address
  std::vector<vxl::Slab> slabs;     // tree arena
  // tuning
  uint32_t sat_epoch = 0;  // setup + SAT calls enqueued (numbers the plan's parity slots)
  int sat_waves_per_cu = 32;  // resident SAT waves per CU (8 per SIMD; persistent, work-queue fed)
  int sat_units_per_wave = 1;  // work units per resident wave
  int sat_merge_words = 64;  // leaf-grid words a SAT wave merges in LDS before it writes them
                             // out (1..64: the 128-slot table takes 64 more per insert)
  int dda_refill = -1;  // > 0: persistent traversal lanes, refilled when that many are
                        // idle; 0: one lane per ray; -1: chosen per tree by its size
  int dda_block = 64;  // one-wave workgroups: a finished wave frees its slot at once
  int dda_batch_refill = 0;  // batched traversals: 0 one lane per ray, > 0 persistent lanes
                             // refilled once that many of a wave's lanes are idle
  uint64_t batch_pl_min_rays = 0;  // 4 x the rays the persistent batch kernel holds at once
  uint64_t pl_min_rays = 0;        // 4 x the rays the persistent ray kernel holds at once
  // last-call timing (ms), filled when timing is enabled
  bool timing = false;
  hipEvent_t ev[8] = {};  // ev[0] call start, ev[1]/ev[2] around the SAT kernel, ev[3] end
  float last_ms[8] = {0};
  std::set<vxl_tree *> trees;  // live trees created on this context
  // batch scratch: pooled leaf grids / brick bitmaps of a group of meshes (zero between
  // calls), per-group descriptors, pinned staging for their upload
  vxl::DevBuf pool_grid, pool_b1, batch_scratch, batch_l2, batch_desc, trav_desc;
  vxl::DevBuf trav_ctr;         // work counters of persistent traversals (a ring of slots)
  vxl::DevBuf owner_io, color_io, owner_scratch;  // per-voxel owners / colours (n^3 u32)
  vxl::DevBuf leaf_map;          // brick code -> leaf index of the tree being given materials
  vxl::DevBuf compact_io;        // compact voxels staged for the host
  vxl::DevBuf edit_io, edit_scratch;  // voxel edits and their per-edit results
  uint32_t trav_ctr_next = 0;
  bool pool_dirty = false;
  uint64_t pool_budget = 16ull << 30;  // leaf-grid bytes one batch group may use
  void *batch_host = nullptr, *batch_host_dev = nullptr;  // mapped: kernels may write it
  size_t batch_host_bytes = 0, batch_host_used = 0;
  hipEvent_t staging_free = nullptr;  // recorded after the last upload from batch_host
  hipEvent_t trav_done = nullptr;     // recorded after the last batched traversal kernel
  hipEvent_t ev_wait = nullptr;       // wait_stream's completion marker
  uint32_t size_seq = 0;              // single-tree builds so far (their size read-back tag)
  // host triangle uploads: two pinned chunk buffers, each free once its event has completed
  void *up_host[2] = {nullptr, nullptr};
  hipEvent_t up_ev[2] = {nullptr, nullptr};
  bool up_pending[2] = {false, false};
};


struct vxl_tree {
  vxl_ctx *ctx = nullptr;
  uint32_t n = 0, k = 0;
  uint64_t node_count = 0, leaf_count = 0;
  uint32_t *nodes = nullptr;   // 3 x u32 per node (>= 1 entry allocated)
  uint64_t *leaves = nullptr;  // 8 x u64 per leaf (>= 1 entry allocated)
  uint32_t *bounds = nullptr;  // 1 x u32 per leaf (>= 1 entry allocated)
  uint64_t *codes = nullptr;   // 1 x u64 per leaf (>= 1 entry allocated)
  uint16_t *materials = nullptr;  // 512 x u16 per leaf (intra-brick Morton order) or none
  uint64_t level_nodes[16] = {0};
  size_t bytes = 0;
  int slab = -1;  // arena slab holding the four arrays (-1: separate device allocations)
  vxl::SharedBlock *shared = nullptr;  // set for trees of a batch: arrays live in `shared`
};

namespace vxl {
struct GraphGroup;
struct TraverseSeg;
}  // namespace vxl

// A whole batch (setup, SAT, tree build, traversal of every mesh) captured in one HIP graph.
// The graph runs on a private context whose scratch nothing else touches, so replays see
// the buffers the capture recorded; the trees go to an arena sized at creation.
struct vxl_batch_graph {
  vxl_ctx *ctx = nullptr;  // the caller's context: default launch stream, result trees
  vxl_ctx *own = nullptr;  // private context (its stream, scratch and pooled grid)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t count = 0;                    // meshes
  std::vector<vxl::GraphGroup> *groups = nullptr;
  std::vector<uint32_t> n, k;            // per mesh grid
  void *desc = nullptr;                  // device descriptors of every group + segments
  void *arena = nullptr;                 // the trees
  size_t arena_bytes = 0, desc_bytes = 0;
  vxl::TraverseSeg *d_segs = nullptr;
  hipEvent_t done = nullptr;             // recorded after each launch
  bool launched = false;
  size_t kernel_nodes = 0;
};

namespace vxl {

// A tree the device path may read: non-null and still attached to a live context.
// vxl_ctx_destroy releases the device memory of the context's trees and detaches them
// (ctx = nullptr, counts zeroed); every entry point that reads a tree refuses such a
// handle on the host instead of launching a kernel over freed memory.
inline void check_tree(const vxl_tree *t) {
  if (!t) throw Error(VXL_INVALID_ARG, "tree is null");
  if (!t->ctx)
    throw Error(VXL_INVALID_ARG, "tree is detached: its context (or server) was destroyed");
}

// Records timing event i on the context stream when timing is enabled.
inline void mark(vxl_ctx *ctx, int i) {
  if (ctx->timing) (void)hipEventRecord(ctx->ev[i], ctx->stream);
}

// ---------------------------------------------------------------------------
// host-side helpers
// ---------------------------------------------------------------------------
bool resolution_k(uint32_t n, uint32_t *k);  // n = 8*4^k ?
bool is_device_ptr(const void *p);

// Waits until every operation enqueued on s so far has completed, polling an event
// rather than sleeping in hipStreamSynchronize: the sizing read-backs between the
// voxelize stages sit on the critical path of every call.
void wait_stream(vxl_ctx *ctx, hipStream_t s);

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
  uint32_t has_mat;  // world_to_grid override: g_a = ((m[0][a] px + m[1][a] py) + m[2][a] pz) + m[3][a]
  float m[12];       // columns x, y, z, w of the override, rows 0..2 (m[3 * col + row])
};

// One mesh of a batch group, as the device sees it. Mesh g owns the leaf-grid words
// [b1_begin * 512, (b1_begin + w1r) * 512) of the pooled grid and brick-bitmap words
// [b1_begin, b1_begin + w1r), with w1r = w1 rounded up to 64 so that every upper bitmap
// word belongs to one mesh; its records occupy slots [slot_begin, slot_end).
struct BatchMesh {
  const float *tris;    // the mesh's triangles (T x 9 f32, device)
  uint64_t T;           // triangles
  uint64_t b1_begin;    // first brick-bitmap word of the mesh's region
  uint64_t w1;          // brick-bitmap words in use: ceil((n/8)^3 / 64)
  uint64_t slot_begin, slot_end;
  uint32_t n, k;
  VoxelizeParams p;
};

// Per-level arrays of one tree's upper levels (levels 2..5, n <= 2048).
constexpr int kLevels = 6;
struct LevelPtrs {
  uint64_t win[kLevels];
  uint32_t *cb[kLevels], *ni[kLevels], *desc[kLevels], *pos[kLevels], *bs[kLevels];
  unsigned long long *a_mask[kLevels];
  uint32_t *a_base[kLevels], *a_code[kLevels];
  uint32_t *tot;  // per level L: tot[2L] = children, tot[2L+1] = nodes
};

// Per-mesh sizes of a batch group, read back once before the trees are allocated.
struct BatchTotals {
  uint64_t leaves, n2, leaf_begin, n2_begin, pairs, nodes[3];  // nodes of levels 3, 4, 5
};

// Pinned staging a batch group of G meshes uses (descriptors, block starts, level
// pointers, sizes, node-array pointers), each part 256-byte aligned.
inline size_t batch_staging_bytes(uint32_t G) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  return up(G * sizeof(BatchMesh)) + up((G + 1) * 4) + up(G * 8) + up(G * sizeof(LevelPtrs)) +
         up(G * sizeof(BatchTotals)) + up(G * 8) + 1024;
}
// The SAT launch plan the device derives from the pair total (voxelize.hip).
constexpr int kSkipSlots = 64;
constexpr int kUnitCounters = 8;
constexpr size_t kPlanBytes = 2048;  // SatPlan, then the unit start table
struct SatPlan {
  uint64_t P;       // decoded pairs (what the SAT kernel walks)
  uint64_t cpu;     // chunks of 64 pairs per work unit
  uint64_t nunits;  // work units
  uint64_t P_ref;   // the reference's candidate pairs (bounding-box cells), reported
  // cells the windows of call e skip accumulate in skipped[e & 1] (spread over
  // kSkipSlots slots by setup block, so the atomics do not meet on one address); the plan
  // of call e sums them and clears skipped[(e + 1) & 1] for the next call on the stream
  unsigned long long skipped[2][kSkipSlots];
  // units past the SAT grid's waves' own first units, handed out by kUnitCounters counters
  // 64 bytes apart (counter c serves the workgroups with index c mod kUnitCounters and
  // hands out every kUnitCounters-th unit), so the atomics do not queue on one address
  unsigned int counter[kUnitCounters][16];
};
// Enqueues setup + SAT into ctx->leafgrid / ctx->bitmap (level-1); no host round trip.
// The reference's pair count stays on the device at sat_pairs_dev(ctx) until the next call.
// owner (nullable): n^3 u32 preset to UINT32_MAX; each voxel's lowest covering triangle.
// With leaf_map (brick code -> leaf index of a tree built from the same call), owner is
// leaf_count x 512 u32 indexed leaf * 512 + intra-brick Morton index, and the pass writes
// nothing else (the leaf grid stays as it was).
void run_sat(vxl_ctx *ctx, const float *d_tris, uint64_t T, const VoxelizeParams &p,
             uint32_t *owner = nullptr, const uint32_t *leaf_map = nullptr);
// color[i] = hash_color(owner[i]) for owned voxels, 0 elsewhere (N voxels).
void owner_colors(vxl_ctx *ctx, const uint32_t *owner, uint32_t *color, uint64_t N);
// materials.hip: the tree of a mesh with per-voxel materials from a per-triangle table (two
// u16 per u32); the occupied voxels of a tree as (g + coord, material) int4 rows (returns
// the count, writes when capacity allows); a tree with materials from such rows.
vxl_tree *voxelize_tree_materials(vxl_ctx *ctx, const float *d_tris, uint64_t T, uint32_t n,
                                  uint32_t k, const VoxelizeParams &p, const uint32_t *d_table,
                                  uint32_t table_len, uint64_t *pairs);
uint64_t compact_tree(vxl_ctx *ctx, const vxl_tree *t, const int32_t g[3], int32_t *d_out,
                      uint64_t capacity);
vxl_tree *tree_from_compact(vxl_ctx *ctx, uint32_t n, uint32_t k, const int32_t *d_vox,
                            uint64_t count, uint64_t *dropped);
// The owner pass of the call run_sat just enqueued (same triangles): k_sat in leaf-owner
// mode over the records, offsets and plan that call left, with the work counter re-armed.
void sat_owner_pass(vxl_ctx *ctx, uint64_t T, uint32_t *owner, const uint32_t *leaf_map,
                    uint32_t own_n);
const uint64_t *sat_pairs_dev(vxl_ctx *ctx);
// Setup + SAT of a batch group into the pooled grid; mesh g owns setup blocks
// [sb_begin[g], sb_begin[g+1]) (setup_slots_per_block() record slots each).
// d_ref_pairs: G zeroed counters that receive the candidates each mesh's windows skip.
void run_sat_batch(vxl_ctx *ctx, const BatchMesh *d_mesh, const uint32_t *d_sb_begin,
                   uint32_t G, uint32_t setup_blocks, unsigned long long *leafgrid,
                   unsigned long long *brickbits, unsigned long long *d_ref_pairs);
uint64_t setup_slots_per_block();

// tree.hip
// Builds a School-B tree from the brick bitmap + leafgrid in ctx (clearing both). When
// pairs is non-null it receives *pairs_dev (0 if pairs_dev is null), read with the sizes.
vxl_tree *build_tree_from_grid(vxl_ctx *ctx, uint32_t n, uint32_t k,
                               const uint64_t *pairs_dev = nullptr, uint64_t *pairs = nullptr);
// Tree arena (tree.hip)
void *arena_alloc(vxl_ctx *ctx, size_t bytes, int *slab_out);
void arena_release(vxl_ctx *ctx, int slab);
void arena_destroy(vxl_ctx *ctx);
// Converts the leafgrid to x-major dense words (n^3/32 u32) and clears grid + bitmap.
void grid_to_words(vxl_ctx *ctx, uint32_t n, uint32_t *d_words);
// Loads x-major words into the leafgrid + brick bitmap.
void words_to_grid(vxl_ctx *ctx, uint32_t n, const uint32_t *d_words);
// Ensures leafgrid/bitmap capacity for n; clears them if dirty.
void prepare_grid(vxl_ctx *ctx, uint32_t n);
size_t bitmap_level1_words(uint32_t n);
// Trees of a batch group from the pooled grid (cleared again on the way). meshes[g] as
// uploaded to d_mesh; W1 = brick-bitmap words of the group. Writes trees[g] and, when
// non-null, pairs[g]. One host sync.
void build_trees_batch(vxl_ctx *ctx, const std::vector<BatchMesh> &meshes,
                       const BatchMesh *d_mesh, uint64_t W1, vxl_tree **trees, uint64_t *pairs,
                       const unsigned long long *d_ref_pairs);

// dda.hip
// counter (nullable): the persistent schedule's work counter; null takes a slot of the
// context's ring (allocated on first use).
void launch_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                     uint32_t *d_hits, double *d_t, hipStream_t s,
                     unsigned long long *counter = nullptr);
// dda.hip: the allocation-free traversal's counter slots (a per-device static array): its
// address found once per device at context creation (hipGetSymbolAddress), one slot per call.
void noalloc_setup(int device);
unsigned long long *noalloc_counter(int device);
// One segment of a batched traversal: R rays at `rays` (6 f32 each) against `tree`,
// results to hits (4 u32 each) and t (nullable).
struct RaySegment {
  const vxl_tree *tree;
  const float *rays;
  uint64_t R;
  uint32_t *hits;
  double *t;
};
// All segments in one kernel launch on stream s.
void launch_traverse_batch(vxl_ctx *ctx, const RaySegment *segs, uint32_t S, hipStream_t s);
// One segment of a batched traversal, as the kernel sees it.
struct TraverseSeg {
  const uint32_t *nodes, *leafw, *bounds;
  const float *rays;
  uint4 *hits;
  double *t;
  uint64_t leaf_count, R;
  uint32_t n, k;
};
// The batched traversal over descriptors already in device memory, in the context's batch
// lane schedule: one lane per ray (segment g owns blocks [d_blk[g], d_blk[g+1]) of `blocks`
// workgroups of ctx->dda_block lanes), or persistent lanes over the run table d_runs
// (segment g owns runs [d_runs[g], d_runs[g+1]) of 256 rays; `counter` is cleared by a
// kernel first, and t comes from a second pass when want_t). kmax is the largest tree level
// count k of any segment. Kernel launches only (graph capture uses it).
void launch_traverse_segs(vxl_ctx *ctx, const TraverseSeg *d_segs, const uint32_t *d_blk,
                          const uint64_t *d_runs, uint32_t S, uint64_t blocks, uint64_t runs,
                          uint32_t kmax, bool want_t, unsigned long long *counter,
                          hipStream_t s);
// Rays that fill the device's resident persistent batch waves four times over (the
// server's threshold for its persistent-lane schedule).
uint64_t persistent_batch_min_rays(vxl_ctx *ctx);
// Rays per run of the persistent batched traversal's run table.
constexpr uint64_t kBatchRayRun = 256;
// tree.hip: SparseTree::set_voxel over distinct in-bounds edits on the device; patches
// leaves in place or, when a brick appears or empties, rebuilds the tree's arrays (the
// handle stays the same).
struct EditResult {
  uint64_t changed;
  bool topology;
};
EditResult edit_tree(vxl_ctx *ctx, vxl_tree *t, const vxl_voxel_edit *d_edits, uint64_t E);
// noise.hip: the NoiseField fixture's tree, built on the device.
vxl_tree *generate_noise_tree(vxl_ctx *ctx, uint32_t n, uint32_t k, const vxl_noise &f);
// dda.hip: one lane per ray, nothing but the kernel launch (graph capture uses it).
void launch_traverse_plain(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                           uint32_t *d_hits, double *d_t, hipStream_t s);
// dda.hip: the walk statistics of a ray batch {leaf descents, node descents, visits}.
void traverse_stats(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                    uint64_t out[3]);
// dda.hip: one RGBA8 image of the tree (render.wgsl); d_colors: the global colour table.
void launch_render(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera &cam,
                   const uint32_t *d_colors, uint32_t ncolors, uint32_t *d_rgba, hipStream_t s);

// capi.hip: argument checks of a voxelize call (grid, triangle count, epsilon; returns k)
// and the world-to-grid parameters.
uint32_t check_voxelize_args(const vxl_grid *grid, uint64_t T, float eps);
VoxelizeParams voxelize_params(const vxl_grid *grid, float eps);
// tree.hip: an empty tree (one padded entry per array, like upload_structure's padding)
vxl_tree *empty_tree_device(vxl_ctx *ctx, uint32_t n, uint32_t k);
// tree.hip: a tree of ctx holding copies (enqueued on s) of device arrays of another
// (leaves = 0: an empty tree); level_nodes: 16 per-level node counts.
vxl_tree *tree_copy_arrays(vxl_ctx *ctx, uint32_t n, uint32_t k, uint64_t leaves,
                           uint64_t nodes, const uint64_t *level_nodes, const void *src_nodes,
                           const void *src_leaves, const void *src_bounds,
                           const void *src_codes, hipStream_t s);

// batch.hip
// Pinned host staging for descriptor uploads: staging_begin waits until earlier uploads
// from it are done and hands out `bytes` (more via staging_take); staging_end records
// the point after which it may be reused.
void *staging_begin(vxl_ctx *ctx, size_t bytes);
void *staging_take(vxl_ctx *ctx, size_t bytes);
void staging_end(vxl_ctx *ctx, hipStream_t s);
void *staging_dev(vxl_ctx *ctx, void *host);  // device address of a staging pointer
// One mesh handed to voxelize_batch: out[index] / pairs[index] receive its results.
struct BatchInput {
  uint32_t index;
  const float *tris;  // device
  uint64_t T;
  uint32_t n, k;  // 2 <= k <= 4
  VoxelizeParams p;
};
void voxelize_batch(vxl_ctx *ctx, const std::vector<BatchInput> &in, vxl_tree **out,
                    uint64_t *pairs);
// The groups voxelize_batch splits `in` into: [first, end) runs whose pooled leaf grids fit
// the context's batch budget.
std::vector<std::pair<size_t, size_t>> batch_groups(const vxl_ctx *ctx,
                                                    const std::vector<BatchInput> &in);
// The group descriptors (BatchMesh, setup-block starts) of in[first, end), as run_group
// uploads them; returns the group's setup blocks and brick-bitmap words.
void batch_group_meshes(const BatchInput *in, uint32_t G, std::vector<BatchMesh> &meshes,
                        std::vector<uint32_t> &sb_begin, uint64_t *blocks, uint64_t *W1);

// A batch group of a graph (batch.hip builds it, tree.hip enqueues its tree build). Device
// descriptors are the graph's own and stay put between replays; the level arrays and the
// group-wide scratch come from the graph's private context.
struct GraphState {  // written on the device by each replay, per group
  uint32_t ok;       // 1: the group's trees fit its arena region
  uint32_t pad;
  uint64_t node_bytes, leaves;  // what the group's trees needed
};
struct GraphGroup {
  uint32_t first = 0, G = 0;     // meshes [first, first + G) of the batch
  uint32_t setup_blocks = 0;
  uint64_t W1 = 0;               // brick-bitmap words of the group
  std::vector<BatchMesh> meshes;
  // device, graph-owned
  BatchMesh *d_mesh = nullptr;
  uint32_t *d_sb = nullptr;
  unsigned long long *d_ref = nullptr;
  LevelPtrs *d_lps = nullptr;
  BatchTotals *d_tot = nullptr;
  uint64_t *n2_beg = nullptr;    // G + 1
  uint32_t **nodes_ptr = nullptr;
  uint64_t *node_off = nullptr;  // byte offset of mesh g's nodes in node_region
  GraphState *state = nullptr;
  // the group's arena region
  char *node_region = nullptr;
  uint64_t node_cap = 0;         // bytes
  char *leaves_base = nullptr, *bounds_base = nullptr, *codes_base = nullptr;
  uint64_t leaf_cap = 0;         // leaves
};
// batch.hip: the graph of a batch (vxl_batch_graph_create): every mesh of `in` (in[g].index
// = g, 2 <= k <= 4); rays [roff[g], roff[g+1]) of d_rays walk mesh g's tree. Launch enqueues
// a replay on s; result copies the last replay's trees (trees and pairs nullable) or throws
// VXL_STORAGE_EXCEEDED; release frees what build made (a partly built graph too).
void batch_graph_build(vxl_batch_graph *bg, const std::vector<BatchInput> &in,
                       const float *d_rays, const uint64_t *roff, uint32_t *d_hits,
                       double *d_t, double headroom);
void batch_graph_launch(vxl_batch_graph *bg, hipStream_t s);
void batch_graph_result(vxl_batch_graph *bg, vxl_tree **trees, uint64_t *pairs);
void batch_graph_release(vxl_batch_graph *bg);
// tree.hip: the group-wide scratch of a graph group's tree build (level-2 scan and
// bitmap, School-A level-2 arrays) and every mesh's upper-level arrays, carved from one
// buffer of group_scratch_bytes(meshes, W1) bytes; lps receives one LevelPtrs per mesh.
struct GroupScratch {
  uint64_t *pn2;
  unsigned long long *b2, *a_mask2;
  uint32_t *a_base2, *a_code2;
};
size_t group_scratch_bytes(const std::vector<BatchMesh> &meshes, uint64_t W1);
GroupScratch carve_group_scratch(const std::vector<BatchMesh> &meshes, uint64_t W1, char *p,
                                 LevelPtrs *lps);
// tree.hip: the group's tree build with the sizes left on the device: kernels only, trees
// placed in the group's arena region by the device, segs[first + g] given mesh g's tree.
void enqueue_trees_graph(vxl_ctx *ctx, const GraphGroup &gg, const GroupScratch &gs,
                         TraverseSeg *segs, hipStream_t s);
// voxelize.hip: a u64 fill kernel (graphs clear counters with a kernel, not a memset node).
void fill_u64(unsigned long long *p, uint64_t count, unsigned long long v, hipStream_t s);

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
// Largest g in [0, G) with begin[g] <= x, for a non-decreasing begin[] (the segment
// holding x when segment g spans [begin[g], begin[g+1])). Callers pass a wave-uniform x.
template <class T>
__device__ __forceinline__ uint32_t upper_index(const T *__restrict__ begin, uint32_t G, T x) {
  uint32_t lo = 0, hi = G - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (begin[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
// 6-bit child index of a 4^3 child (voxel-core/src/node.rs:78-83).
__host__ __device__ __forceinline__ uint32_t child_bit(uint32_t x, uint32_t y, uint32_t z) {
  return (x & 1) | ((y & 1) << 1) | ((z & 1) << 2) | ((x & 2) << 2) | ((y & 2) << 3) |
         ((z & 2) << 4);
}

}  // namespace dev
}  // namespace vxl
