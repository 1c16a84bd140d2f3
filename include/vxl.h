/*
 * vxl.h -- C ABI of the MI355X voxelize + sparse-tree + DDA hot path (libvxl.so).
 *
 * Drop-in boundary for the reference's GPU voxelizer and GPU traverser. Each entry
 * point names the reference interface it replaces (paths relative to the reference
 * repository root). Plain pointers and sizes only; no C++ or torch types.
 *
 * Conventions (mirroring the reference boundary):
 *  - Inputs are borrowed; outputs go to caller-owned buffers or library-owned
 *    opaque handles released with an explicit *_free / *_destroy.
 *  - Entry points without a `_dev` suffix take HOST pointers and are synchronous,
 *    like the reference's blocking `voxelize_surface` / `traverse` calls.
 *    `_dev` entry points take DEVICE pointers and enqueue on the given HIP stream
 *    (NULL = the context's stream); outputs are valid once that stream is synced.
 *  - Validation happens before any device work, in the reference's order:
 *    grid -> mesh -> opts -> storage (voxelizer/src/gpu/dense.rs:30-40).
 *  - An empty mesh short-circuits to the all-empty output (gpu/dense.rs:45-47).
 *  - A context serialises its calls (internal mutex); use one context per stream
 *    of work for concurrency.
 *  - vxl_last_error() returns the calling thread's last message.
 */
#ifndef VXL_H
#define VXL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes mirror voxelizer/src/error.rs:121-177 (VoxelizeGpuError) and
 * voxel-gpu/src/error.rs:9-50 (GpuError). */
typedef enum vxl_status {
  VXL_OK = 0,
  VXL_NO_DEVICE = 1,            /* NoAdapter */
  VXL_INVALID_ARG = 2,          /* PipelineValidation (grid / mesh / opts checks) */
  VXL_STORAGE_EXCEEDED = 3,     /* StorageExceeded / BufferTooLarge */
  VXL_WORKGROUPS_EXCEEDED = 4,  /* WorkgroupsExceeded */
  VXL_OOM = 5,                  /* device allocation failed */
  VXL_BUSY = 6,                 /* serving queue full / admission refused */
  VXL_DEVICE_ERROR = 7,         /* DeviceRequest / Poll / BufferMap */
  VXL_NOT_REPRESENTABLE = 8     /* Resolution::new rejected n (resolution.rs:76-118) */
} vxl_status;

typedef struct vxl_ctx vxl_ctx;
typedef struct vxl_tree vxl_tree;

/* A cubic grid: n = 8*4^k voxels per axis, placed by origin + uniform voxel size
 * (VoxelGrid, voxelizer/src/core.rs:25-35), or by an explicit world-to-grid affine
 * (VoxelGrid::world_to_grid, core.rs:30-34,101-104): when has_world_to_grid is non-zero,
 * world_to_grid is the glam Mat4 in its column-major storage (x_axis, y_axis, z_axis,
 * w_axis, four floats each) and each vertex is mapped as Mat4::transform_point3 does, per
 * component ((x_axis*p.x + y_axis*p.y) + z_axis*p.z) + w_axis (the fourth row is not used);
 * the 16 values must be finite and voxel_size, origin still finite (VoxelGrid::validate,
 * core.rs:116-132). Zero-initialise the struct to leave the override off. */
typedef struct vxl_grid {
  uint32_t n;
  float origin[3];
  float voxel_size;
  uint32_t has_world_to_grid;
  float world_to_grid[16];
} vxl_grid;

/* School-B node, 12 bytes: voxel-core/src/node.rs:24-31 (GpuNode). */
typedef struct vxl_node {
  uint32_t mask_lo;
  uint32_t mask_hi;
  uint32_t child_base;
} vxl_node;

typedef struct vxl_tree_info {
  uint32_t n;
  uint32_t k;
  uint64_t node_count;
  uint64_t leaf_count;
  uint64_t level_nodes[16]; /* School-A node count per traversal level L (2..k+1) */
  uint64_t device_bytes;
} vxl_tree_info;

/* ---- context (replaces GpuContext::try_new, voxel-gpu/src/context.rs:24, and
 *      GpuVoxelizer::from_device / new_standalone, voxelizer/src/gpu/mod.rs:172,222) */
vxl_status vxl_ctx_create(int device, vxl_ctx **out);
void vxl_ctx_destroy(vxl_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work. */
void *vxl_ctx_stream(vxl_ctx *ctx);
const char *vxl_last_error(void);
const char *vxl_status_str(vxl_status s);

/* ---- voxelize, dense output: replaces GpuVoxelizer::voxelize_surface
 *      (voxelizer/src/gpu/dense.rs:23-58) and matches voxelize_surface_cpu
 *      (voxelizer/src/reference_cpu.rs:110-200) bit for bit.
 * tris: T x 9 f32 world-space vertices. occ_words: ceil(n^3/32) u32, x-major,
 * bit = x + n*(y + n*z), LSB first (VoxelOccupancy, core.rs:135-253).
 * pairs_out (nullable): number of (triangle, candidate voxel) pairs tested. */
vxl_status vxl_voxelize_dense(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                              float epsilon, uint32_t *occ_words, uint64_t *pairs_out);
vxl_status vxl_voxelize_dense_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                  const vxl_grid *grid, float epsilon, uint32_t *d_occ_words,
                                  uint64_t *pairs_out, void *stream);

/* ---- voxelize, dense output with per-voxel owners and colours: the store_owner /
 *      store_color options of the same call (VoxelizeOpts, voxelizer/src/core.rs:465-501;
 *      VoxelizationOutput.owner_id / color_rgba). owner (nullable): n^3 u32, x-major like
 *      the words, the lowest index of a triangle covering the voxel, UINT32_MAX where none
 *      does (reference_cpu.rs:118-122,163-169). color (nullable): n^3 u32,
 *      hash_color(owner) (reference_cpu.rs:91-99) where a triangle covers the voxel, 0
 *      elsewhere; with owner given this is reference_cpu.rs:173-186, and a colour alone
 *      follows the GPU call, which hashes the owner without storing it
 *      (voxelizer/shaders/voxelize.wgsl:264-281). */
vxl_status vxl_voxelize_dense_owner(vxl_ctx *ctx, const float *tris, uint64_t T,
                                    const vxl_grid *grid, float epsilon, uint32_t *occ_words,
                                    uint32_t *owner, uint32_t *color, uint64_t *pairs_out);
vxl_status vxl_voxelize_dense_owner_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                        const vxl_grid *grid, float epsilon,
                                        uint32_t *d_occ_words, uint32_t *d_owner,
                                        uint32_t *d_color, uint64_t *pairs_out, void *stream);

/* ---- per-voxel materials: the mesh path of the viewer (voxel-viewer/src/main.rs:984-1003),
 *      GpuVoxelizer::compact_surface_sparse (voxelizer/src/gpu/sparse.rs:61-111,
 *      shaders/compact_voxels.wgsl) then tree_from_compact (voxelizer/src/materials.rs:
 *      153-192) over SparseTree::from_voxels (voxel-core/src/sparse.rs:324-366).
 * material_table: two u16 per u32, triangle t's global material id in bits (t & 1) * 16
 * of word t >> 1 (material_table_for_sparse, materials.rs:122-139); a voxel takes the id of
 * its owner (lowest covering triangle, as store_owner), 0 (MISSING) when the owner is past
 * the table. Trees made here carry 512 u16 per leaf in intra-brick Morton order
 * (SparseTree::leaf_materials, sparse.rs:482-484). */
typedef struct vxl_compact_voxel {
  int32_t vx, vy, vz; /* global voxel coordinate: g_origin + grid coordinate */
  uint32_t material;  /* global material id (CompactVoxel, voxelizer/src/core.rs:537-546) */
} vxl_compact_voxel;
/* voxelize + materials into a device tree (the two calls above fused; the tree is the one
 * vxl_voxelize_tree returns, bit for bit). */
vxl_status vxl_voxelize_tree_materials(vxl_ctx *ctx, const float *tris, uint64_t T,
                                       const vxl_grid *grid, float epsilon,
                                       const uint32_t *material_table, uint32_t table_len,
                                       vxl_tree **out, uint64_t *pairs_out);
/* the same with device triangles and device table; the tree is ready when it returns. */
vxl_status vxl_voxelize_tree_materials_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                           const vxl_grid *grid, float epsilon,
                                           const uint32_t *d_table, uint32_t table_len,
                                           vxl_tree **out, uint64_t *pairs_out, void *stream);
/* compact_surface_sparse: every occupied voxel with its material, in leaf (brick Morton)
 * order then intra-brick Morton order. *count receives the number of voxels; out may be
 * NULL to ask for it; a non-NULL out with capacity < *count is VXL_STORAGE_EXCEEDED. */
vxl_status vxl_compact_surface(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                               float epsilon, const uint32_t *material_table, uint32_t table_len,
                               const int32_t g_origin[3], vxl_compact_voxel *out,
                               uint64_t capacity, uint64_t *count);
/* tree_from_compact: voxels outside [0, n)^3 are dropped (counted in *dropped); the
 * material is the low 16 bits; a repeated coordinate keeps the last entry's material. */
vxl_status vxl_tree_from_compact(vxl_ctx *ctx, uint32_t n, const vxl_compact_voxel *voxels,
                                 uint64_t count, vxl_tree **out, uint64_t *dropped);
/* leaf_count x 512 u16 materials of a tree (zeros for a tree made without materials). */
vxl_status vxl_tree_materials_download(const vxl_tree *tree, uint16_t *materials);

/* ---- voxelize straight into a device-resident School-B tree: the composition
 *      voxelize_surface -> VoxelOccupancy::to_school_b (core.rs:224-233) =
 *      SparseTree::build (sparse.rs:241-305) + SchoolBBuffer::from_sparse
 *      (school_b.rs:113-151), without the n^3 host scan. */
vxl_status vxl_voxelize_tree(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                             float epsilon, vxl_tree **out, uint64_t *pairs_out);
vxl_status vxl_voxelize_tree_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                 const vxl_grid *grid, float epsilon, vxl_tree **out,
                                 uint64_t *pairs_out, void *stream);

/* ---- occupancy field -> tree: VoxelOccupancy::to_school_b (core.rs:231-233).
 * words: ceil(n^3/32) x-major u32 (host pointer). */
vxl_status vxl_tree_from_words(vxl_ctx *ctx, const uint32_t *words, uint32_t n, vxl_tree **out);

/* ---- tree handles (replace SchoolBBuffer accessors nodes/leaves/leaf_bounds_words,
 *      school_b.rs:266-281, and upload_structure, voxel-gpu/src/buffers.rs:21-77) */
vxl_status vxl_tree_info_get(const vxl_tree *tree, vxl_tree_info *info);
/* Copies the School-B arrays to host buffers sized from vxl_tree_info. Any output
 * pointer may be NULL. leaf_words: 8 u64 per leaf; codes: brick Morton code per leaf. */
vxl_status vxl_tree_download(const vxl_tree *tree, vxl_node *nodes, uint64_t *leaf_words,
                             uint32_t *bounds, uint64_t *codes);
/* Uploads a host School-B structure (GpuTraverser::new's upload, traverse.rs:70-150).
 * The structure is checked on the host first: every node reachable from the root must
 * address its children inside the arrays (child_base + popcount(mask) <= node_count for
 * nodes above level 2, <= leaf_count for level-2 nodes) and be reached at one level only;
 * otherwise VXL_INVALID_ARG and nothing is uploaded. */
vxl_status vxl_tree_upload(vxl_ctx *ctx, uint32_t n, const vxl_node *nodes, uint64_t node_count,
                           const uint64_t *leaf_words, const uint32_t *bounds,
                           uint64_t leaf_count, vxl_tree **out);
/* Frees a tree once the context's stream has drained; its device memory goes back to the
 * context for the next trees. Work the caller enqueued on its own streams that reads the
 * tree must be complete before the call (synchronise those streams first).
 * A tree lives no longer than the context that made it: vxl_ctx_destroy releases the device
 * memory of its live trees and detaches their handles, and every call given a detached tree
 * returns VXL_INVALID_ARG without device work; vxl_tree_free still releases the handle.
 * Trees collected with vxl_server_wait belong to the server's lane contexts and are detached
 * by vxl_server_destroy in the same way. */
void vxl_tree_free(vxl_tree *tree);

/* ---- traversal: replaces GpuTraverser::traverse (voxel-gpu/src/traverse.rs:154-278),
 *      hit-for-hit identical to mirror_traverse (voxel-core/src/mirror.rs:254-326).
 * rays: R x 6 f32 {origin.xyz, dir.xyz} in voxel space. hits: R x 4 u32 {x,y,z,hit}
 * (the kernel's vec4<u32>, hdda.wgsl:29-33). t (nullable): R f64 ray parameter at
 * which the hit voxel is entered (units of |dir|; 0 for a miss) -- the t_enter of
 * layout::traverse (voxel-core/src/layout.rs:176-183). */
vxl_status vxl_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *rays, uint64_t R,
                        uint32_t *hits, double *t);
vxl_status vxl_traverse_dev(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                            uint32_t *d_hits, double *d_t, void *stream);

/* ---- allocation-free traversal. No counterpart in the reference: it is for callers that
 *      cannot allocate device memory (code under HIP graph capture, tight serving loops,
 *      harnesses that forbid hipMalloc). The same results and lane schedule as
 *      vxl_traverse_dev; the call allocates nothing (no hipMalloc, hipFree or pinned host
 *      memory, no growth of context scratch) and enqueues kernels only. The persistent
 *      schedule's work counter lives in a fixed per-GPU array inside the library (64 counters,
 *      512 bytes, found at vxl_ctx_create and held until the process ends), one counter per
 *      call in turn, zeroed by a kernel on the call's stream. More than 64 of these calls in
 *      flight at once on one GPU race on a counter; replays of a captured call reuse its
 *      counter, so replays must not overlap each other. R above 2^40 rays is
 *      VXL_STORAGE_EXCEEDED. */
vxl_status vxl_traverse_dev_noalloc(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays,
                                    uint64_t R, uint32_t *d_hits, double *d_t, void *stream);

/* ---- procedural noise fixtures: generate_noise_tree (voxel-gpu/src/generate.rs:99) over
 *      NoiseField (voxel-core/src/fixtures.rs:259-346, noise.rs): fBm or ridged Perlin
 *      noise, optionally domain-warped, thresholded; evaluated in f64 with the CPU field's
 *      operations, so the tree is SparseTree::build of the CPU NoiseField bit for bit. Presets
 *      (fixtures.rs:284-314): perlin = {0x5eed1234, 5 octaves, fBm, frequency 5, lacunarity 2,
 *      gain 0.5, warp 0, threshold 0.12}; caves = {0xca5e5eed, 5, ridged, 4, 2, 0.5, 0.45,
 *      0.48}. */
typedef struct vxl_noise {
  uint32_t seed, octaves, ridged, pad;
  double frequency, lacunarity, gain, warp, threshold;
} vxl_noise;
vxl_status vxl_generate_noise_tree(vxl_ctx *ctx, uint32_t n, const vxl_noise *field,
                                   vxl_tree **out);

/* ---- incremental edits on a resident tree: SparseTree::set_voxel (voxel-core/src/
 *      sparse.rs:384-436) for a batch, applied in order (a voxel edited twice takes the
 *      last edit); out-of-bounds edits change nothing. A voxel in an existing brick is
 *      patched in place with its leaf's bounds (Edit::Leaf, the GpuRenderer::update_leaf /
 *      SchoolBBuffer::patch_leaf case, school_b.rs:315-326); a newly set voxel's material
 *      reads 0 again. When a brick appears or empties (Edit::Topology) the tree's arrays are
 *      rebuilt, equal to a fresh build of the edited occupancy (sparse.rs:856-901); the
 *      handle stays valid and surviving bricks keep their materials. *changed (nullable):
 *      voxels whose state changed; *topology (nullable): 1 after a rebuild, else 0. */
typedef struct vxl_voxel_edit {
  uint32_t x, y, z;
  uint32_t occupied; /* 1: set, 0: clear */
} vxl_voxel_edit;
vxl_status vxl_tree_set_voxels(vxl_ctx *ctx, vxl_tree *tree, const vxl_voxel_edit *edits,
                               uint64_t count, uint64_t *changed, int *topology);

/* ---- the plain per-tree traversal launches captured in a HIP graph (C4's many ray batches
 *      as one replayable submission; no reference counterpart: the reference submits each
 *      traversal on its own). Captures, for g < count, the one-lane-per-ray traversal of
 *      rays [ray_offsets[g], ray_offsets[g+1]) against trees[g] into d_hits / d_t, as four
 *      independent chains that fork and join (so traversals overlap on the device); the
 *      graph holds the trees' and buffers' device addresses, which must outlive it. Replays
 *      give exactly vxl_traverse_dev's results. */
typedef struct vxl_graph vxl_graph;
vxl_status vxl_traverse_graph_create(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                                     const float *d_rays, const uint64_t *ray_offsets,
                                     uint32_t *d_hits, double *d_t, vxl_graph **out);
vxl_status vxl_graph_launch(vxl_graph *graph, void *stream);
void vxl_graph_destroy(vxl_graph *graph);

/* ---- a whole batch as one HIP graph: every mesh voxelized, its tree built and its rays
 *      walked, with no host round trip (C4's hundreds of meshes and uneven ray batches as
 *      one replayable submission; no reference counterpart: the reference voxelizes mesh by
 *      mesh, crates/voxel-cli/src/main.rs:1288-1300, and traverses each tree on its own,
 *      main.rs:785-792). Creation runs the batch once on the current triangles to size an
 *      arena for the trees (headroom x their size; 0 means 1.5), then captures, for every
 *      batch group, setup, SAT and tree build -- the tree sizes stay on the device, which
 *      places each tree in the group's arena region -- and one traversal kernel over all
 *      segments: rays [ray_offsets[g], ray_offsets[g+1]) of d_rays against mesh g's tree,
 *      hits to d_hits (4 u32 per ray) and t to d_t (nullable); ray_offsets null: no rays.
 *      Every mesh needs n in 128..2048 (k = 2..4). Each replay reads d_tris and d_rays and
 *      writes d_hits and d_t: their contents may change between replays, their sizes and
 *      the grids may not. Replays use the graph's own scratch, so a replay launched on
 *      another stream first waits for the previous one. vxl_batch_graph_trees, after a launch:
 *      waits for it and returns copies of its trees (owned by ctx; trees and pairs
 *      nullable), or VXL_STORAGE_EXCEEDED when a group's trees did not fit its region (the
 *      hits of that group's meshes are then misses; a later replay whose trees fit is whole
 *      again). vxl_batch_graph_info: the arena's bytes, the batch groups, the graph's nodes. */
typedef struct vxl_batch_graph vxl_batch_graph;
vxl_status vxl_batch_graph_create(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                                  uint32_t count, const vxl_grid *grids, float epsilon,
                                  const float *d_rays, const uint64_t *ray_offsets,
                                  uint32_t *d_hits, double *d_t, double headroom,
                                  vxl_batch_graph **out);
vxl_status vxl_batch_graph_launch(vxl_batch_graph *graph, void *stream);
vxl_status vxl_batch_graph_trees(vxl_batch_graph *graph, vxl_tree **trees, uint64_t *pairs);
vxl_status vxl_batch_graph_info(const vxl_batch_graph *graph, uint64_t *arena_bytes,
                                uint32_t *groups, uint64_t *graph_nodes);
void vxl_batch_graph_destroy(vxl_batch_graph *graph);

/* ---- walk statistics of a ray batch (the reference's measurement counters,
 *      voxel-core/src/measure.rs:189-235, over the mirror's walk mirror.rs:263-326): sums
 *      over the rays of descents into leaves, descents into nodes and cell visits, in
 *      out[0..2]. Device rays, host result; for rooflines, not on the hot path. */
vxl_status vxl_traverse_stats(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                              uint64_t out[3]);

/* ---- device ray generation + shading: the render path (GpuRenderer,
 *      voxel-gpu/src/render.rs:211-589, shaders/render.wgsl:54-91), one invocation per pixel
 *      with no host rays and no readback of hits. Camera layout = render.wgsl's Camera: eye
 *      in grid units, tan of half the vertical field of view, unit forward / right / up,
 *      aspect = width / height, n = the tree's grid size. d_colors: the global colour table
 *      (MaterialTable words, RGBA8 with R in the low byte), indexed by a voxel's material;
 *      material 0 (or a tree without materials) is shaded by position. d_rgba: width x height
 *      u32, row-major from the top row. */
typedef struct vxl_camera {
  float eye[3], tan_half_fov;
  float forward[3], aspect;
  float right[3], n;
  float up[3], pad0;
  uint32_t width, height, pad1, pad2;
} vxl_camera;
vxl_status vxl_render_dev(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera *camera,
                          const uint32_t *d_colors, uint32_t ncolors, uint32_t *d_rgba,
                          void *stream);
/* the same with a host colour table and a host image (copied back; synchronous). */
vxl_status vxl_render(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera *camera,
                      const uint32_t *colors, uint32_t ncolors, uint32_t *rgba);

/* ---- batches (no single reference counterpart: each replaces a host loop of the calls
 *      above, e.g. the per-mesh GpuVoxelizer::voxelize_surface -> to_school_b ->
 *      GpuTraverser::new/traverse sequence of voxel-cli/src/main.rs:1288-1300 and
 *      :785-792, over many meshes and ray batches).
 * Meshes: d_tris holds every mesh's triangles back to back (T x 9 f32, device);
 * tri_offsets[count+1] (host) delimits mesh g = triangles [tri_offsets[g],
 * tri_offsets[g+1]); grids[count]; out[count] receives one tree per mesh; pairs_out
 * (nullable) one pair count per mesh. Each tree equals the one vxl_voxelize_tree_dev
 * returns for that mesh alone. Meshes run in groups whose dense grids fit the batch
 * budget (vxl_set_batch_budget): per group one setup, one SAT and one tree-build pass
 * over all its meshes, with two host syncs in total. Trees of one group share one device
 * allocation, released when the last of them is freed. */
vxl_status vxl_voxelize_batch_dev(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                                  uint32_t count, const vxl_grid *grids, float epsilon,
                                  vxl_tree **out, uint64_t *pairs_out, void *stream);
/* Host-pointer form; tri_offsets[0] must be 0. Synchronous. */
vxl_status vxl_voxelize_batch(vxl_ctx *ctx, const float *tris, const uint64_t *tri_offsets,
                              uint32_t count, const vxl_grid *grids, float epsilon, vxl_tree **out,
                              uint64_t *pairs_out);
/* Rays: segment g = rays [ray_offsets[g], ray_offsets[g+1]) of d_rays (R x 6 f32) against
 * trees[g], in ONE kernel launch; hits / t as for vxl_traverse_dev, indexed like the rays.
 * Results equal per-segment vxl_traverse_dev calls. */
vxl_status vxl_traverse_batch_dev(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                                  const float *d_rays, const uint64_t *ray_offsets,
                                  uint32_t *d_hits, double *d_t, void *stream);
/* Host-pointer form; ray_offsets[0] must be 0. Synchronous. */
vxl_status vxl_traverse_batch(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                              const float *rays, const uint64_t *ray_offsets, uint32_t *hits,
                              double *t);
/* Dense-grid bytes one batch group may use (default 16 GiB; 1..128 GiB). */
vxl_status vxl_set_batch_budget(vxl_ctx *ctx, uint64_t bytes);

typedef struct vxl_tuning {
  int sat_waves_per_cu, sat_units_per_wave, dda_block;
  float voxelize_ms, traverse_ms; /* batch times with the chosen values */
  int sat_merge_words, dda_batch_refill;
} vxl_tuning;
/* Autotuner: times vxl_voxelize_batch_dev on the given meshes for each SAT launch shape
 * (resident waves per CU x work units per wave: the work-queue chunk a wave takes is the
 * pairs / (CUs x waves x units) candidate pairs of one unit), keeps the fastest, then for
 * it each merge-tile size (leaf-grid words a wave merges in LDS before writing them out:
 * 16, 32, 48, 64), then times vxl_traverse_batch_dev on the given rays (nullable) for each
 * DDA workgroup size with one lane per ray and for persistent lanes refilled at 8, 16 or 32
 * idle lanes (vxl_set_batch_refill); the chosen values stay set on ctx and are returned in
 * *result. Trees made while tuning are freed; d_hits / d_t are overwritten. */
vxl_status vxl_autotune_batch(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                              uint32_t count, const vxl_grid *grids, float epsilon,
                              const float *d_rays, const uint64_t *ray_offsets, uint32_t *d_hits,
                              double *d_t, vxl_tuning *result);

/* ---- serving (no reference counterpart: the reference runs one blocking call at a time;
 *      this is a request loop over the calls above). Trees stay resident in HBM; ray
 *      queries and voxelize requests are admitted by device-byte footprint under a budget,
 *      batched as they arrive (continuous batching) and overlapped on the device. Each
 *      result equals the one the request gets alone through the per-request entry points.
 * Footprints: a ray query R * 48 bytes (rays, hits, t); a voxelize request
 * T * 116 bytes + its dense-grid region + an output-tree estimate. */
typedef struct vxl_server vxl_server;
typedef struct vxl_server_config {
  uint64_t budget_bytes;       /* admission budget over in-flight footprints + held bytes
                                  (0 = 200e9) */
  uint32_t batching;           /* 1: continuous batching on two lanes (rays, voxelize);
                                  0: plain, one request per launch in arrival order */
  uint32_t max_batch_requests; /* requests per launch (0 = 1024) */
  uint64_t max_batch_rays;     /* rays per traversal launch (0 = 64 Mi) */
  uint64_t max_batch_tris;     /* triangles per voxelize batch (0 = 16 Mi) */
  uint32_t batch_refill;       /* lane schedule of the batched ray launches: 0 one lane per
                                  ray; 1..64 persistent lanes (vxl_set_batch_refill) for
                                  launches of at least four times the rays the device's
                                  resident waves hold (smaller ones: one lane per ray) */
} vxl_server_config;
typedef struct vxl_server_stats {
  uint64_t submitted, completed, failed, batches;
  double mean_batch_requests;
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion, all requests */
  double ray_p50_ms, ray_p99_ms, vox_p50_ms, vox_p99_ms;
  double throughput_rps;                  /* completed / (last completion - first submit) */
  uint64_t peak_bytes;                    /* peak held + in-flight footprint */
} vxl_server_stats;
vxl_status vxl_server_create(int device, const vxl_server_config *cfg, vxl_server **out);
/* Serves every pending request, then stops. Voxelize results not yet collected with
 * vxl_server_wait are released; collected ones are detached (see vxl_tree_free). */
void vxl_server_destroy(vxl_server *server);
/* Adds (or, negative, releases) bytes held resident, e.g. the trees queries run against. */
vxl_status vxl_server_hold(vxl_server *server, int64_t delta_bytes);
/* Enqueue; device buffers must stay valid until the request completes. VXL_STORAGE_EXCEEDED
 * if the footprint can never fit the budget. */
vxl_status vxl_server_submit_rays(vxl_server *server, const vxl_tree *tree, const float *d_rays,
                                  uint64_t R, uint32_t *d_hits, double *d_t, uint64_t *id);
vxl_status vxl_server_submit_voxelize(vxl_server *server, const float *d_tris, uint64_t T,
                                      const vxl_grid *grid, float epsilon, uint64_t *id);
vxl_status vxl_server_poll(vxl_server *server, uint64_t id, int *done);
/* Blocks until request id completes and retires it: its status is returned; a voxelize
 * request's tree goes to *tree_out (freed when tree_out is NULL). One wait per request: a
 * second wait on an id already being waited on (or retired) returns VXL_INVALID_ARG. */
vxl_status vxl_server_wait(vxl_server *server, uint64_t id, vxl_tree **tree_out,
                           uint64_t *pairs_out, double *latency_ms);
vxl_status vxl_server_stats_get(vxl_server *server, vxl_server_stats *out);
vxl_status vxl_server_reset_stats(vxl_server *server);

/* ---- timing: when enabled, the last voxelize/traverse call records per-stage
 * device times (ms) measured with HIP events on the stream the kernels ran on.
 * stage 0 = setup+scan, 1 = SAT, 2 = tree build, 3 = traverse, 4 = whole voxelize
 * batch, 5 = traverse batch, 6 = whole vxl_voxelize_tree_materials_dev, 7 = render. */
void vxl_set_timing(vxl_ctx *ctx, int enable);
void vxl_last_timing(vxl_ctx *ctx, float out_ms[8]);

/* ---- tuning knobs (the autotuner's surface; no reference counterpart: the reference's
 * GpuVoxelizerConfig{workgroup_size, tiles_per_workgroup}, voxelizer/src/gpu/mod.rs:61-75,
 * plays this role for its WGSL kernel). sat_waves_per_cu: resident SAT waves per CU;
 * sat_units_per_wave: contiguous work units per resident wave; dda_block: DDA workgroup
 * size. 0 keeps the current value. */
vxl_status vxl_set_tuning(vxl_ctx *ctx, int sat_waves_per_cu, int sat_units_per_wave,
                          int dda_block);
/* Lane schedule of the batched traversals (vxl_traverse_batch[_dev], the batch graph):
 * 0 (default) one lane per ray; refill_min in [1, 64] persistent lanes that take runs of 256
 * rays of one segment and hand idle lanes the next ray once refill_min of a wave's lanes are
 * idle. Results do not depend on it. */
vxl_status vxl_set_batch_refill(vxl_ctx *ctx, int refill_min);
/* The SAT's merge tile: leaf-grid words (4^3 sub-bricks) a wave merges in its LDS table
 * before it ORs them into the grid, in [1, 64] (default 64). Results do not depend on it. */
vxl_status vxl_set_merge_words(vxl_ctx *ctx, int words);
/* DDA lane schedule of vxl_traverse / vxl_traverse_dev: refill_min in [1, 64] runs a
 * persistent grid whose waves hand a lane the next ray once refill_min of their lanes are
 * idle; 0 launches one lane per ray; -1 (the default) picks per call: persistent lanes
 * (refill 8) for trees whose leaves, bounds and nodes exceed 256 MiB, and for trees of
 * k >= 4 levels when the call holds at least 4 x the rays the device's resident waves walk
 * at once; else one lane per ray. Results are the same either way. */
vxl_status vxl_set_dda_refill(vxl_ctx *ctx, int refill_min);

#ifdef __cplusplus
}
#endif
#endif /* VXL_H */
