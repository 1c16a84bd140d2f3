/*
 * vxl_oracle.h -- CPU restatement of the reference voxelize -> sparse tree -> DDA path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker and the CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (friendly-journey_amd/libvxl.so) never links, loads or calls it.
 *
 * Every function cites the reference source it restates (paths relative to the
 * reference repository root). Arithmetic is IEEE binary32/binary64 with no FMA
 * contraction (built -ffp-contract=off, no fast-math), in the reference's order.
 */
#ifndef VXL_ORACLE_H
#define VXL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- morton.rs / node.rs / leaf.rs / resolution.rs ---- */
uint64_t vo_morton_encode(uint32_t x, uint32_t y, uint32_t z);
void vo_morton_decode(uint64_t code, uint32_t out_xyz[3]);
uint32_t vo_encode_brick(uint32_t x, uint32_t y, uint32_t z);
uint32_t vo_child_bit(uint32_t cx, uint32_t cy, uint32_t cz);
uint32_t vo_child_slot(uint32_t mask_lo, uint32_t mask_hi, uint32_t child_base, uint32_t bit);
/* 0 ok (k set), 1 too small, 2 too large, 3 not representable (below/above set) */
int vo_resolution(uint32_t n, uint32_t *k, uint32_t *below, uint32_t *above);
uint32_t vo_bounds_pack(const uint32_t mn[3], const uint32_t mx[3]);
void vo_bounds_unpack(uint32_t p, uint32_t mn[3], uint32_t mx[3]);
uint32_t vo_leaf_occupied_bounds(const uint64_t bits[8]);

/* ---- reference_cpu.rs ---- */
int vo_triangle_box_overlap(const float center[3], const float half[3], const float v0[3],
                            const float v1[3], const float v2[3]);
/* world->grid point transform of core.rs:101-107 (glam Mat4 product, see .c) */
void vo_world_to_grid(const float p[3], const float origin[3], float voxel_size, float out[3]);
/* voxelize_surface_cpu: tris T x 9 f32 world space; writes ceil(n^3/32) x-major words
 * (zeroed here) and, when owner != NULL, n^3 lowest-triangle owners. Returns the number
 * of (triangle, candidate voxel) pairs the epsilon-padded AABB scan enumerates. */
uint64_t vo_voxelize_surface(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                             float voxel_size, float eps, uint32_t *words, uint32_t *owner);
/* the world_to_grid override (core.rs:30-34,101-104): glam Mat4::transform_point3 of a point,
 * column-major m; and voxelize_surface_cpu on a grid carrying that override (serial). */
void vo_world_to_grid_mat(const float p[3], const float m[16], float out[3]);
uint64_t vo_voxelize_surface_mat(const float *tris, uint64_t T, uint32_t n, const float m[16],
                                 float eps, uint32_t *words, uint32_t *owner);
/* vo_voxelize_surface with the scan split over threads by z-plane rows: the same words,
 * owners and pair count (occupancy is an OR, the owner a minimum); for fast checks only. */
uint64_t vo_voxelize_surface_mt(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                                float voxel_size, float eps, uint32_t *words, uint32_t *owner,
                                int threads);
/* hash_color (reference_cpu.rs:91-99) and the colour pass of reference_cpu.rs:173-186. */
uint32_t vo_hash_color(uint32_t id);
void vo_owner_colors(const uint32_t *owner, uint64_t N, uint32_t *color);
/* materials: compact_voxels.wgsl lookup, compact_surface_sparse from dense outputs, and
 * tree_from_compact / SparseTree::from_voxels (see .c). */
uint32_t vo_resolve_material(uint32_t tri, const uint32_t *table, uint32_t table_len);
uint64_t vo_compact(const uint32_t *words, const uint32_t *owner, uint32_t n,
                    const uint32_t *table, uint32_t table_len, const int32_t g_origin[3],
                    int32_t *out);
uint64_t vo_compact_to_words(const int32_t *vox, uint64_t count, uint32_t n, uint32_t *words);
void vo_compact_materials(const int32_t *vox, uint64_t count, uint32_t n, const uint64_t *codes,
                          uint64_t L, uint16_t *mats);
/* noise.rs / fixtures.rs NoiseField in f64: f = {frequency, lacunarity, gain, warp,
 * threshold}, u = {seed, octaves, ridged}. */
double vo_noise_grad(uint32_t hash, double x, double y, double z);
double vo_noise_fade(double t);
double vo_perlin3(uint32_t seed, double x, double y, double z);
double vo_noise_value(const double f[5], const uint32_t u[3], uint32_t n, uint32_t x, uint32_t y,
                      uint32_t z);
uint64_t vo_noise_words(const double f[5], const uint32_t u[3], uint32_t n, uint32_t *words);
/* Same candidate count without doing the SAT (for roofline accounting). */
uint64_t vo_count_candidates(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                             float voxel_size, float eps);

/* ---- sparse.rs + school_b.rs ---- */
typedef struct vo_tree vo_tree;
/* SparseTree::build over an x-major word field, then SchoolBBuffer::from_sparse. */
vo_tree *vo_tree_build(const uint32_t *words, uint32_t n, int threads);
/* Wall seconds of the last build's stages: out[0] SparseTree::build (brick scan, Morton sort,
 * build_levels), out[1] SchoolBBuffer::from_sparse (parent-first emit + leaf bounds). */
void vo_tree_stage_seconds(const vo_tree *t, double out[2]);
/* The same tree, reading each 8-voxel brick row of the words as one byte (for checks). */
vo_tree *vo_tree_build_rows(const uint32_t *words, uint32_t n, int threads);
/* Wrap School-B arrays produced elsewhere (e.g. by the GPU) for traversal. */
vo_tree *vo_tree_from_school_b(uint32_t n, const uint32_t *nodes, uint64_t node_count,
                               const uint64_t *leaves, uint64_t leaf_count,
                               const uint32_t *bounds);
void vo_tree_free(vo_tree *t);
/* counts[0]=leaves, counts[1]=School-B nodes, counts[2+L]=School-A nodes at level L
 * (L in 0..15; 0 and 1 unused). */
void vo_tree_counts(const vo_tree *t, uint64_t counts[18]);
void vo_tree_export(const vo_tree *t, uint32_t *nodes_b /*3*N*/, uint64_t *leaves /*8*L*/,
                    uint32_t *bounds /*L*/, uint64_t *codes /*L, may be NULL*/);
/* School-A level L nodes (3 u32 each). */
void vo_tree_export_level(const vo_tree *t, uint32_t level, uint32_t *nodes_a);
int vo_tree_is_occupied(const vo_tree *t, uint32_t x, uint32_t y, uint32_t z);

/* ---- mirror.rs (f32) ---- rays R x 6 f32; hits R x 4 u32 (x,y,z,flag).
 * stats (nullable) R x 3 u32: leaf descents, node descents, cell steps. */
void vo_mirror_traverse(const vo_tree *t, const float *rays, uint64_t R, uint32_t *hits,
                        int enable_skip, uint32_t *stats, int threads);
/* render.wgsl:54-91: camera = {eye[3], tan, forward[3], aspect, right[3], n, up[3], pad}. */
void vo_camera_ray(const float cam[16], uint32_t width, uint32_t height, uint32_t px,
                   uint32_t py, float ray[6]);
void vo_render(const vo_tree *t, const float cam[16], uint32_t width, uint32_t height,
               const uint16_t *mats, const uint32_t *colors, uint32_t ncolors, uint32_t *out,
               int threads);

/* ---- layout.rs (f64) ---- rays R x 6 f64; school 0 = School-A arrays, 1 = School-B. */
void vo_layout_traverse(const vo_tree *t, const double *rays, uint64_t R, uint32_t *hits,
                        double *t_enter, int school, int threads);
/* ---- oracle.rs (dense f64 tier-A) ---- */
void vo_first_hit_dense(const uint32_t *words, uint32_t n, const double *rays, uint64_t R,
                        uint32_t *hits, double *t_enter);

/* ---- fixtures.rs procedural fields as x-major words ----
 * kind: 0 Empty, 1 Solid, 2 SingleVoxel(p0,p1,p2), 3 Checkerboard, 4 OctantFractal(keep=p0),
 *       5 WireLattice(period=p0, thickness=p1), 6 Dust(density=p0) */
void vo_fixture_words(int kind, uint32_t n, uint32_t p0, uint32_t p1, uint32_t p2,
                      uint32_t *words);

#ifdef __cplusplus
}
#endif
#endif
