/*
 * vxl_oracle.c -- CPU restatement of the reference voxelize -> sparse tree -> DDA path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline ("port").
 * The product library never links or calls this code.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (x86-64 SSE2: binary32 ops are
 * evaluated in binary32, FLT_EVAL_METHOD == 0, denormals preserved), which is the
 * arithmetic the reference's Rust f32/f64 code gets on the same host.
 *
 * Rust semantics restated explicitly:
 *   f32::min/max       -> fminf/fmaxf (IEEE minNum/maxNum: the non-NaN operand wins)
 *   f32::clamp(lo,hi)  -> if x<lo x=lo; if x>hi x=hi  (NaN passes through)
 *   `as i32`/`as u32`  -> saturating conversion, NaN -> 0 (sat_i32 / sat_u32 below)
 *   glam Vec3::dot     -> (x*x' + y*y') + z*z'   (left to right)
 *   glam Vec3::cross   -> (y*z' - y'*z, z*x' - z'*x, x*y' - x'*y)
 */
#include "vxl_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------- */
/* Rust numeric-cast semantics                                               */
/* ------------------------------------------------------------------------- */
static int32_t sat_i32(float v) {
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}
static uint32_t sat_u32f(float v) {
  if (v != v || v <= 0.0f) return 0;
  if (v >= 4294967296.0f) return UINT32_MAX;
  return (uint32_t)v;
}
static uint32_t sat_u32d(double v) {
  if (v != v || v <= 0.0) return 0;
  if (v >= 4294967296.0) return UINT32_MAX;
  return (uint32_t)v;
}

/* ------------------------------------------------------------------------- */
/* morton.rs:25-89                                                           */
/* ------------------------------------------------------------------------- */
static uint64_t split_by_3(uint32_t v) { /* morton.rs:25-33 */
  uint64_t x = (uint64_t)v & 0x1fffffULL;
  x = (x | (x << 32)) & 0x001f00000000ffffULL;
  x = (x | (x << 16)) & 0x001f0000ff0000ffULL;
  x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2)) & 0x1249249249249249ULL;
  return x;
}
static uint32_t compact_by_3(uint64_t v) { /* morton.rs:39-47 */
  uint64_t x = v & 0x1249249249249249ULL;
  x = (x | (x >> 2)) & 0x10c30c30c30c30c3ULL;
  x = (x | (x >> 4)) & 0x100f00f00f00f00fULL;
  x = (x | (x >> 8)) & 0x001f0000ff0000ffULL;
  x = (x | (x >> 16)) & 0x001f00000000ffffULL;
  x = (x | (x >> 32)) & 0x1fffffULL;
  return (uint32_t)x;
}
uint64_t vo_morton_encode(uint32_t x, uint32_t y, uint32_t z) { /* morton.rs:55-61 */
  return split_by_3(x) | (split_by_3(y) << 1) | (split_by_3(z) << 2);
}
void vo_morton_decode(uint64_t code, uint32_t out[3]) { /* morton.rs:71-77 */
  out[0] = compact_by_3(code);
  out[1] = compact_by_3(code >> 1);
  out[2] = compact_by_3(code >> 2);
}
uint32_t vo_encode_brick(uint32_t x, uint32_t y, uint32_t z) { /* morton.rs:86-89 */
  return (uint32_t)vo_morton_encode(x & 7, y & 7, z & 7);
}
uint32_t vo_child_bit(uint32_t cx, uint32_t cy, uint32_t cz) { /* node.rs:78-83 */
  return vo_encode_brick(cx, cy, cz) & 63;
}
uint32_t vo_child_slot(uint32_t lo, uint32_t hi, uint32_t base, uint32_t bit) { /* node.rs:64-72 */
  uint64_t mask = ((uint64_t)hi << 32) | lo;
  uint64_t below = bit == 0 ? 0 : (mask & ((1ULL << bit) - 1));
  return base + (uint32_t)__builtin_popcountll(below);
}

/* resolution.rs:76-118 */
int vo_resolution(uint32_t n, uint32_t *k, uint32_t *below, uint32_t *above) {
  const uint32_t MIN_RES = 8, MAX_RES = 1u << 31;
  if (n < MIN_RES) return 1;
  if (n > MAX_RES) return 2;
  if (n % 8 == 0) {
    uint32_t m = n / 8;
    if ((m & (m - 1)) == 0 && (__builtin_ctz(m) % 2) == 0) {
      if (k) *k = (uint32_t)__builtin_ctz(m) / 2;
      return 0;
    }
  }
  uint32_t b = MIN_RES, a = MIN_RES, r = MIN_RES;
  while (r < n) {
    b = r;
    r *= 4;
    a = r;
  }
  if (below) *below = b;
  if (above) *above = a;
  return 3;
}

/* leaf.rs:67-92 */
uint32_t vo_bounds_pack(const uint32_t mn[3], const uint32_t mx[3]) {
  return (mn[0] & 7) | ((mn[1] & 7) << 3) | ((mn[2] & 7) << 6) | ((mx[0] & 7) << 9) |
         ((mx[1] & 7) << 12) | ((mx[2] & 7) << 15);
}
void vo_bounds_unpack(uint32_t p, uint32_t mn[3], uint32_t mx[3]) {
  mn[0] = p & 7; mn[1] = (p >> 3) & 7; mn[2] = (p >> 6) & 7;
  mx[0] = (p >> 9) & 7; mx[1] = (p >> 12) & 7; mx[2] = (p >> 15) & 7;
}
static const uint32_t FULL_BOUNDS = (7u << 9) | (7u << 12) | (7u << 15); /* leaf.rs:44-47 */

/* leaf.rs:185-213 (BOUNDS_SCAN_LIMIT = 64) */
uint32_t vo_leaf_occupied_bounds(const uint64_t bits[8]) {
  uint32_t count = 0;
  for (int w = 0; w < 8; w++) count += (uint32_t)__builtin_popcountll(bits[w]);
  if (count > 64) return FULL_BOUNDS;
  uint32_t mn[3] = {7, 7, 7}, mx[3] = {0, 0, 0};
  int any = 0;
  for (uint32_t w = 0; w < 8; w++) {
    uint64_t b = bits[w];
    while (b) {
      any = 1;
      uint32_t i = w * 64 + (uint32_t)__builtin_ctzll(b);
      uint32_t c[3];
      vo_morton_decode(i, c);
      for (int a = 0; a < 3; a++) {
        if (c[a] < mn[a]) mn[a] = c[a];
        if (c[a] > mx[a]) mx[a] = c[a];
      }
      b &= b - 1;
    }
  }
  return any ? vo_bounds_pack(mn, mx) : FULL_BOUNDS;
}

/* ------------------------------------------------------------------------- */
/* reference_cpu.rs:27-85 -- Akenine-Moller 13-axis SAT, in the reference order */
/* ------------------------------------------------------------------------- */
static float dot3(const float a[3], const float b[3]) { /* glam Vec3::dot */
  float s = a[0] * b[0];
  s = s + a[1] * b[1];
  s = s + a[2] * b[2];
  return s;
}

int vo_triangle_box_overlap(const float bc[3], const float bh[3], const float a0[3],
                            const float a1[3], const float a2[3]) {
  float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
  for (int i = 0; i < 3; i++) { /* :34-36 */
    v0[i] = a0[i] - bc[i];
    v1[i] = a1[i] - bc[i];
    v2[i] = a2[i] - bc[i];
  }
  for (int i = 0; i < 3; i++) { /* :38-40 */
    e0[i] = v1[i] - v0[i];
    e1[i] = v2[i] - v1[i];
    e2[i] = v0[i] - v2[i];
  }
  const float *es[3] = {e0, e1, e2};
  float axes[9][3];
  for (int j = 0; j < 3; j++) { /* :42-52 */
    const float *e = es[j];
    axes[j][0] = 0.0f; axes[j][1] = -e[2]; axes[j][2] = e[1];
    axes[3 + j][0] = e[2]; axes[3 + j][1] = 0.0f; axes[3 + j][2] = -e[0];
    axes[6 + j][0] = -e[1]; axes[6 + j][1] = e[0]; axes[6 + j][2] = 0.0f;
  }
  for (int j = 0; j < 9; j++) { /* :54-64 */
    const float *ax = axes[j];
    float p0 = dot3(v0, ax), p1 = dot3(v1, ax), p2 = dot3(v2, ax);
    float min_p = fminf(p0, fminf(p1, p2));
    float max_p = fmaxf(p0, fmaxf(p1, p2));
    float r = bh[0] * fabsf(ax[0]);
    r = r + bh[1] * fabsf(ax[1]);
    r = r + bh[2] * fabsf(ax[2]);
    if (min_p > r || max_p < -r) return 0;
  }
  for (int i = 0; i < 3; i++) { /* :66-74 */
    if (fminf(v0[i], fminf(v1[i], v2[i])) > bh[i] || fmaxf(v0[i], fmaxf(v1[i], v2[i])) < -bh[i])
      return 0;
  }
  float nrm[3]; /* :76 e0.cross(e1) */
  nrm[0] = e0[1] * e1[2] - e1[1] * e0[2];
  nrm[1] = e0[2] * e1[0] - e1[2] * e0[0];
  nrm[2] = e0[0] * e1[1] - e1[0] * e0[1];
  float d = -dot3(nrm, v0); /* :77 */
  float r = bh[0] * fabsf(nrm[0]);
  r = r + bh[1] * fabsf(nrm[1]);
  r = r + bh[2] * fabsf(nrm[2]); /* :78 */
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  float s = dot3(nrm, zero) + d; /* :79 */
  if (fabsf(s) > r) return 0;
  return 1;
}

/* core.rs:101-107: scale(1/vs) * translate(-origin), applied with transform_point3.
 * Per component the glam Mat4 product is ((inv*p + 0*q) + 0*r) + inv*(-o); the zero
 * products are exact no-ops for finite p, leaving (inv*p) + (inv*(-o)). */
void vo_world_to_grid(const float p[3], const float origin[3], float voxel_size, float out[3]) {
  float inv = 1.0f / voxel_size;
  for (int i = 0; i < 3; i++) {
    float t = inv * (-origin[i]);
    float s = inv * p[i];
    out[i] = s + t;
  }
}

static void tri_box(const float g[9], uint32_t n, float eps, int32_t mn[3], int32_t mx[3]) {
  /* reference_cpu.rs:139-150 */
  for (int a = 0; a < 3; a++) {
    float lo = fminf(fminf(g[a], g[3 + a]), g[6 + a]) - eps;
    float hi = fmaxf(fmaxf(g[a], g[3 + a]), g[6 + a]) + eps;
    mn[a] = sat_i32(fmaxf(floorf(lo), 0.0f));
    mx[a] = sat_i32(fminf(floorf(hi), (float)(n - 1)));
  }
}

/* core.rs:101-104 with the world_to_grid override, through glam's Mat4::transform_point3
 * (column-major m: m[0..3] = x_axis, m[4..7] = y_axis, m[8..11] = z_axis, m[12..15] = w_axis):
 * res = x_axis * p.x; res = y_axis * p.y + res; res = z_axis * p.z + res; res = w_axis + res,
 * lane by lane, each product and sum rounded (glam 0.33 src/f32/{scalar,sse2}/mat4.rs). */
void vo_world_to_grid_mat(const float p[3], const float m[16], float out[3]) {
  for (int i = 0; i < 3; i++) {
    float r = m[i] * p[0];
    r = m[4 + i] * p[1] + r;
    r = m[8 + i] * p[2] + r;
    out[i] = m[12 + i] + r;
  }
}

static const float *g_mat = NULL; /* set by vo_voxelize_surface_mat for its call (serial) */

static void tri_to_grid(const float *tri, const float origin[3], float vs, float g[9]) {
  for (int v = 0; v < 3; v++) {
    if (g_mat)
      vo_world_to_grid_mat(tri + 3 * v, g_mat, g + 3 * v);
    else
      vo_world_to_grid(tri + 3 * v, origin, vs, g + 3 * v);
  }
}

uint64_t vo_count_candidates(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                             float vs, float eps) {
  uint64_t pairs = 0;
  for (uint64_t t = 0; t < T; t++) {
    float g[9];
    int32_t mn[3], mx[3];
    tri_to_grid(tris + 9 * t, origin, vs, g);
    tri_box(g, n, eps, mn, mx);
    if (mx[0] < mn[0] || mx[1] < mn[1] || mx[2] < mn[2]) continue;
    pairs += (uint64_t)(mx[0] - mn[0] + 1) * (uint64_t)(mx[1] - mn[1] + 1) *
             (uint64_t)(mx[2] - mn[2] + 1);
  }
  return pairs;
}

/* reference_cpu.rs:91-99: three steps of the LCG x -> 1664525 x + 1013904223 (mod 2^32)
 * from the id, the low byte of each as r, g, b; alpha 255; bytes little-endian. */
uint32_t vo_hash_color(uint32_t id) {
  uint32_t x = id * 1664525u + 1013904223u;
  uint8_t r = (uint8_t)(x & 0xff);
  x = x * 1664525u + 1013904223u;
  uint8_t g = (uint8_t)(x & 0xff);
  x = x * 1664525u + 1013904223u;
  uint8_t b = (uint8_t)(x & 0xff);
  return (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | (255u << 24);
}

/* reference_cpu.rs:173-186 with store_owner set: colour of each owned voxel, 0 elsewhere. */
void vo_owner_colors(const uint32_t *owner, uint64_t N, uint32_t *color) {
  for (uint64_t i = 0; i < N; i++) color[i] = owner[i] != UINT32_MAX ? vo_hash_color(owner[i]) : 0u;
}

/* compact_voxels.wgsl:46-66: the material of triangle `tri` from a table of two u16 per
 * u32 (triangle t in bits (t & 1) * 16 of word t >> 1); 0 (MISSING) for no owner, an empty
 * table or a triangle past its end. */
uint32_t vo_resolve_material(uint32_t tri, const uint32_t *table, uint32_t table_len) {
  if (tri == UINT32_MAX || table_len == 0) return 0;
  const uint32_t w = tri >> 1;
  if (w >= table_len) return 0;
  return (table[w] >> ((tri & 1u) << 4)) & 0xFFFFu;
}

/* compact_surface_sparse (voxelizer/src/gpu/sparse.rs:61-111, compact_voxels.wgsl) restated
 * from the dense outputs: every occupied voxel as (g_origin + (x, y, z), material of its
 * owner), x fastest; returns the count. out (count x 4 i32) may be NULL to count only. */
uint64_t vo_compact(const uint32_t *words, const uint32_t *owner, uint32_t n,
                    const uint32_t *table, uint32_t table_len, const int32_t g_origin[3],
                    int32_t *out) {
  const uint64_t nn = n, nv = nn * nn * nn;
  uint64_t c = 0;
  for (uint64_t lin = 0; lin < nv; lin++) {
    if (!((words[lin >> 5] >> (lin & 31)) & 1u)) continue;
    if (out) {
      const uint32_t x = (uint32_t)(lin % nn), y = (uint32_t)((lin / nn) % nn),
                     z = (uint32_t)(lin / (nn * nn));
      out[4 * c + 0] = g_origin[0] + (int32_t)x;
      out[4 * c + 1] = g_origin[1] + (int32_t)y;
      out[4 * c + 2] = g_origin[2] + (int32_t)z;
      out[4 * c + 3] = (int32_t)vo_resolve_material(owner[lin], table, table_len);
    }
    c++;
  }
  return c;
}

/* tree_from_compact's filter (materials.rs:153-192): the occupancy words (n^3/32, zeroed
 * here) of the voxels inside [0, n)^3; returns the number dropped. */
uint64_t vo_compact_to_words(const int32_t *vox, uint64_t count, uint32_t n, uint32_t *words) {
  const uint64_t nn = n;
  memset(words, 0, (size_t)((nn * nn * nn + 31) / 32) * 4);
  uint64_t dropped = 0;
  for (uint64_t i = 0; i < count; i++) {
    const int64_t x = vox[4 * i], y = vox[4 * i + 1], z = vox[4 * i + 2];
    if (x < 0 || y < 0 || z < 0 || x >= (int64_t)n || y >= (int64_t)n || z >= (int64_t)n) {
      dropped++;
      continue;
    }
    const uint64_t lin = (uint64_t)x + nn * ((uint64_t)y + nn * (uint64_t)z);
    words[lin >> 5] |= 1u << (lin & 31);
  }
  return dropped;
}

/* SparseTree::from_voxels materials (sparse.rs:324-366): per leaf (leaf i has brick code
 * codes[i], ascending) 512 u16 in intra-brick Morton order (morton.rs:86-89), 0 where no
 * voxel set one; voxels in input order, so a repeated coord keeps the last; the material
 * is the low 16 bits (materials.rs:174-181). */
void vo_compact_materials(const int32_t *vox, uint64_t count, uint32_t n, const uint64_t *codes,
                          uint64_t L, uint16_t *mats) {
  memset(mats, 0, (size_t)(L * 512) * 2);
  for (uint64_t i = 0; i < count; i++) {
    const int64_t x = vox[4 * i], y = vox[4 * i + 1], z = vox[4 * i + 2];
    if (x < 0 || y < 0 || z < 0 || x >= (int64_t)n || y >= (int64_t)n || z >= (int64_t)n)
      continue;
    const uint64_t code = vo_morton_encode((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3);
    uint64_t lo = 0, hi = L;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (codes[mid] < code) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= L || codes[lo] != code) continue; /* not a leaf of this tree */
    const uint32_t m = vo_encode_brick((uint32_t)x & 7u, (uint32_t)y & 7u, (uint32_t)z & 7u);
    mats[lo * 512 + m] = (uint16_t)((uint32_t)vox[4 * i + 3] & 0xFFFFu);
  }
}

/* reference_cpu.rs:110-200 (occupancy + owner; colours from owners: vo_owner_colors) */
uint64_t vo_voxelize_surface(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                             float vs, float eps, uint32_t *words, uint32_t *owner) {
  uint64_t nn = n;
  uint64_t num_voxels = nn * nn * nn;
  memset(words, 0, (size_t)((num_voxels + 31) / 32) * 4);
  if (owner)
    for (uint64_t i = 0; i < num_voxels; i++) owner[i] = UINT32_MAX;
  const float half[3] = {0.5f, 0.5f, 0.5f};
  uint64_t pairs = 0;
  for (uint64_t t = 0; t < T; t++) {
    float g[9];
    int32_t mn[3], mx[3];
    tri_to_grid(tris + 9 * t, origin, vs, g);
    tri_box(g, n, eps, mn, mx);
    for (int32_t z = mn[2]; z <= mx[2]; z++) {
      for (int32_t y = mn[1]; y <= mx[1]; y++) {
        for (int32_t x = mn[0]; x <= mx[0]; x++) {
          pairs++;
          float c[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
          if (vo_triangle_box_overlap(c, half, g, g + 3, g + 6)) {
            uint64_t lin = (uint64_t)x + nn * ((uint64_t)y + nn * (uint64_t)z);
            words[lin >> 5] |= 1u << (lin & 31);
            if (owner && (uint32_t)t < owner[lin]) owner[lin] = (uint32_t)t;
          }
        }
      }
    }
  }
  return pairs;
}

/* voxelize_surface_cpu on a grid whose world_to_grid override is m (serial). */
uint64_t vo_voxelize_surface_mat(const float *tris, uint64_t T, uint32_t n, const float m[16],
                                 float eps, uint32_t *words, uint32_t *owner) {
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  g_mat = m;
  uint64_t pairs = vo_voxelize_surface(tris, T, n, zero, 1.0f, eps, words, owner);
  g_mat = NULL;
  return pairs;
}

/* The same scan split over threads by z (test speed only, not a baseline): thread i takes
 * the z-planes whose 8-plane row index (z >> 3) is i modulo the thread count, for every
 * triangle. Occupancy is an OR and the owner a minimum, so neither depends on the order in
 * which pairs are tested; each thread writes only its own planes (n^2 voxels, whole words
 * for n >= 8), and the pair count is the serial one. */
typedef struct {
  const float *tris;
  uint64_t T;
  uint32_t n;
  const float *origin;
  float vs, eps;
  uint32_t *words, *owner;
  uint32_t part, parts;
  uint64_t pairs;
} vox_job;

static void *vox_job_run(void *arg) {
  vox_job *j = (vox_job *)arg;
  const uint64_t nn = j->n;
  const float half[3] = {0.5f, 0.5f, 0.5f};
  uint64_t pairs = 0;
  for (uint64_t t = 0; t < j->T; t++) {
    float g[9];
    int32_t mn[3], mx[3];
    tri_to_grid(j->tris + 9 * t, j->origin, j->vs, g);
    tri_box(g, j->n, j->eps, mn, mx);
    for (int32_t z = mn[2]; z <= mx[2]; z++) {
      if ((uint32_t)(z >> 3) % j->parts != j->part) continue;
      for (int32_t y = mn[1]; y <= mx[1]; y++) {
        for (int32_t x = mn[0]; x <= mx[0]; x++) {
          pairs++;
          float c[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
          if (vo_triangle_box_overlap(c, half, g, g + 3, g + 6)) {
            uint64_t lin = (uint64_t)x + nn * ((uint64_t)y + nn * (uint64_t)z);
            j->words[lin >> 5] |= 1u << (lin & 31);
            if (j->owner && (uint32_t)t < j->owner[lin]) j->owner[lin] = (uint32_t)t;
          }
        }
      }
    }
  }
  j->pairs = pairs;
  return NULL;
}

uint64_t vo_voxelize_surface_mt(const float *tris, uint64_t T, uint32_t n, const float origin[3],
                                float vs, float eps, uint32_t *words, uint32_t *owner,
                                int threads) {
  uint64_t nn = n;
  uint64_t num_voxels = nn * nn * nn;
  uint32_t th = threads < 1 ? 1 : (uint32_t)threads;
  if (n < 8 || th == 1) return vo_voxelize_surface(tris, T, n, origin, vs, eps, words, owner);
  if (th > n / 8) th = n / 8;
  memset(words, 0, (size_t)((num_voxels + 31) / 32) * 4);
  if (owner)
    for (uint64_t i = 0; i < num_voxels; i++) owner[i] = UINT32_MAX;
  vox_job *jobs = (vox_job *)calloc(th, sizeof(vox_job));
  pthread_t *tids = (pthread_t *)calloc(th, sizeof(pthread_t));
  for (uint32_t i = 0; i < th; i++) {
    vox_job j = {tris, T, n, origin, vs, eps, words, owner, i, th, 0};
    jobs[i] = j;
    pthread_create(&tids[i], NULL, vox_job_run, &jobs[i]);
  }
  uint64_t pairs = 0;
  for (uint32_t i = 0; i < th; i++) {
    pthread_join(tids[i], NULL);
    pairs += jobs[i].pairs;
  }
  free(jobs);
  free(tids);
  return pairs;
}

/* ------------------------------------------------------------------------- */
/* sparse.rs:128-305 -- SparseTree::build; school_b.rs:68-151 -- from_sparse  */
/* ------------------------------------------------------------------------- */
typedef struct {
  uint32_t lo, hi, base;
} vnode;

struct vo_tree {
  uint32_t n, k;
  uint64_t nleaves;
  uint64_t *leaves; /* 8 words per leaf, Morton-sorted */
  uint64_t *codes;  /* brick codes (NULL when imported from School-B) */
  uint32_t *bounds;
  vnode *levels[16]; /* School-A: levels[L] for L in 2..k+1 */
  uint64_t level_count[16];
  vnode *nodes_b; /* School-B single buffer */
  uint64_t nb;
  double stage_s[2]; /* wall seconds: SparseTree::build (scan, sort, levels), from_sparse */
};

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static inline uint64_t vmask(vnode v) { return ((uint64_t)v.hi << 32) | v.lo; }

typedef struct {
  const uint32_t *words;
  uint32_t n, bpa, bz_lo, bz_hi;
  uint64_t count, cap;
  uint64_t *codes;
  uint64_t *leaves;
  int rows; /* 1: read each brick row as one byte (vo_tree_build_rows) */
} slab_job;

/* VoxelOccupancy::is_occupied, core.rs:214-253 */
static inline int occ_probe(const uint32_t *words, uint64_t n, uint32_t x, uint32_t y, uint32_t z) {
  if (x >= n || y >= n || z >= n) return 0;
  uint64_t i = (uint64_t)x + n * ((uint64_t)y + n * (uint64_t)z);
  return (words[i >> 5] >> (i & 31)) & 1;
}

static void *enumerate_slab(void *arg) { /* sparse.rs:197-225 */
  slab_job *j = (slab_job *)arg;
  for (uint32_t bz = j->bz_lo; bz < j->bz_hi; bz++)
    for (uint32_t by = 0; by < j->bpa; by++)
      for (uint32_t bx = 0; bx < j->bpa; bx++) {
        uint64_t leaf[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int any = 0;
        if (j->rows) {
          /* the 8 voxels of a brick row are 8 consecutive bits starting at a multiple of 8 */
          const uint8_t *bytes = (const uint8_t *)j->words;
          const uint64_t nn = j->n;
          for (uint32_t lz = 0; lz < 8; lz++)
            for (uint32_t ly = 0; ly < 8; ly++) {
              const uint64_t lin = (uint64_t)bx * 8 + nn * ((uint64_t)(by * 8 + ly) +
                                                            nn * (uint64_t)(bz * 8 + lz));
              const uint32_t row = bytes[lin >> 3];
              for (uint32_t lx = 0; row && lx < 8; lx++)
                if ((row >> lx) & 1u) {
                  uint32_t i = vo_encode_brick(lx, ly, lz);
                  leaf[i >> 6] |= 1ULL << (i & 63);
                  any = 1;
                }
            }
        } else {
          for (uint32_t lz = 0; lz < 8; lz++)
            for (uint32_t ly = 0; ly < 8; ly++)
              for (uint32_t lx = 0; lx < 8; lx++)
                if (occ_probe(j->words, j->n, bx * 8 + lx, by * 8 + ly, bz * 8 + lz)) {
                  uint32_t i = vo_encode_brick(lx, ly, lz);
                  leaf[i >> 6] |= 1ULL << (i & 63);
                  any = 1;
                }
        }
        if (!any) continue;
        if (j->count == j->cap) {
          j->cap = j->cap ? j->cap * 2 : 1024;
          j->codes = (uint64_t *)realloc(j->codes, j->cap * 8);
          j->leaves = (uint64_t *)realloc(j->leaves, j->cap * 64);
        }
        j->codes[j->count] = vo_morton_encode(bx, by, bz);
        memcpy(j->leaves + 8 * j->count, leaf, 64);
        j->count++;
      }
  return NULL;
}

typedef struct {
  uint64_t code;
  uint64_t idx;
} code_idx;
static int cmp_code(const void *a, const void *b) {
  uint64_t x = ((const code_idx *)a)->code, y = ((const code_idx *)b)->code;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* sparse.rs:128-144 build_parents */
static uint64_t build_parents(const uint64_t *child_codes, uint64_t nc, vnode **out_nodes,
                              uint64_t **out_codes) {
  vnode *nodes = (vnode *)malloc((nc ? nc : 1) * sizeof(vnode));
  uint64_t *pc = (uint64_t *)malloc((nc ? nc : 1) * 8);
  uint64_t i = 0, m = 0;
  while (i < nc) {
    uint64_t parent = child_codes[i] >> 6;
    uint32_t base = (uint32_t)i;
    uint64_t mask = 0;
    while (i < nc && (child_codes[i] >> 6) == parent) {
      mask |= 1ULL << (child_codes[i] & 63);
      i++;
    }
    nodes[m].lo = (uint32_t)mask;
    nodes[m].hi = (uint32_t)(mask >> 32);
    nodes[m].base = base;
    pc[m] = parent;
    m++;
  }
  *out_nodes = nodes;
  *out_codes = pc;
  return m;
}

/* school_b.rs:68-98 emit_subtree */
static void emit_subtree(vo_tree *t, uint64_t pos, uint32_t level, vnode node) {
  if (level == 2) {
    t->nodes_b[pos].base = node.base;
    return;
  }
  const vnode *children = t->levels[level - 1];
  uint32_t nch = (uint32_t)__builtin_popcountll(vmask(node));
  uint64_t block = t->nb;
  for (uint32_t j = 0; j < nch; j++) {
    vnode c = children[node.base + j];
    t->nodes_b[t->nb].lo = c.lo;
    t->nodes_b[t->nb].hi = c.hi;
    t->nodes_b[t->nb].base = 0;
    t->nb++;
  }
  t->nodes_b[pos].base = (uint32_t)block;
  for (uint32_t j = 0; j < nch; j++)
    emit_subtree(t, block + j, level - 1, children[node.base + j]);
}

static void finish_tree(vo_tree *t) {
  /* sparse.rs:151-163 build_levels */
  uint64_t nc = t->nleaves;
  uint64_t *codes = t->codes;
  for (uint32_t L = 2; L <= t->k + 1 && nc > 0; L++) {
    vnode *nodes;
    uint64_t *pcodes;
    uint64_t m = build_parents(codes, nc, &nodes, &pcodes);
    t->levels[L] = nodes;
    t->level_count[L] = m;
    if (codes != t->codes) free(codes);
    codes = pcodes;
    nc = m;
  }
  if (codes != t->codes) free(codes);
  const double t_levels = now_s();
  t->stage_s[0] = t_levels - t->stage_s[0]; /* started by tree_build */
  /* school_b.rs:113-136 */
  uint64_t total = 0;
  for (int L = 2; L < 16; L++) total += t->level_count[L];
  t->nodes_b = (vnode *)malloc((total ? total : 1) * sizeof(vnode));
  t->nb = 0;
  if (t->k >= 1 && t->nleaves > 0) {
    vnode root = t->levels[t->k + 1][0];
    t->nodes_b[0].lo = root.lo;
    t->nodes_b[0].hi = root.hi;
    t->nodes_b[0].base = 0;
    t->nb = 1;
    emit_subtree(t, 0, t->k + 1, root);
  }
  t->bounds = (uint32_t *)malloc((t->nleaves ? t->nleaves : 1) * 4);
  for (uint64_t i = 0; i < t->nleaves; i++)
    t->bounds[i] = vo_leaf_occupied_bounds(t->leaves + 8 * i);
  t->stage_s[1] = now_s() - t_levels;
}

void vo_tree_stage_seconds(const vo_tree *t, double out[2]) {
  out[0] = t->stage_s[0];
  out[1] = t->stage_s[1];
}

static vo_tree *tree_build(const uint32_t *words, uint32_t n, int threads, int rows);

vo_tree *vo_tree_build(const uint32_t *words, uint32_t n, int threads) {
  return tree_build(words, n, threads, 0);
}

/* The same tree with each brick row read as one byte of the x-major words instead of eight
 * voxel probes (an equal leaf set, faster at 2048^3; for checks, not the baseline). */
vo_tree *vo_tree_build_rows(const uint32_t *words, uint32_t n, int threads) {
  return tree_build(words, n, threads, 1);
}

static vo_tree *tree_build(const uint32_t *words, uint32_t n, int threads, int rows) {
  uint32_t k = 0;
  if (vo_resolution(n, &k, NULL, NULL) != 0) return NULL;
  vo_tree *t = (vo_tree *)calloc(1, sizeof(vo_tree));
  t->stage_s[0] = now_s();
  t->n = n;
  t->k = k;
  uint32_t bpa = n / 8;
  /* sparse.rs:248-267: threads clamped to [1, bpa], chunk = ceil(bpa/threads) */
  uint32_t th = threads < 1 ? 1 : (uint32_t)threads;
  if (th > bpa) th = bpa;
  uint32_t chunk = (bpa + th - 1) / th;
  slab_job *jobs = (slab_job *)calloc(th, sizeof(slab_job));
  pthread_t *tids = (pthread_t *)calloc(th, sizeof(pthread_t));
  for (uint32_t i = 0; i < th; i++) {
    jobs[i].words = words;
    jobs[i].n = n;
    jobs[i].bpa = bpa;
    jobs[i].bz_lo = i * chunk < bpa ? i * chunk : bpa;
    jobs[i].bz_hi = (i + 1) * chunk < bpa ? (i + 1) * chunk : bpa;
    jobs[i].rows = rows;
    if (th == 1)
      enumerate_slab(&jobs[i]);
    else
      pthread_create(&tids[i], NULL, enumerate_slab, &jobs[i]);
  }
  uint64_t total = 0;
  for (uint32_t i = 0; i < th; i++) {
    if (th > 1) pthread_join(tids[i], NULL);
    total += jobs[i].count;
  }
  /* sparse.rs:281-305 from_bricks: sort by Morton code */
  code_idx *ci = (code_idx *)malloc((total ? total : 1) * sizeof(code_idx));
  uint64_t *tmp_codes = (uint64_t *)malloc((total ? total : 1) * 8);
  uint64_t *tmp_leaves = (uint64_t *)malloc((total ? total : 1) * 64);
  uint64_t o = 0;
  for (uint32_t i = 0; i < th; i++) {
    memcpy(tmp_codes + o, jobs[i].codes, jobs[i].count * 8);
    memcpy(tmp_leaves + 8 * o, jobs[i].leaves, jobs[i].count * 64);
    o += jobs[i].count;
    free(jobs[i].codes);
    free(jobs[i].leaves);
  }
  for (uint64_t i = 0; i < total; i++) {
    ci[i].code = tmp_codes[i];
    ci[i].idx = i;
  }
  qsort(ci, total, sizeof(code_idx), cmp_code);
  t->nleaves = total;
  t->codes = (uint64_t *)malloc((total ? total : 1) * 8);
  t->leaves = (uint64_t *)malloc((total ? total : 1) * 64);
  for (uint64_t i = 0; i < total; i++) {
    t->codes[i] = ci[i].code;
    memcpy(t->leaves + 8 * i, tmp_leaves + 8 * ci[i].idx, 64);
  }
  free(ci);
  free(tmp_codes);
  free(tmp_leaves);
  free(jobs);
  free(tids);
  finish_tree(t);
  return t;
}

vo_tree *vo_tree_from_school_b(uint32_t n, const uint32_t *nodes, uint64_t node_count,
                               const uint64_t *leaves, uint64_t leaf_count,
                               const uint32_t *bounds) {
  uint32_t k = 0;
  if (vo_resolution(n, &k, NULL, NULL) != 0) return NULL;
  vo_tree *t = (vo_tree *)calloc(1, sizeof(vo_tree));
  t->n = n;
  t->k = k;
  t->nleaves = leaf_count;
  t->leaves = (uint64_t *)malloc((leaf_count ? leaf_count : 1) * 64);
  memcpy(t->leaves, leaves, leaf_count * 64);
  t->bounds = (uint32_t *)malloc((leaf_count ? leaf_count : 1) * 4);
  memcpy(t->bounds, bounds, leaf_count * 4);
  t->nodes_b = (vnode *)malloc((node_count ? node_count : 1) * sizeof(vnode));
  memcpy(t->nodes_b, nodes, node_count * 12);
  t->nb = node_count;
  return t;
}

void vo_tree_free(vo_tree *t) {
  if (!t) return;
  free(t->leaves);
  free(t->codes);
  free(t->bounds);
  for (int L = 0; L < 16; L++) free(t->levels[L]);
  free(t->nodes_b);
  free(t);
}

void vo_tree_counts(const vo_tree *t, uint64_t c[18]) {
  c[0] = t->nleaves;
  c[1] = t->nb;
  for (int L = 0; L < 16; L++) c[2 + L] = t->level_count[L];
}

void vo_tree_export(const vo_tree *t, uint32_t *nodes_b, uint64_t *leaves, uint32_t *bounds,
                    uint64_t *codes) {
  if (nodes_b) memcpy(nodes_b, t->nodes_b, t->nb * 12);
  if (leaves) memcpy(leaves, t->leaves, t->nleaves * 64);
  if (bounds) memcpy(bounds, t->bounds, t->nleaves * 4);
  if (codes && t->codes) memcpy(codes, t->codes, t->nleaves * 8);
}

void vo_tree_export_level(const vo_tree *t, uint32_t level, uint32_t *nodes_a) {
  if (level < 16 && t->levels[level]) memcpy(nodes_a, t->levels[level], t->level_count[level] * 12);
}

/* sparse.rs:614-632 is_occupied (School-A point query) */
int vo_tree_is_occupied(const vo_tree *t, uint32_t x, uint32_t y, uint32_t z) {
  if (t->nleaves == 0 || x >= t->n || y >= t->n || z >= t->n) return 0;
  uint32_t bx = x >> 3, by = y >> 3, bz = z >> 3;
  uint32_t level = t->k + 1;
  uint64_t idx = 0;
  while (level >= 2) {
    vnode nd = t->levels[level][idx];
    uint32_t sh = 2 * (level - 2);
    uint32_t bit = vo_child_bit((bx >> sh) & 3, (by >> sh) & 3, (bz >> sh) & 3);
    if (!((vmask(nd) >> bit) & 1)) return 0;
    idx = vo_child_slot(nd.lo, nd.hi, nd.base, bit);
    level--;
  }
  uint32_t i = vo_encode_brick(x & 7, y & 7, z & 7);
  return (int)((t->leaves[8 * idx + (i >> 6)] >> (i & 63)) & 1);
}

/* ------------------------------------------------------------------------- */
/* mirror.rs:22-326 -- f32 iterative explicit-stack HDDA over School-B        */
/* ------------------------------------------------------------------------- */
#define MAX_DEPTH 6 /* mirror.rs:22 */

static uint32_t floor_to_cell_f(float v, uint32_t dim) { /* mirror.rs:31-33 */
  float x = floorf(v);
  float hi = (float)(dim - 1);
  if (x < 0.0f) x = 0.0f;
  if (x > hi) x = hi;
  return sat_u32f(x);
}

typedef struct {
  uint32_t cell[3];
  int32_t step[3];
  float t_max[3], t_delta[3];
  float t_entry;
  uint32_t dim;
} fwalker;

static void fwalker_enter(fwalker *w, const float o[3], const float d[3], const float org[3],
                          uint32_t dim, float cs, float t_enter) { /* mirror.rs:48-85 */
  for (int a = 0; a < 3; a++) {
    w->cell[a] = 0;
    w->step[a] = 0;
    w->t_max[a] = INFINITY;
    w->t_delta[a] = INFINITY;
  }
  for (int a = 0; a < 3; a++) {
    float entry = o[a] + t_enter * d[a];
    float local = (entry - org[a]) / cs;
    uint32_t idx = floor_to_cell_f(local, dim);
    w->cell[a] = idx;
    if (d[a] > 0.0f) {
      w->step[a] = 1;
      float next = org[a] + ((float)idx + 1.0f) * cs;
      w->t_max[a] = t_enter + (next - entry) / d[a];
      w->t_delta[a] = cs / d[a];
    } else if (d[a] < 0.0f) {
      w->step[a] = -1;
      float next = org[a] + (float)idx * cs;
      w->t_max[a] = t_enter + (next - entry) / d[a];
      w->t_delta[a] = -cs / d[a];
    }
  }
  w->t_entry = t_enter;
  w->dim = dim;
}

static int fwalker_step(fwalker *w) { /* mirror.rs:87-116 */
  int a = 0;
  if (w->t_max[1] < w->t_max[a]) a = 1;
  if (w->t_max[2] < w->t_max[a]) a = 2;
  if (!isfinite(w->t_max[a])) return 0;
  if (w->step[a] == 1) {
    if (w->cell[a] + 1 >= w->dim) return 0;
    w->cell[a] += 1;
  } else if (w->step[a] == -1) {
    if (w->cell[a] == 0) return 0;
    w->cell[a] -= 1;
  } else {
    return 0;
  }
  w->t_entry = w->t_max[a];
  w->t_max[a] += w->t_delta[a];
  return 1;
}

typedef struct {
  uint32_t node, level;
  uint32_t origin[3];
  fwalker w;
} fframe;

static uint32_t frame_cell_size(uint32_t level) { /* mirror.rs:129-131 */
  return level == 1 ? 1u : (1u << (2 * level - 1));
}

static void make_frame(fframe *f, const float o[3], const float d[3], uint32_t node,
                       uint32_t level, const uint32_t origin[3], float t_enter) { /* :134-151 */
  float fo[3] = {(float)origin[0], (float)origin[1], (float)origin[2]};
  uint32_t dim = level == 1 ? 8 : 4;
  float cs = (float)frame_cell_size(level);
  f->node = node;
  f->level = level;
  f->origin[0] = origin[0];
  f->origin[1] = origin[1];
  f->origin[2] = origin[2];
  fwalker_enter(&f->w, o, d, fo, dim, cs, t_enter);
}

static int ascend(fframe *stack, int *sp) { /* mirror.rs:155-165 */
  for (;;) {
    *sp -= 1;
    if (*sp == 0) return 0;
    if (fwalker_step(&stack[*sp - 1].w)) return 1;
  }
}

static int ray_aabb_f32(const float o[3], const float d[3], float n, float *tn, float *tf) {
  float t_near = -INFINITY, t_far = INFINITY; /* mirror.rs:169-192 */
  for (int a = 0; a < 3; a++) {
    if (d[a] == 0.0f) {
      if (o[a] < 0.0f || o[a] > n) return 0;
    } else {
      float inv = 1.0f / d[a];
      float t1 = (0.0f - o[a]) * inv;
      float t2 = (n - o[a]) * inv;
      if (t1 > t2) {
        float s = t1;
        t1 = t2;
        t2 = s;
      }
      t_near = fmaxf(t_near, t1);
      t_far = fminf(t_far, t2);
      if (t_near > t_far) return 0;
    }
  }
  *tn = t_near;
  *tf = t_far;
  return 1;
}

static int leaf_reaches_f32(const vo_tree *t, uint32_t slot, const float o[3], const float d[3],
                            const uint32_t origin[3], float t_enter) { /* mirror.rs:202-249 */
  uint32_t b = t->bounds[slot];
  uint32_t mn[3], mx[3];
  vo_bounds_unpack(b, mn, mx);
  if (mn[0] == 0 && mn[1] == 0 && mn[2] == 0 && mx[0] == 7 && mx[1] == 7 && mx[2] == 7) return 1;
  const float m = 1.0f; /* SKIP_MARGIN, layout.rs:34 */
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (float)(origin[a] + mn[a]) - m;
    hi[a] = (float)(origin[a] + mx[a] + 1) + m;
  }
  float t_near = -1e30f, t_far = 1e30f;
  for (int a = 0; a < 3; a++) {
    if (d[a] == 0.0f) {
      if (o[a] < lo[a] || o[a] > hi[a]) return 0;
    } else {
      float inv = 1.0f / d[a];
      float t1 = (lo[a] - o[a]) * inv;
      float t2 = (hi[a] - o[a]) * inv;
      if (t1 > t2) {
        float s = t1;
        t1 = t2;
        t2 = s;
      }
      t_near = fmaxf(t_near, t1);
      t_far = fminf(t_far, t2);
    }
  }
  if (t_near > t_far) return 0;
  return t_far >= t_enter;
}

static int leaf_bit(const uint64_t *leaf, uint32_t x, uint32_t y, uint32_t z) {
  uint32_t i = vo_encode_brick(x, y, z);
  return (int)((leaf[i >> 6] >> (i & 63)) & 1);
}

/* mirror.rs:263-326 mirror_with; stats = {leaf descents, node descents, steps} */
static int mirror_one(const vo_tree *t, const float o[3], const float d[3], int skip,
                      uint32_t out[3], uint32_t stats[3]) {
  if (t->nleaves == 0) return 0;
  float t0, t1;
  if (!ray_aabb_f32(o, d, (float)t->n, &t0, &t1)) return 0;
  if (t1 < 0.0f) return 0;
  float t_entry = fmaxf(t0, 0.0f);
  uint32_t root_level = t->k == 0 ? 1 : t->k + 1;
  fframe stack[MAX_DEPTH];
  const uint32_t zero3[3] = {0, 0, 0};
  make_frame(&stack[0], o, d, 0, root_level, zero3, t_entry);
  int sp = 1;
  for (;;) {
    int top = sp - 1;
    if (stats) stats[2]++;
    if (stack[top].level == 1) {
      const uint32_t *v = stack[top].w.cell;
      if (leaf_bit(t->leaves + 8 * (uint64_t)stack[top].node, v[0], v[1], v[2])) {
        for (int a = 0; a < 3; a++) out[a] = stack[top].origin[a] + v[a];
        return 1;
      }
      if (fwalker_step(&stack[top].w)) continue;
      if (!ascend(stack, &sp)) return 0;
    } else {
      fframe fr = stack[top];
      const uint32_t *c = fr.w.cell;
      vnode node = t->nodes_b[fr.node];
      uint32_t bit = vo_child_bit(c[0], c[1], c[2]);
      uint32_t child_level = fr.level - 1;
      uint32_t size = frame_cell_size(fr.level);
      uint32_t corg[3] = {fr.origin[0] + c[0] * size, fr.origin[1] + c[1] * size,
                          fr.origin[2] + c[2] * size};
      int descend = (int)((vmask(node) >> bit) & 1);
      uint32_t slot = descend ? vo_child_slot(node.lo, node.hi, node.base, bit) : 0;
      if (skip && descend && child_level == 1)
        descend = leaf_reaches_f32(t, slot, o, d, corg, fr.w.t_entry);
      if (descend) {
        if (sp >= MAX_DEPTH) return 0; /* unreachable for k <= 5 */
        make_frame(&stack[sp], o, d, slot, child_level, corg, fr.w.t_entry);
        sp++;
        if (stats) stats[child_level == 1 ? 0 : 1]++;
      } else if (!fwalker_step(&stack[top].w) && !ascend(stack, &sp)) {
        return 0;
      }
    }
  }
}

typedef struct {
  const vo_tree *t;
  const void *rays;
  uint32_t *hits;
  double *tt;
  uint32_t *stats;
  uint64_t lo, hi;
  int skip, school;
  const uint32_t *words;
  uint32_t n;
} trav_job;

static void *mirror_job(void *arg) {
  trav_job *j = (trav_job *)arg;
  const float *rays = (const float *)j->rays;
  for (uint64_t r = j->lo; r < j->hi; r++) {
    uint32_t v[3] = {0, 0, 0};
    uint32_t *st = j->stats ? j->stats + 3 * r : NULL;
    if (st) st[0] = st[1] = st[2] = 0;
    int h = mirror_one(j->t, rays + 6 * r, rays + 6 * r + 3, j->skip, v, st);
    j->hits[4 * r + 0] = h ? v[0] : 0;
    j->hits[4 * r + 1] = h ? v[1] : 0;
    j->hits[4 * r + 2] = h ? v[2] : 0;
    j->hits[4 * r + 3] = (uint32_t)h;
  }
  return NULL;
}

static void run_jobs(void *(*fn)(void *), trav_job proto, uint64_t R, int threads) {
  int th = threads < 1 ? 1 : threads;
  if ((uint64_t)th > R) th = R ? (int)R : 1;
  trav_job *jobs = (trav_job *)calloc(th, sizeof(trav_job));
  pthread_t *tids = (pthread_t *)calloc(th, sizeof(pthread_t));
  uint64_t chunk = (R + th - 1) / th;
  for (int i = 0; i < th; i++) {
    jobs[i] = proto;
    jobs[i].lo = (uint64_t)i * chunk < R ? (uint64_t)i * chunk : R;
    jobs[i].hi = (uint64_t)(i + 1) * chunk < R ? (uint64_t)(i + 1) * chunk : R;
    if (th == 1)
      fn(&jobs[i]);
    else
      pthread_create(&tids[i], NULL, fn, &jobs[i]);
  }
  if (th > 1)
    for (int i = 0; i < th; i++) pthread_join(tids[i], NULL);
  free(jobs);
  free(tids);
}

void vo_mirror_traverse(const vo_tree *t, const float *rays, uint64_t R, uint32_t *hits,
                        int enable_skip, uint32_t *stats, int threads) {
  trav_job p;
  memset(&p, 0, sizeof(p));
  p.t = t;
  p.rays = rays;
  p.hits = hits;
  p.stats = stats;
  p.skip = enable_skip;
  run_jobs(mirror_job, p, R, threads);
}

/* ------------------------------------------------------------------------- */
/* noise.rs + fixtures.rs NoiseField -- the procedural noise fixtures, f64      */
/* ------------------------------------------------------------------------- */
static uint32_t noise_hash(uint32_t seed, int64_t ix, int64_t iy, int64_t iz) { /* noise.rs lattice_hash */
  const uint32_t x = (uint32_t)ix, y = (uint32_t)iy, z = (uint32_t)iz;
  uint32_t h = seed ^ (x * 0x8da6b343u) ^ (y * 0xd8163841u) ^ (z * 0xcb1ab31fu);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h;
}

double vo_noise_grad(uint32_t hash, double x, double y, double z) { /* noise.rs grad */
  const uint32_t h = hash & 15u;
  double u = h < 8 ? x : y;
  double v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
  if (h & 1u) u = -u;
  if (h & 2u) v = -v;
  return u + v;
}

double vo_noise_fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }

static double noise_lerp(double a, double b, double t) { return a + t * (b - a); }

double vo_perlin3(uint32_t seed, double x, double y, double z) { /* noise.rs perlin3 */
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
  const double xf = x - fx, yf = y - fy, zf = z - fz;
  const double u = vo_noise_fade(xf), v = vo_noise_fade(yf), w = vo_noise_fade(zf);
#define G(dx, dy, dz)                                                                     \
  vo_noise_grad(noise_hash(seed, ix + (dx), iy + (dy), iz + (dz)), xf - (double)(dx),      \
                yf - (double)(dy), zf - (double)(dz))
  const double x00 = noise_lerp(G(0, 0, 0), G(1, 0, 0), u);
  const double x10 = noise_lerp(G(0, 1, 0), G(1, 1, 0), u);
  const double x01 = noise_lerp(G(0, 0, 1), G(1, 0, 1), u);
  const double x11 = noise_lerp(G(0, 1, 1), G(1, 1, 1), u);
#undef G
  const double y0 = noise_lerp(x00, x10, v), y1 = noise_lerp(x01, x11, v);
  return noise_lerp(y0, y1, w);
}

static double noise_fractal(uint32_t seed, const double p[3], uint32_t octaves, double lac,
                            double gain, int ridged) { /* noise.rs fractal */
  double amp = 1.0, freq = 1.0, sum = 0.0, norm = 0.0;
  for (uint32_t o = 0; o < octaves; o++) {
    const double off = (double)o;
    const double n = vo_perlin3(seed + o * 0x9e3779b9u, p[0] * freq + off * 17.13,
                                p[1] * freq - off * 9.71, p[2] * freq + off * 5.37);
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

/* NoiseField::value (fixtures.rs:316-346): the sample point, the optional domain warp
 * (noise.rs domain_warp, at most 3 octaves), the fractal. f = {frequency, lacunarity, gain,
 * warp, threshold}, u = {seed, octaves, ridged}. */
double vo_noise_value(const double f[5], const uint32_t u[3], uint32_t n, uint32_t x, uint32_t y,
                      uint32_t z) {
  const double s = f[0] / (double)n;
  double p[3] = {((double)x + 0.5) * s, ((double)y + 0.5) * s, ((double)z + 0.5) * s};
  if (f[3] > 0.0) {
    const uint32_t wo = u[1] < 3 ? u[1] : 3;
    const double a[3] = {p[0] + 1.7, p[1] + 9.2, p[2] + 3.3};
    const double b[3] = {p[0] - 5.1, p[1] + 2.8, p[2] - 7.4};
    const double c[3] = {p[0] + 8.6, p[1] - 4.5, p[2] + 1.9};
    const double wx = noise_fractal(u[0] ^ 0x11111111u, a, wo, f[1], f[2], (int)u[2]);
    const double wy = noise_fractal(u[0] ^ 0x22222222u, b, wo, f[1], f[2], (int)u[2]);
    const double wz = noise_fractal(u[0] ^ 0x33333333u, c, wo, f[1], f[2], (int)u[2]);
    p[0] = p[0] + f[3] * wx;
    p[1] = p[1] + f[3] * wy;
    p[2] = p[2] + f[3] * wz;
  }
  return noise_fractal(u[0], p, u[1], f[1], f[2], (int)u[2]);
}

/* NoiseField::is_occupied over the grid as x-major words (zeroed here); returns the count. */
uint64_t vo_noise_words(const double f[5], const uint32_t u[3], uint32_t n, uint32_t *words) {
  const uint64_t nn = n;
  memset(words, 0, (size_t)((nn * nn * nn + 31) / 32) * 4);
  uint64_t c = 0;
  for (uint64_t z = 0; z < nn; z++)
    for (uint64_t y = 0; y < nn; y++)
      for (uint64_t x = 0; x < nn; x++)
        if (vo_noise_value(f, u, n, (uint32_t)x, (uint32_t)y, (uint32_t)z) > f[4]) {
          const uint64_t lin = x + nn * (y + nn * z);
          words[lin >> 5] |= 1u << (lin & 31);
          c++;
        }
  return c;
}

/* ------------------------------------------------------------------------- */
/* render.wgsl:54-91 -- camera rays, traversal and shading per pixel            */
/* ------------------------------------------------------------------------- */
static uint32_t unorm8(float v) { /* rgba8unorm store: clamp, scale, round to nearest */
  if (!(v > 0.0f)) v = 0.0f;
  if (v > 1.0f) v = 1.0f;
  return (uint32_t)rintf(v * 255.0f);
}

static uint32_t rgba8(float r, float g, float b, float a) {
  return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}

/* The pixel's camera ray (render.wgsl:66-70) in binary32; normalize(v) is v / sqrt(v.v)
 * with the dot product summed (x x + y y) + z z. */
void vo_camera_ray(const float cam[16], uint32_t width, uint32_t height, uint32_t px,
                   uint32_t py, float ray[6]) {
  const float *eye = cam, tanf_ = cam[3], *fwd = cam + 4, aspect = cam[7], *right = cam + 8,
              *up = cam + 12;
  const float w = (float)width, h = (float)height;
  const float ndc_x = (((float)px + 0.5f) / w * 2.0f - 1.0f) * tanf_ * aspect;
  const float ndc_y = (1.0f - ((float)py + 0.5f) / h * 2.0f) * tanf_;
  float v[3];
  for (int a = 0; a < 3; a++) v[a] = (fwd[a] + right[a] * ndc_x) + up[a] * ndc_y;
  const float len = sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  for (int a = 0; a < 3; a++) {
    ray[a] = eye[a];
    ray[3 + a] = v[a] / len;
  }
}

/* render_main (render.wgsl:57-91): per pixel the camera ray walked as the mirror walks it;
 * a hit takes its material (mats: L x 512 u16 in the leaf order of t, or NULL = none) and
 * is shaded by position (material 0: voxel / n) or by the colour table (RGBA8, R in the
 * low byte; an id past the table reads 0); a miss takes the sky gradient. */
void vo_render(const vo_tree *t, const float cam[16], uint32_t width, uint32_t height,
               const uint16_t *mats, const uint32_t *colors, uint32_t ncolors, uint32_t *out,
               int threads) {
  const uint64_t P = (uint64_t)width * height;
  float *rays = (float *)malloc((size_t)P * 24);
  uint32_t *hits = (uint32_t *)malloc((size_t)P * 16);
  for (uint32_t y = 0; y < height; y++)
    for (uint32_t x = 0; x < width; x++) vo_camera_ray(cam, width, height, x, y, rays + 6 * ((uint64_t)y * width + x));
  vo_mirror_traverse(t, rays, P, hits, 1, NULL, threads);
  const float camn = cam[11];
  for (uint64_t i = 0; i < P; i++) {
    const uint32_t y = (uint32_t)(i / width);
    const uint32_t *hh = hits + 4 * i;
    if (hh[3] == 1) {
      uint32_t mat = 0;
      if (mats && t->codes) {
        const uint64_t code = vo_morton_encode(hh[0] >> 3, hh[1] >> 3, hh[2] >> 3);
        uint64_t lo = 0, hi = t->nleaves;
        while (lo < hi) {
          const uint64_t mid = (lo + hi) / 2;
          if (t->codes[mid] < code) lo = mid + 1;
          else hi = mid;
        }
        if (lo < t->nleaves && t->codes[lo] == code)
          mat = mats[lo * 512 + vo_encode_brick(hh[0] & 7u, hh[1] & 7u, hh[2] & 7u)];
      }
      if (mat == 0)
        out[i] = rgba8((float)hh[0] / camn, (float)hh[1] / camn, (float)hh[2] / camn, 1.0f);
      else
        out[i] = mat < ncolors ? colors[mat] : 0u;
    } else {
      const float tt = (float)y / (float)height;
      out[i] = rgba8(0.08f * (1.0f - tt), 0.10f * (1.0f - tt), 0.16f + 0.12f * tt, 1.0f);
    }
  }
  free(rays);
  free(hits);
}

/* ------------------------------------------------------------------------- */
/* ray.rs:55-85, dda.rs:18-131, layout.rs:78-232 -- the f64 reference          */
/* ------------------------------------------------------------------------- */
static int ray_aabb(const double o[3], const double d[3], const double lo[3], const double hi[3],
                    double *tn, double *tf) {
  double t_near = -INFINITY, t_far = INFINITY;
  for (int a = 0; a < 3; a++) {
    if (d[a] == 0.0) {
      if (o[a] < lo[a] || o[a] > hi[a]) return 0;
    } else {
      double inv = 1.0 / d[a];
      double t1 = (lo[a] - o[a]) * inv;
      double t2 = (hi[a] - o[a]) * inv;
      if (t1 > t2) {
        double s = t1;
        t1 = t2;
        t2 = s;
      }
      t_near = fmax(t_near, t1);
      t_far = fmin(t_far, t2);
      if (t_near > t_far) return 0;
    }
  }
  *tn = t_near;
  *tf = t_far;
  return 1;
}

static uint32_t floor_to_cell_d(double v, uint32_t dim) { /* dda.rs:18-21 */
  double x = floor(v);
  double hi = (double)(dim - 1);
  if (x < 0.0) x = 0.0;
  if (x > hi) x = hi;
  return sat_u32d(x);
}

typedef struct {
  uint32_t cell[3];
  int step[3];
  double t_max[3], t_delta[3];
  uint32_t dim;
  double t_entry;
} dwalker;

static void dwalker_enter(dwalker *w, const double o[3], const double d[3], const double org[3],
                          uint32_t dim, double cs, double t_enter) { /* dda.rs:41-83 */
  for (int a = 0; a < 3; a++) {
    w->cell[a] = 0;
    w->step[a] = 0;
    w->t_max[a] = INFINITY;
    w->t_delta[a] = INFINITY;
  }
  for (int a = 0; a < 3; a++) {
    double entry = o[a] + t_enter * d[a];
    double local = (entry - org[a]) / cs;
    uint32_t idx = floor_to_cell_d(local, dim);
    w->cell[a] = idx;
    if (d[a] > 0.0) {
      w->step[a] = 1;
      double next = org[a] + ((double)idx + 1.0) * cs;
      w->t_max[a] = t_enter + (next - entry) / d[a];
      w->t_delta[a] = cs / d[a];
    } else if (d[a] < 0.0) {
      w->step[a] = -1;
      double next = org[a] + (double)idx * cs;
      w->t_max[a] = t_enter + (next - entry) / d[a];
      w->t_delta[a] = -cs / d[a];
    }
  }
  w->dim = dim;
  w->t_entry = t_enter;
}

static int dwalker_step(dwalker *w) { /* dda.rs:100-130 */
  int a = 0;
  if (w->t_max[1] < w->t_max[a]) a = 1;
  if (w->t_max[2] < w->t_max[a]) a = 2;
  if (!isfinite(w->t_max[a])) return 0;
  if (w->step[a] == 1) {
    if (w->cell[a] + 1 >= w->dim) return 0;
    w->cell[a] += 1;
  } else if (w->step[a] == -1) {
    if (w->cell[a] == 0) return 0;
    w->cell[a] -= 1;
  } else {
    return 0;
  }
  w->t_entry = w->t_max[a];
  w->t_max[a] += w->t_delta[a];
  return 1;
}

/* Cell handles: kind 0 empty, 1 leaf, 2 node (layout.rs:38-45) */
static void layout_child(const vo_tree *t, int school, uint64_t node, uint32_t level,
                         uint32_t bit, int *kind, uint64_t *idx) { /* sparse.rs:667 / school_b.rs:380 */
  vnode nd = school ? t->nodes_b[node] : t->levels[level][node];
  if (!((vmask(nd) >> bit) & 1)) {
    *kind = 0;
    return;
  }
  *idx = vo_child_slot(nd.lo, nd.hi, nd.base, bit);
  *kind = level == 2 ? 1 : 2;
}

static int leaf_chord_reaches(const vo_tree *t, uint64_t idx, const double o[3], const double d[3],
                              const uint32_t origin[3], double t_enter) { /* layout.rs:78-107 */
  uint32_t mn[3], mx[3];
  vo_bounds_unpack(t->bounds[idx], mn, mx);
  if (mn[0] == 0 && mn[1] == 0 && mn[2] == 0 && mx[0] == 7 && mx[1] == 7 && mx[2] == 7) return 1;
  double lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = (double)(origin[a] + mn[a]) - 1.0;
    hi[a] = (double)(origin[a] + mx[a] + 1) + 1.0;
  }
  double tn, tf;
  if (!ray_aabb(o, d, lo, hi, &tn, &tf)) return 0;
  return tf >= t_enter - 1e-9; /* SKIP_EPS, layout.rs:24 */
}

static int walk(const vo_tree *t, int school, const double o[3], const double d[3], int kind,
                uint64_t cell, uint32_t level, const uint32_t origin[3], double t_enter,
                uint32_t out[3], double *t_out) { /* layout.rs:151-232 */
  double fo[3] = {(double)origin[0], (double)origin[1], (double)origin[2]};
  if (kind == 0) return 0;
  if (kind == 1) {
    const uint64_t *leaf = t->leaves + 8 * cell;
    dwalker w;
    dwalker_enter(&w, o, d, fo, 8, 1.0, t_enter);
    for (;;) {
      if (leaf_bit(leaf, w.cell[0], w.cell[1], w.cell[2])) {
        for (int a = 0; a < 3; a++) out[a] = origin[a] + w.cell[a];
        *t_out = w.t_entry;
        return 1;
      }
      if (!dwalker_step(&w)) return 0;
    }
  }
  uint32_t child_size = 1u << (2 * (level - 1) + 1);
  dwalker w;
  dwalker_enter(&w, o, d, fo, 4, (double)child_size, t_enter);
  for (;;) {
    int ck;
    uint64_t ci = 0;
    layout_child(t, school, cell, level, vo_child_bit(w.cell[0], w.cell[1], w.cell[2]), &ck, &ci);
    if (ck != 0) {
      uint32_t corg[3] = {origin[0] + w.cell[0] * child_size, origin[1] + w.cell[1] * child_size,
                          origin[2] + w.cell[2] * child_size};
      int descend = ck == 1 ? leaf_chord_reaches(t, ci, o, d, corg, w.t_entry) : 1;
      if (descend && walk(t, school, o, d, ck, ci, level - 1, corg, w.t_entry, out, t_out))
        return 1;
    }
    if (!dwalker_step(&w)) return 0;
  }
}

static int layout_one(const vo_tree *t, int school, const double o[3], const double d[3],
                      uint32_t out[3], double *t_out) { /* layout.rs:126-147 */
  const double lo[3] = {0.0, 0.0, 0.0};
  const double hi[3] = {(double)t->n, (double)t->n, (double)t->n};
  double t0, t1;
  if (!ray_aabb(o, d, lo, hi, &t0, &t1)) return 0;
  if (t1 < 0.0) return 0;
  if (t->nleaves == 0) return 0;
  int kind = t->k == 0 ? 1 : 2;
  uint32_t level = t->k == 0 ? 1 : t->k + 1;
  const uint32_t z3[3] = {0, 0, 0};
  return walk(t, school, o, d, kind, 0, level, z3, fmax(t0, 0.0), out, t_out);
}

static void *layout_job(void *arg) {
  trav_job *j = (trav_job *)arg;
  const double *rays = (const double *)j->rays;
  for (uint64_t r = j->lo; r < j->hi; r++) {
    uint32_t v[3] = {0, 0, 0};
    double tt = 0.0;
    int h = layout_one(j->t, j->school, rays + 6 * r, rays + 6 * r + 3, v, &tt);
    j->hits[4 * r + 0] = h ? v[0] : 0;
    j->hits[4 * r + 1] = h ? v[1] : 0;
    j->hits[4 * r + 2] = h ? v[2] : 0;
    j->hits[4 * r + 3] = (uint32_t)h;
    if (j->tt) j->tt[r] = h ? tt : 0.0;
  }
  return NULL;
}

void vo_layout_traverse(const vo_tree *t, const double *rays, uint64_t R, uint32_t *hits,
                        double *t_enter, int school, int threads) {
  trav_job p;
  memset(&p, 0, sizeof(p));
  p.t = t;
  p.rays = rays;
  p.hits = hits;
  p.tt = t_enter;
  p.school = school ? 1 : 0;
  if (!school && t->codes == NULL) p.school = 1; /* imported trees only carry School-B */
  run_jobs(layout_job, p, R, threads);
}

/* oracle.rs:46-131 dense tier-A first hit */
void vo_first_hit_dense(const uint32_t *words, uint32_t n, const double *rays, uint64_t R,
                        uint32_t *hits, double *t_enter) {
  for (uint64_t r = 0; r < R; r++) {
    const double *o = rays + 6 * r, *d = rays + 6 * r + 3;
    uint32_t *h = hits + 4 * r;
    h[0] = h[1] = h[2] = h[3] = 0;
    if (t_enter) t_enter[r] = 0.0;
    const double lo[3] = {0, 0, 0}, hi[3] = {(double)n, (double)n, (double)n};
    double t0, t1;
    if (!ray_aabb(o, d, lo, hi, &t0, &t1) || t1 < 0.0) continue;
    double te = fmax(t0, 0.0);
    double entry[3];
    uint32_t v[3];
    int step[3] = {0, 0, 0};
    double tmax[3] = {INFINITY, INFINITY, INFINITY}, tdel[3] = {INFINITY, INFINITY, INFINITY};
    for (int a = 0; a < 3; a++) {
      entry[a] = o[a] + te * d[a];
      v[a] = floor_to_cell_d(entry[a], n);
    }
    for (int a = 0; a < 3; a++) {
      if (d[a] > 0.0) {
        step[a] = 1;
        tmax[a] = te + ((double)(v[a] + 1) - entry[a]) / d[a];
        tdel[a] = 1.0 / d[a];
      } else if (d[a] < 0.0) {
        step[a] = -1;
        tmax[a] = te + ((double)v[a] - entry[a]) / d[a];
        tdel[a] = -1.0 / d[a];
      }
    }
    double tc = te;
    for (;;) {
      if (occ_probe(words, n, v[0], v[1], v[2])) {
        h[0] = v[0]; h[1] = v[1]; h[2] = v[2]; h[3] = 1;
        if (t_enter) t_enter[r] = tc;
        break;
      }
      int a = 0;
      if (tmax[1] < tmax[a]) a = 1;
      if (tmax[2] < tmax[a]) a = 2;
      if (!isfinite(tmax[a])) break;
      if (step[a] == 1) {
        if (v[a] + 1 >= n) break;
        v[a]++;
      } else {
        if (v[a] == 0) break;
        v[a]--;
      }
      tc = tmax[a];
      tmax[a] += tdel[a];
    }
  }
}

/* ------------------------------------------------------------------------- */
/* fixtures.rs:12-237                                                        */
/* ------------------------------------------------------------------------- */
static uint32_t hash3(uint32_t x, uint32_t y, uint32_t z) { /* fixtures.rs:220-227 */
  uint32_t h = (x * 0x8da6b343u) ^ (y * 0xd8163841u) ^ (z * 0xcb1ab31fu);
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  return h;
}

static int fixture_occ(int kind, uint32_t n, uint32_t p0, uint32_t p1, uint32_t p2, uint32_t x,
                       uint32_t y, uint32_t z) {
  switch (kind) {
  case 0: return 0;
  case 1: return 1;
  case 2: return x == p0 && y == p1 && z == p2;
  case 3: return ((x + y + z) % 2) == 0;
  case 4: {
    uint32_t levels = (uint32_t)__builtin_ctz(n);
    for (uint32_t b = 0; b < levels; b++) {
      uint32_t oct = ((x >> b) & 1) | (((y >> b) & 1) << 1) | (((z >> b) & 1) << 2);
      if (((p0 >> oct) & 1) == 0) return 0;
    }
    return 1;
  }
  case 5: {
    int nx = x % p0 < p1, ny = y % p0 < p1, nz = z % p0 < p1;
    return nx + ny + nz >= 2;
  }
  case 6: return hash3(x, y, z) % p0 == 0;
  default: return 0;
  }
}

void vo_fixture_words(int kind, uint32_t n, uint32_t p0, uint32_t p1, uint32_t p2,
                      uint32_t *words) {
  uint64_t nn = n, nv = nn * nn * nn;
  memset(words, 0, (size_t)((nv + 31) / 32) * 4);
  for (uint32_t z = 0; z < n; z++)
    for (uint32_t y = 0; y < n; y++)
      for (uint32_t x = 0; x < n; x++)
        if (fixture_occ(kind, n, p0, p1, p2, x, y, z)) {
          uint64_t i = (uint64_t)x + nn * ((uint64_t)y + nn * z);
          words[i >> 5] |= 1u << (i & 31);
        }
}
