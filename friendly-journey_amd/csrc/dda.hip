// Hierarchical DDA over the School-B tree: one lane per ray, hit-for-hit identical to
// the f32 iterative mirror (voxel-core/src/mirror.rs:254-326).
//
// Every floating-point operation is the mirror's, in its order, in binary32, with no
// FMA contraction (-ffp-contract=off) and IEEE division. Two exact rewrites:
//  * (entry - origin) / cell_size with cell_size = 2^e is computed as a multiply by
//    2^-e: both are the correctly rounded value of the same real number;
//  * cell_size / d is computed as cell_size * (1/d) when |d| < 2^100: scaling by a
//    power of two commutes with rounding unless 1/d is subnormal, which that bound
//    excludes. Otherwise the division is done as written.
//
// A frame's cells are tested against a 64-bit mask held in registers: a node frame's child
// mask, or, in a leaf frame, the leaf's u64 word of the 4^3 sub-brick the cell is in (the 6 low
// bits of the intra-brick Morton index are the child_bit formula of the cell's low two
// coordinate bits; the 3 high bits pick the word), reloaded only when a step crosses into
// another sub-brick. So a leaf is read once per sub-brick visited, not once per cell.
//
// Frame storage: the active frame lives in registers. Each ancestor frame is kept in an
// LDS row named by its level (the stack index of a frame is root_level - level, so "the
// frame at level L" is unique at any time), one column per lane.
//
// The hit's ray parameter t is the entry parameter of the f64 slab of the hit voxel for
// the ray widened to f64: max(0, max_a min((v_a - o_a)/d_a, (v_a + 1 - o_a)/d_a)) over
// axes with d_a != 0, which is the t_enter the f64 reference traversal reports for that
// voxel (voxel-core/src/layout.rs:176-183).
#include "vxl_internal.hpp"

#include <atomic>

// Work counters of vxl_traverse_dev_noalloc (the entry point for callers that cannot
// allocate): kNoAllocSlots 8-byte counters, one per persistent traversal in flight, used in
// turn. Each GPU the library runs on holds its own copy until the process ends. Host code
// never passes this name to a kernel: it is a host-side stand-in, so each device's address is
// looked up once with hipGetSymbolAddress (noalloc_setup) and only that address is used.
constexpr uint32_t kNoAllocSlots = 64;
static __device__ __align__(256) unsigned long long scratch[kNoAllocSlots];

namespace vxl {
namespace {

constexpr uint32_t kFull = (7u << 9) | (7u << 12) | (7u << 15);
// Descents one walk may make (a safety bound a legitimate walk never comes close to: a
// frame steps at most 3 (dim - 1) times and is entered only by a descent, so this bounds
// every walk's visits; checked at descents, not at every visit).
constexpr uint32_t kMaxDescents = 1u << 18;

// A suspended ancestor frame lives in LDS, one row per (level, field), one column per
// thread: node, packed cell, t_max x 3. Its t_delta is not kept (a function of the frame's
// cell size and the ray, recomputed with the same operations on restore), nor its t_entry
// (a restored frame steps at once, which overwrites it, or the walk ends).
constexpr int kFrameWords = 5;

__device__ __forceinline__ float axis_delta(float cs, float d, float rinv, bool big) {
  float q = cs * rinv;
  if (__builtin_expect(big, 0)) {
    // the division only for |d| >= 2^100; the empty volatile asm keeps the compiler from
    // computing it on every lane and selecting
    asm volatile("");
    q = cs / d;
  }
  return d > 0.0f ? q : (d < 0.0f ? -q : INFINITY);
}

__device__ __forceinline__ uint32_t floor_to_cell(float v, uint32_t dim) {
  // mirror.rs:31-33: v.floor().clamp(0, dim-1) as u32 (NaN passes the clamp -> 0)
  float x = floorf(v);
  const float hi = float(dim - 1);
  if (x < 0.0f) x = 0.0f;
  if (x > hi) x = hi;
  return (x == x) ? uint32_t(x) : 0u;
}

struct Ray {
  float o[3], d[3], rinv[3];
  bool big[3];
  uint32_t rsel;  // per axis (a byte each, x low): the row of the ray's direction in the
                  // range-mask table (below)
};

// Walker::enter (mirror.rs:48-85) for one axis. The d > 0 and d < 0 cases differ only in
// which cell face is next (idx + 1 or idx), so one division serves both; a lane with
// d == 0 (or NaN) keeps t_max = t_delta = +inf.
__device__ __forceinline__ void axis_enter(const Ray &r, int a, float org, uint32_t dim, float cs,
                                           float inv_cs, float t_enter, uint32_t &cell,
                                           float &tmax, float &tdel) {
  const float o = r.o[a], d = r.d[a];
  const float entry = o + t_enter * d;
  const float local = (entry - org) * inv_cs;
  const uint32_t idx = floor_to_cell(local, dim);
  cell = idx;
  const float face = d > 0.0f ? float(idx) + 1.0f : float(idx);
  const float next = org + face * cs;
  const float tm = t_enter + (next - entry) / d;
  tmax = (d > 0.0f || d < 0.0f) ? tm : INFINITY;  // a NaN component keeps +inf, as in enter
  tdel = axis_delta(cs, d, r.rinv[a], r.big[a]);
}

__device__ __forceinline__ uint32_t frame_cell_size(uint32_t level) {
  return level == 1 ? 1u : (1u << (2 * level - 1));
}

// leaf_reaches_f32 (mirror.rs:202-249), with 1/d precomputed (same value per call).
// Each axis is evaluated branch-free (selects instead of the d == 0 branch), with the same
// values: for d == 0 the products are computed and discarded.
__device__ __forceinline__ bool leaf_reaches(uint32_t packed, const Ray &r, uint32_t ox,
                                             uint32_t oy, uint32_t oz, float t_enter) {
  if (packed == kFull) return true;
  const uint32_t mn[3] = {packed & 7u, (packed >> 3) & 7u, (packed >> 6) & 7u};
  const uint32_t mx[3] = {(packed >> 9) & 7u, (packed >> 12) & 7u, (packed >> 15) & 7u};
  const uint32_t org[3] = {ox, oy, oz};
  float t_near = -1e30f, t_far = 1e30f;
  bool miss = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float lo = float(org[a] + mn[a]) - 1.0f;
    const float hi = float(org[a] + mx[a] + 1u) + 1.0f;
    const bool flat = r.d[a] == 0.0f;
    miss = miss || (flat && (r.o[a] < lo || r.o[a] > hi));
    const float t1 = (lo - r.o[a]) * r.rinv[a];
    const float t2 = (hi - r.o[a]) * r.rinv[a];
    const bool sw = t1 > t2;
    const float tn = sw ? t2 : t1, tf = sw ? t1 : t2;
    t_near = flat ? t_near : fmaxf(t_near, tn);
    t_far = flat ? t_far : fminf(t_far, tf);
  }
  if (miss || t_near > t_far) return false;
  return t_far >= t_enter;
}

// The walk of one lane: the active frame (level, node, its origin, cell, t_max, t_delta,
// t_entry) and its 64-bit scan mask: for a node frame the node's child mask with child_base
// in `cbase`, for a leaf frame the leaf word of the current 4^3 sub-brick with that
// sub-brick's index in `cbase`. Ancestor frames are in LDS (above).
struct Walk {
  uint32_t level, node, ox, oy, oz, cx, cy, cz;
  float tm0, tm1, tm2, td0, td1, td2, te;
  uint32_t mlo, mhi, cbase;
  uint32_t root_level, iter;
  uint32_t lbad;  // leaf frame: per axis (4 bits each, x low) 1 + the first cell past the
                  // leaf's occupied bounds in the ray's direction (15: the axis never steps)
};

// Cells of a 4^3 node frame (64 bits in child_bit order) whose coordinate along axis a is
// >= c (s = 0) or <= c (s = 1); table row a * 8 + s * 4 + c.
constexpr unsigned long long range_mask(int a, int s, int c) {
  unsigned long long m = 0;
  for (int b = 0; b < 64; b++) {
    const int coord = a == 0 ? ((b & 1) | ((b >> 2) & 2))
                             : (a == 1 ? (((b >> 1) & 1) | ((b >> 3) & 2))
                                       : (((b >> 2) & 1) | ((b >> 4) & 2)));
    if (s ? coord <= c : coord >= c) m |= 1ull << b;
  }
  return m;
}
#define VXL_RANGE4(a, s) range_mask(a, s, 0), range_mask(a, s, 1), range_mask(a, s, 2), \
                         range_mask(a, s, 3)
__constant__ unsigned long long kRangeMasks[24] = {VXL_RANGE4(0, 0), VXL_RANGE4(0, 1),
                                                   VXL_RANGE4(1, 0), VXL_RANGE4(1, 1),
                                                   VXL_RANGE4(2, 0), VXL_RANGE4(2, 1)};
#undef VXL_RANGE4
static_assert(range_mask(0, 0, 1) == 0xFFAAFFAAFFAAFFAAull, "x >= 1: x bit 0 or x bit 1");
static_assert(range_mask(2, 1, 1) == 0x00000000FFFFFFFFull, "z <= 1: z bit 1 clear");

// Every block of a walking kernel copies the table into LDS before any thread leaves.
__device__ __forceinline__ void load_range_masks(unsigned long long *rmask) {
  if (threadIdx.x < 24) rmask[threadIdx.x] = kRangeMasks[threadIdx.x];
  __syncthreads();
}

// 1/d, the |d| >= 2^100 flags and the range-mask rows of a ray whose o and d are set. The
// row follows the sign bit: a ray that moves towards +axis only reaches cells >= its own
// along it, one towards -axis cells <= its own; d = +-0 or NaN never steps along the axis,
// so either row holds the cells it reaches.
__device__ __forceinline__ void finish_ray(Ray &r) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    r.rinv[a] = 1.0f / r.d[a];
    r.big[a] = !(fabsf(r.d[a]) < 0x1p100f);
  }
  r.rsel = ((__float_as_uint(r.d[0]) >> 31) * 4) |
           ((8 + (__float_as_uint(r.d[1]) >> 31) * 4) << 8) |
           ((16 + (__float_as_uint(r.d[2]) >> 31) * 4) << 16);
}

__device__ __forceinline__ void load_ray(const float *__restrict__ rays, uint64_t ri, Ray &r) {
  const float2 a = *reinterpret_cast<const float2 *>(rays + 6 * ri);
  const float2 b = *reinterpret_cast<const float2 *>(rays + 6 * ri + 2);
  const float2 c = *reinterpret_cast<const float2 *>(rays + 6 * ri + 4);
  r.o[0] = a.x; r.o[1] = a.y; r.o[2] = b.x;
  r.d[0] = b.y; r.d[1] = c.x; r.d[2] = c.y;
  finish_ray(r);
}

// Grid clip, ray_aabb_f32 (mirror.rs:169-192) + the t1 < 0 check (:273-277); false when
// the ray misses the grid or the tree is empty, else t0 is the entry parameter. Branch-free:
// the slab interval only narrows, so "empty after some axis" (the early return) is "empty
// after all three"; an axis with d == 0 tests the origin instead and leaves the interval.
__device__ __forceinline__ bool clip_grid(const Ray &r, uint32_t n, uint64_t leaf_count,
                                          float &t0_out) {
  const float nf = float(n);
  float t0 = -INFINITY, t1 = INFINITY;
  bool missed = leaf_count == 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const bool flat = r.d[a] == 0.0f;
    missed = missed || (flat && (r.o[a] < 0.0f || r.o[a] > nf));
    const float ta = (0.0f - r.o[a]) * r.rinv[a];
    const float tb = (nf - r.o[a]) * r.rinv[a];
    const bool sw = ta > tb;
    t0 = flat ? t0 : fmaxf(t0, sw ? tb : ta);
    t1 = flat ? t1 : fminf(t1, sw ? ta : tb);
  }
  missed = missed || t0 > t1 || t1 < 0.0f;
  t0_out = t0;
  return !missed;
}

// Walker::enter (mirror.rs:48-85) of a frame at `level` with origin (ox, oy, oz).
__device__ __forceinline__ void make_frame(const Ray &r, Walk &w, uint32_t level, float t_enter) {
  const uint32_t dim = level == 1 ? 8u : 4u;
  const uint32_t csu = frame_cell_size(level);
  const float cs = float(csu);
  const float ics = __uint_as_float((127u - (31u - __clz(csu))) << 23);  // 2^-e
  axis_enter(r, 0, float(w.ox), dim, cs, ics, t_enter, w.cx, w.tm0, w.td0);
  axis_enter(r, 1, float(w.oy), dim, cs, ics, t_enter, w.cy, w.tm1, w.td1);
  axis_enter(r, 2, float(w.oz), dim, cs, ics, t_enter, w.cz, w.tm2, w.td2);
  w.te = t_enter;
}

// Walker::step (mirror.rs:87-116) on the active frame; false when exhausted. The axis
// choice (lowest t_max, ties to the lowest axis) and the update are selects, so lanes
// stepping along different axes do not split the wave.
// The fields are read into locals first: a conditional expression over struct fields would
// become a load from a selected address, which keeps the walk state out of registers.
template <bool kSkip>
__device__ __forceinline__ bool step_frame(const Ray &r, Walk &w, bool &past) {
  const uint32_t dim = w.level == 1 ? 8u : 4u;
  const float tm0 = w.tm0, tm1 = w.tm1, tm2 = w.tm2;
  const float d0 = r.d[0], d1 = r.d[1], d2 = r.d[2];
  const uint32_t cx = w.cx, cy = w.cy, cz = w.cz;
  const bool a1 = tm1 < tm0;
  float best = a1 ? tm1 : tm0;
  const bool a2 = tm2 < best;
  best = a2 ? tm2 : best;
  const bool on0 = !a1 && !a2, on1 = a1 && !a2;
  const float da = a2 ? d2 : (a1 ? d1 : d0);
  const uint32_t ca = a2 ? cz : (a1 ? cy : cx);
  const bool ok = isfinite(best) && ((da > 0.0f && ca + 1 < dim) || (da < 0.0f && ca != 0));
  if (ok) {
    const uint32_t inc = da > 0.0f ? 1u : 0xFFFFFFFFu;
    const float td0 = w.td0, td1 = w.td1, td2 = w.td2;
    w.cx = cx + (on0 ? inc : 0u);
    w.cy = cy + (on1 ? inc : 0u);
    w.cz = cz + (a2 ? inc : 0u);
    w.tm0 = on0 ? tm0 + td0 : tm0;
    w.tm1 = on1 ? tm1 + td1 : tm1;
    w.tm2 = a2 ? tm2 + td2 : tm2;
    w.te = best;
    // the stepped axis's new cell against the leaf's bound on it: past it, every later cell
    // of the frame is too (the walk is monotone along each axis)
    if (kSkip) {
      const uint32_t sh = a2 ? 8u : (a1 ? 4u : 0u);
      past = w.level == 1 && ((w.lbad >> sh) & 15u) == ca + inc + 1u;
    }
  }
  return ok;
}

// The leaf frame just entered: lbad from the leaf's packed bounds (LeafBounds), and whether
// its entry cell already lies past them on some axis (then no cell of the frame is occupied).
__device__ __forceinline__ bool leaf_limits(const Ray &r, Walk &w, uint32_t packed) {
  const uint32_t c[3] = {w.cx, w.cy, w.cz};
  uint32_t bad = 0;
  bool past = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const uint32_t mn = (packed >> (3 * a)) & 7u, mx = (packed >> (9 + 3 * a)) & 7u;
    const float d = r.d[a];
    const uint32_t f = d > 0.0f ? mx + 2u : (d < 0.0f ? mn : 15u);
    bad |= f << (4 * a);
    past = past || (d > 0.0f ? c[a] > mx : (d < 0.0f ? c[a] < mn : (c[a] < mn || c[a] > mx)));
  }
  w.lbad = bad;
  return past;
}

// The 4^3 sub-brick of a leaf-frame cell: the high bits of the intra-brick Morton index.
__device__ __forceinline__ uint32_t sub_brick(uint32_t cx, uint32_t cy, uint32_t cz) {
  return (cx >> 2) | ((cy >> 2) << 1) | ((cz >> 2) << 2);
}

// The scan mask of the active frame: a node's record, or a leaf's sub-brick word.
__device__ __forceinline__ void load_node(Walk &w, const uint32_t *__restrict__ nodes) {
  w.mlo = nodes[3ull * w.node];
  w.mhi = nodes[3ull * w.node + 1];
  w.cbase = nodes[3ull * w.node + 2];
}
__device__ __forceinline__ void load_leaf_word(Walk &w, const uint32_t *__restrict__ leafw) {
  w.cbase = sub_brick(w.cx, w.cy, w.cz);
  const uint2 v = *reinterpret_cast<const uint2 *>(leafw + (uint64_t(w.node) * 8 + w.cbase) * 2);
  w.mlo = v.x;
  w.mhi = v.y;
}

// The root frame of a ray that entered the grid at t0 (mirror.rs:278-283).
__device__ __forceinline__ void begin_walk(const Ray &r, Walk &w, uint32_t k, float t0,
                                           const uint32_t *__restrict__ nodes,
                                           const uint32_t *__restrict__ leafw) {
  w.root_level = k == 0 ? 1u : k + 1;
  w.level = w.root_level;
  w.node = 0;
  w.ox = w.oy = w.oz = 0;
  w.iter = 0;
  make_frame(r, w, w.level, fmaxf(t0, 0.0f));
  w.lbad = 0xFFFu;
  if (w.level >= 2) load_node(w, nodes);
  else load_leaf_word(w, leafw);
}

constexpr uint32_t kWalking = 0, kHit = 1, kLeft = 2, kCapped = 3;

// Bit of Walk::iter set while the active frame was just restored by an ascent: its next
// visit steps it without testing its cell again (ascend, mirror.rs:155-165).
constexpr uint32_t kPendingStep = 1u << 31;
// Bit of Walk::iter set while the active leaf frame was entered at a cell past the leaf's
// occupied bounds: its next visit pops it without testing or stepping.
constexpr uint32_t kLeafPast = 1u << 30;

// One iteration of the walk (mirror.rs:284-325): test the active cell against the frame's
// scan mask (a node's child mask, or the leaf word of the cell's sub-brick, both in
// registers); on a hit voxel stop; on an occupied child whose leaf the ray reaches descend;
// else step, reloading the leaf word when the step crosses into another sub-brick. A frame
// that cannot step is popped, one level per iteration: the parent is restored and steps at
// its next iteration, as the mirror's ascend steps the parent. Popping one level per
// iteration (rather than looping until some ancestor steps) keeps each iteration's code
// short and lets the other lanes of the wave walk on while one climbs several levels.
// With kSkip a leaf frame is also popped as soon as the walk leaves the leaf's occupied
// bounds (LeafBounds) along the axis it steps on, or enters past them: every cell it would
// still visit is empty, so the mirror would only step through them and pop the frame at its
// end; the hit is the same, with fewer visits and leaf-word loads. It pays where lanes wait
// on leaf loads from HBM (the persistent-lane walk of trees beyond the Infinity Cache); on
// cache-resident trees the extra instructions of every step cost more than the visits saved.
// kSkip also tests the cell a step lands on within the same iteration and, when it is empty,
// steps once more, so a lane crossing empty cells needs fewer iterations (soup_2048: -2.5%;
// a third step per iteration, or the second on the cache-resident sphere, costs more than
// it saves).
// After a step, a node frame whose child mask holds no cell in the box still ahead of the ray
// pops at once (the cells on the ray's side of the new cell along each axis, from the
// range-mask table `rmask` in LDS: the walk is monotone along each axis, so every cell the
// mirror would still visit in the frame is in that box and empty). The walk's hit is the
// mirror's; the empty visits are skipped (terrain_1024's tree: 4.40 -> 3.96 ms with one lane
// per ray, 3.88 -> 3.49 ms on persistent lanes; sphere_256 -2%, soup_2048 -2%).
// `frames` is this lane's column of the LDS frame rows, `fstride` the row pitch. kStats
// counts the descents into leaves and into nodes and the visits (cell tests; the mirror's
// walk statistics), so it walks every cell as the mirror does.
template <bool kStats = false, bool kSkip = false>
__device__ __forceinline__ uint32_t walk_visit(const Ray &r, Walk &w,
                                               const uint32_t *__restrict__ nodes,
                                               const uint32_t *__restrict__ leafw,
                                               const uint32_t *__restrict__ bounds,
                                               uint32_t *frames, uint32_t fstride,
                                               const unsigned long long *rmask,
                                               uint32_t *stats = nullptr) {
  if (!(w.iter & (kSkip ? kPendingStep | kLeafPast : kPendingStep))) {
    if (kStats) stats[2] += 1;
    const uint32_t mlo = w.mlo, mhi = w.mhi;
    const uint32_t bit = dev::child_bit(w.cx & 3u, w.cy & 3u, w.cz & 3u);
    const bool has = ((bit < 32 ? mlo : mhi) >> (bit & 31u)) & 1u;
    if (has && w.level == 1) return kHit;
    if (has) {
      const uint32_t size = frame_cell_size(w.level);
      const uint32_t cox = w.ox + w.cx * size, coy = w.oy + w.cy * size, coz = w.oz + w.cz * size;
      const uint32_t below = bit < 32 ? __popc(mlo & ((1u << bit) - 1u))
                                      : __popc(mlo) + __popc(mhi & ((1u << (bit - 32)) - 1u));
      const uint32_t slot = w.cbase + below;
      const uint32_t lb = w.level == 2 ? bounds[slot] : kFull;
      const bool descend = w.level != 2 || leaf_reaches(lb, r, cox, coy, coz, w.te);
      if (descend) {
        if (++w.iter > kMaxDescents) return kCapped;
        {
          uint32_t *f = frames + (w.level - 2) * kFrameWords * fstride;
          f[0] = w.node;
          f[fstride] = w.cx | (w.cy << 8) | (w.cz << 16);
          f[2 * fstride] = __float_as_uint(w.tm0);
          f[3 * fstride] = __float_as_uint(w.tm1);
          f[4 * fstride] = __float_as_uint(w.tm2);
        }
        w.level -= 1;
        if (kStats) stats[w.level == 1 ? 0 : 1] += 1;
        w.node = slot;
        w.ox = cox;
        w.oy = coy;
        w.oz = coz;
        make_frame(r, w, w.level, w.te);
        if (w.level >= 2) {
          load_node(w, nodes);
        } else if (kStats || !kSkip) {
          w.lbad = 0xFFFu;
          load_leaf_word(w, leafw);
        } else if (leaf_limits(r, w, lb)) {
          w.iter |= kLeafPast;
        } else {
          load_leaf_word(w, leafw);
        }
        return kWalking;
      }
    }
  }
  w.iter &= ~kPendingStep;
  bool past = false;
  if (!(kSkip && (w.iter & kLeafPast)) && step_frame<kSkip>(r, w, past) && !past) {
    if (w.level == 1 && sub_brick(w.cx, w.cy, w.cz) != w.cbase) {
      load_leaf_word(w, leafw);
      return kWalking;
    }
    // a node frame whose child mask holds nothing among the cells still ahead of the ray
    // (those on the ray's side of the new cell along each axis, the cell included) pops now:
    // the mirror would step through them, all empty, and pop at the frame's end
    bool none_ahead = false;
    if (!kStats && w.level >= 2) {
      const uint32_t sel = r.rsel;
      const unsigned long long m = rmask[(sel & 0xffu) + w.cx] &
                                   rmask[((sel >> 8) & 0xffu) + w.cy] & rmask[(sel >> 16) + w.cz];
      none_ahead = ((w.mlo & uint32_t(m)) | (w.mhi & uint32_t(m >> 32))) == 0;
    }
    if (!none_ahead) {
    if (kSkip && !kStats) {
      // a second visit in the same iteration when the new cell is empty: step again. What
      // this second step cannot finish is left to the next visit through the flags: an
      // exhausted frame (kPendingStep: the step fails again and the frame pops), a step past
      // the leaf's bounds (kLeafPast: pop); an occupied cell is tested again there.
      const uint32_t b2 = dev::child_bit(w.cx & 3u, w.cy & 3u, w.cz & 3u);
      if (!(((b2 < 32 ? w.mlo : w.mhi) >> (b2 & 31u)) & 1u)) {
        bool past2 = false;
        if (!step_frame<kSkip>(r, w, past2)) {
          w.iter |= kPendingStep;
        } else if (past2) {
          w.iter |= kLeafPast;
        } else if (w.level == 1 && sub_brick(w.cx, w.cy, w.cz) != w.cbase) {
          load_leaf_word(w, leafw);
        }
      }
    }
    return kWalking;
    }
  }
  if (kSkip) w.iter &= ~kLeafPast;
  if (w.level == w.root_level) return kLeft;
  w.level += 1;
  const uint32_t size = frame_cell_size(w.level);
  {
    const uint32_t *f = frames + (w.level - 2) * kFrameWords * fstride;
    w.node = f[0];
    const uint32_t cell = f[fstride];
    w.cx = cell & 0xffu;
    w.cy = (cell >> 8) & 0xffu;
    w.cz = cell >> 16;
    w.tm0 = __uint_as_float(f[2 * fstride]);
    w.tm1 = __uint_as_float(f[3 * fstride]);
    w.tm2 = __uint_as_float(f[4 * fstride]);
    const float cs = float(size);
    w.td0 = axis_delta(cs, r.d[0], r.rinv[0], r.big[0]);
    w.td1 = axis_delta(cs, r.d[1], r.rinv[1], r.big[1]);
    w.td2 = axis_delta(cs, r.d[2], r.rinv[2], r.big[2]);
  }
  w.ox -= w.cx * size;
  w.oy -= w.cy * size;
  w.oz -= w.cz * size;
  load_node(w, nodes);  // a restored frame is a node frame
  w.iter |= kPendingStep;
  return kWalking;
}

// t of a hit at voxel v for the ray (o, d) widened to f64 (see the file comment).
__device__ __forceinline__ double hit_t(const float o[3], const float d[3], uint4 v) {
  const double vv[3] = {double(v.x), double(v.y), double(v.z)};
  double tn = 0.0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (d[a] != 0.0f) {
      const double oa = double(o[a]);
      const double inv = 1.0 / double(d[a]);
      const double ta = (vv[a] - oa) * inv, tb = (vv[a] + 1.0 - oa) * inv;
      tn = fmax(tn, fmin(ta, tb));
    }
  }
  return tn;
}

// Writes ray ri's result: (x, y, z, 1) for a hit, (0, 0, 0, 0) for none, (0, 0, 0, ~0) for
// a walk over the descent bound; and t when tout is given.
__device__ __forceinline__ void store_result(uint64_t ri, const Ray &r, const Walk &w,
                                             uint32_t status, uint4 *__restrict__ hits,
                                             double *__restrict__ tout) {
  const bool hit = status == kHit;
  uint4 result = make_uint4(0u, 0u, 0u, status == kCapped ? 0xFFFFFFFFu : 0u);
  if (hit) result = make_uint4(w.ox + w.cx, w.oy + w.cy, w.oz + w.cz, 1u);
  hits[ri] = result;
  if (tout) tout[ri] = hit ? hit_t(r.o, r.d, result) : 0.0;
}

__device__ __forceinline__ void traverse_ray(
    uint64_t ri, const float *__restrict__ rays, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, uint4 *__restrict__ hits, double *__restrict__ tout,
    const unsigned long long *rmask) {
  extern __shared__ uint32_t frames_lds[];
  Ray r;
  load_ray(rays, ri, r);
  Walk w;
  w.ox = w.oy = w.oz = w.cx = w.cy = w.cz = 0;
  uint32_t status = kLeft;
  float t0;
  if (clip_grid(r, n, leaf_count, t0)) {
    begin_walk(r, w, k, t0, nodes, leafw);
    do {
      status = walk_visit(r, w, nodes, leafw, bounds, frames_lds + threadIdx.x, blockDim.x,
                          rmask);
    } while (status == kWalking);
  }
  store_result(ri, r, w, status, hits, tout);
}

__global__ __launch_bounds__(256) void k_traverse(
    const float *__restrict__ rays, uint64_t R, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, uint4 *__restrict__ hits, double *__restrict__ tout) {
  __shared__ unsigned long long rmask[24];
  load_range_masks(rmask);
  const uint64_t ri = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ri >= R) return;
  traverse_ray(ri, rays, nodes, leafw, bounds, n, k, leaf_count, hits, tout, rmask);
}

// t of every ray from its hit record, as store_result computes it: the persistent walk
// leaves t to this pass, where all 64 lanes of a wave work, instead of computing it for the
// few lanes that finish in one visit.
__device__ __forceinline__ void hit_t_one(const float *__restrict__ rays, uint64_t ri,
                                          const uint4 *__restrict__ hits,
                                          double *__restrict__ tout) {
  const uint4 h = hits[ri];
  double t = 0.0;
  if (h.w == 1u) {
    const float2 a = *reinterpret_cast<const float2 *>(rays + 6 * ri);
    const float2 b = *reinterpret_cast<const float2 *>(rays + 6 * ri + 2);
    const float2 c = *reinterpret_cast<const float2 *>(rays + 6 * ri + 4);
    const float o[3] = {a.x, a.y, b.x}, d[3] = {b.y, c.x, c.y};
    t = hit_t(o, d, h);
  }
  tout[ri] = t;
}

__global__ __launch_bounds__(256) void k_hit_t(const float *__restrict__ rays, uint64_t R,
                                               const uint4 *__restrict__ hits,
                                               double *__restrict__ tout) {
  const uint64_t ri = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ri < R) hit_t_one(rays, ri, hits, tout);
}

// The mirror's walk statistics over a ray batch (the reference's measure.rs counters):
// sums of leaf descents, node descents and cell visits, one atomic per wave each.
__global__ __launch_bounds__(64) void k_traverse_stats(
    const float *__restrict__ rays, uint64_t R, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, unsigned long long *__restrict__ sums) {
  extern __shared__ uint32_t frames_lds[];
  const uint64_t ri = uint64_t(blockIdx.x) * 64 + threadIdx.x;
  uint32_t st3[3] = {0, 0, 0};  // leaf descents, node descents, visits
  if (ri < R) {
    Ray r;
    load_ray(rays, ri, r);
    Walk w;
    w.ox = w.oy = w.oz = w.cx = w.cy = w.cz = 0;
    float t0;
    if (clip_grid(r, n, leaf_count, t0)) {
      begin_walk(r, w, k, t0, nodes, leafw);
      uint32_t st;
      do {
        st = walk_visit<true>(r, w, nodes, leafw, bounds, frames_lds + threadIdx.x, 64, nullptr,
                              st3);
      } while (st == kWalking);
    }
  }
  unsigned long long v[3] = {st3[0], st3[1], st3[2]};
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v[i] += (unsigned long long)__shfl_xor((long long)v[i], d, 64);
  }
  if (threadIdx.x == 0) {
    atomicAdd(&sums[0], v[0]);
    atomicAdd(&sums[1], v[1]);
    atomicAdd(&sums[2], v[2]);
  }
}

// Persistent lanes: a grid sized to the device's resident one-wave blocks; each wave takes
// runs of kRayRun consecutive rays from a counter (one atomic per run), and whenever at
// least refill_min of its lanes are idle those lanes take the next rays of its run, so a
// wave is not held by its longest walk. Every ray's walk is the same as in k_traverse;
// only the lane that runs it differs. t is left to k_hit_t.
constexpr uint32_t kRayRun = 256;
static_assert(kRayRun == kBatchRayRun, "the batch run table counts runs of kRayRun rays");

__global__ __launch_bounds__(64) void k_traverse_pl(
    const float *__restrict__ rays, uint64_t R, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, uint4 *__restrict__ hits,
    unsigned long long *__restrict__ counter, uint32_t refill_min) {
  extern __shared__ uint32_t frames_lds[];
  __shared__ unsigned long long rmask[24];
  load_range_masks(rmask);
  uint32_t *const frames = frames_lds + threadIdx.x;
  const uint32_t lane = threadIdx.x;
  const unsigned long long below = (1ull << lane) - 1ull;
  Ray r;
  Walk w;
  w.ox = w.oy = w.oz = w.cx = w.cy = w.cz = 0;
  uint64_t ri = 0;
  uint64_t run = 0, run_end = 0;  // the wave's rays not yet handed out (wave-uniform)
  bool active = false, exhausted = false, first = true;
  for (;;) {
    // refill: the idle lanes take the next rays of the wave's run
    const unsigned long long idle = __ballot(!active);
    const uint32_t ni = __popcll(idle);
    if (!exhausted) {  // wave-uniform; ni >= refill_min here (or the wave just started)
      if (run == run_end) {
        // the wave's first run is its own (run = its block index); later runs come from
        // the counter, past the grid's first runs (one counter address serializes its
        // atomics, so the first runs do not wait on it)
        unsigned long long base = uint64_t(blockIdx.x) * kRayRun;
        if (!first && lane == 0)
          base = atomicAdd(counter, (unsigned long long)kRayRun) +
                 uint64_t(gridDim.x) * kRayRun;
        first = false;
        run = __shfl(base, 0, 64);
        run_end = run + kRayRun < R ? run + kRayRun : R;
        if (run >= R) {
          exhausted = true;
          run = run_end = R;
        }
      }
      const uint64_t take = run_end - run < ni ? run_end - run : ni;
      if (!active) {
        const uint32_t rank = __popcll(idle & below);
        if (rank < take) {
          ri = run + rank;
          load_ray(rays, ri, r);
          float t0;
          if (clip_grid(r, n, leaf_count, t0)) {
            begin_walk(r, w, k, t0, nodes, leafw);
            active = true;
          } else {
            store_result(ri, r, w, kLeft, hits, nullptr);
          }
        }
      }
      run += take;
    } else if (ni == 64) {
      break;  // no rays left and every lane idle (wave-uniform)
    }
    // walk until refill_min lanes are idle (every lane, once the rays are exhausted); the
    // refill stays outside this loop so the walk state is not merged with it every visit
    const uint32_t until = exhausted ? 64u : refill_min;
    for (;;) {
      if (active) {
        const uint32_t status =
            walk_visit<false, true>(r, w, nodes, leafw, bounds, frames, 64, rmask);
        if (status != kWalking) {
          store_result(ri, r, w, status, hits, nullptr);
          active = false;
        }
      }
      if (uint32_t(__popcll(__ballot(!active))) >= until) break;  // wave-uniform
    }
  }
}

// Persistent lanes over the segments of a batched traversal. Segment g's rays form runs of
// kRayRun ([run_begin[g], run_begin[g+1]) in the run table), so a run never spans two trees;
// waves take runs with one atomic and hand idle lanes the next rays of their run, as
// k_traverse_pl does for one tree. A lane keeps the tree of the ray it walks (its pointers
// in vector registers), so lanes of one wave may walk different segments' trees. t is left
// to k_hit_t_batch.
__global__ __launch_bounds__(64) void k_traverse_batch_pl(
    const TraverseSeg *__restrict__ seg, const uint64_t *__restrict__ run_begin, uint32_t S,
    unsigned long long *__restrict__ counter, uint32_t refill_min) {
  extern __shared__ uint32_t frames_lds[];
  __shared__ unsigned long long rmask[24];
  load_range_masks(rmask);
  uint32_t *const frames = frames_lds + threadIdx.x;
  const uint32_t lane = threadIdx.x;
  const unsigned long long below = (1ull << lane) - 1ull;
  const uint64_t nruns = run_begin[S];
  Ray r;
  Walk w;
  w.ox = w.oy = w.oz = w.cx = w.cy = w.cz = 0;
  const uint32_t *nodes = nullptr, *leafw = nullptr, *bounds = nullptr;  // the lane's tree
  uint4 *hit = nullptr;                                                   // the lane's record
  uint32_t cur = 0;               // segment of the wave's run (wave-uniform)
  uint64_t run = 0, run_end = 0;  // its rays not yet handed out (wave-uniform)
  bool active = false, exhausted = false, first = true;
  for (;;) {
    const unsigned long long idle = __ballot(!active);
    const uint32_t ni = __popcll(idle);
    if (!exhausted) {  // wave-uniform
      if (run == run_end) {
        // the wave's first run is its own, later ones from the counter (as k_traverse_pl)
        unsigned long long base = blockIdx.x;
        if (!first && lane == 0) base = atomicAdd(counter, 1ull) + gridDim.x;
        first = false;
        base = __shfl(base, 0, 64);
        if (base >= nruns) {
          exhausted = true;
          run = run_end = 0;
        } else {
          cur = dev::upper_index(run_begin, S, uint64_t(base));
          run = (base - run_begin[cur]) * kRayRun;
          const uint64_t R = seg[cur].R;
          run_end = run + kRayRun < R ? run + kRayRun : R;
        }
      }
      const uint64_t take = run_end - run < ni ? run_end - run : ni;
      if (!active) {
        const uint32_t rank = __popcll(idle & below);
        if (rank < take) {
          const TraverseSeg &sg = seg[cur];
          const uint64_t ri = run + rank;
          load_ray(sg.rays, ri, r);
          hit = sg.hits + ri;
          float t0;
          if (clip_grid(r, sg.n, sg.leaf_count, t0)) {
            nodes = sg.nodes;
            leafw = sg.leafw;
            bounds = sg.bounds;
            begin_walk(r, w, sg.k, t0, nodes, leafw);
            active = true;
          } else {
            store_result(0, r, w, kLeft, hit, nullptr);
          }
        }
      }
      run += take;
    } else if (ni == 64) {
      break;  // no runs left and every lane idle (wave-uniform)
    }
    const uint32_t until = exhausted ? 64u : refill_min;
    for (;;) {
      if (active) {
        const uint32_t status = walk_visit(r, w, nodes, leafw, bounds, frames, 64, rmask);
        if (status != kWalking) {
          store_result(0, r, w, status, hit, nullptr);
          active = false;
        }
      }
      if (uint32_t(__popcll(__ballot(!active))) >= until) break;  // wave-uniform
    }
  }
}

// t of every ray of a batched traversal from its hit record: block b takes run b.
__global__ __launch_bounds__(kRayRun) void k_hit_t_batch(const TraverseSeg *__restrict__ seg,
                                                         const uint64_t *__restrict__ run_begin,
                                                         uint32_t S) {
  const uint64_t b = blockIdx.x;
  const uint32_t g = dev::upper_index(run_begin, S, b);
  const TraverseSeg &sg = seg[g];
  const uint64_t ri = (b - run_begin[g]) * kRayRun + threadIdx.x;
  if (ri < sg.R && sg.t) hit_t_one(sg.rays, ri, sg.hits, sg.t);
}

// ---- device ray generation + shading (render.wgsl:54-91) ----------------------------------

__device__ __forceinline__ uint32_t unorm8(float v) {  // rgba8unorm: clamp, scale, round
  v = v > 0.0f ? v : 0.0f;  // NaN -> 0
  v = v < 1.0f ? v : 1.0f;
  return uint32_t(rintf(v * 255.0f));
}

__device__ __forceinline__ uint32_t rgba8(float r, float g, float b, float a) {
  return unorm8(r) | (unorm8(g) << 8) | (unorm8(b) << 16) | (unorm8(a) << 24);
}

// One wave per 8 x 8 pixel tile (neighbouring pixels walk the same cells). Per pixel the
// camera ray of render.wgsl:66-70 in binary32 (normalize(v) as v / sqrt(v.v), the dot
// product summed (x x + y y) + z z), the walk of k_traverse, and the shading of
// render.wgsl:74-89: a hit takes its voxel's material (0 without materials), shaded by
// position (voxel / n) for material 0 or by colors[material] (RGBA8, R in the low byte;
// an id past the table reads 0); a miss takes the sky gradient. Output RGBA8 per pixel.
// kSkip: the leaf-bounds pop of walk_visit (trees beyond the Infinity Cache).
template <bool kSkip>
__global__ __launch_bounds__(64) void k_render(vxl_camera cam, const uint32_t *__restrict__ nodes,
                                               const uint32_t *__restrict__ leafw,
                                               const uint32_t *__restrict__ bounds, uint32_t n,
                                               uint32_t k, uint64_t leaf_count,
                                               const uint16_t *__restrict__ mats,
                                               const uint32_t *__restrict__ colors,
                                               uint32_t ncolors, uint32_t *__restrict__ out) {
  extern __shared__ uint32_t frames_lds[];
  __shared__ unsigned long long rmask[24];
  load_range_masks(rmask);
  const uint32_t px = blockIdx.x * 8 + (threadIdx.x & 7), py = blockIdx.y * 8 + (threadIdx.x >> 3);
  if (px >= cam.width || py >= cam.height) return;
  const float w = float(cam.width), h = float(cam.height);
  const float ndc_x = ((float(px) + 0.5f) / w * 2.0f - 1.0f) * cam.tan_half_fov * cam.aspect;
  const float ndc_y = (1.0f - (float(py) + 0.5f) / h * 2.0f) * cam.tan_half_fov;
  Ray r;
  {
    float v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) v[a] = (cam.forward[a] + cam.right[a] * ndc_x) + cam.up[a] * ndc_y;
    const float len = __fsqrt_rn((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      r.o[a] = cam.eye[a];
      r.d[a] = v[a] / len;
    }
    finish_ray(r);
  }
  Walk wk;
  wk.ox = wk.oy = wk.oz = wk.cx = wk.cy = wk.cz = 0;
  uint32_t status = kLeft;
  float t0;
  if (clip_grid(r, n, leaf_count, t0)) {
    begin_walk(r, wk, k, t0, nodes, leafw);
    do {
      status = walk_visit<false, kSkip>(r, wk, nodes, leafw, bounds, frames_lds + threadIdx.x, 64,
                                        rmask);
    } while (status == kWalking);
  }
  uint32_t color;
  if (status == kHit) {
    const uint32_t mat = mats ? uint32_t(mats[uint64_t(wk.node) * 512 +
                                              dev::morton9(wk.cx, wk.cy, wk.cz)])
                              : 0u;
    if (mat == 0) {
      color = rgba8(float(wk.ox + wk.cx) / cam.n, float(wk.oy + wk.cy) / cam.n,
                    float(wk.oz + wk.cz) / cam.n, 1.0f);
    } else {
      color = mat < ncolors ? colors[mat] : 0u;
    }
  } else {
    const float t = float(py) / h;
    color = rgba8(0.08f * (1.0f - t), 0.10f * (1.0f - t), 0.16f + 0.12f * t, 1.0f);
  }
  out[uint64_t(py) * cam.width + px] = color;
}

// Segment g owns blocks [blk_begin[g], blk_begin[g+1]), so every block reads one tree.
__global__ __launch_bounds__(256) void k_traverse_batch(const TraverseSeg *__restrict__ seg,
                                                        const uint32_t *__restrict__ blk_begin,
                                                        uint32_t S) {
  const uint32_t g = dev::upper_index(blk_begin, S, blockIdx.x);
  const TraverseSeg &sg = seg[g];
  const uint64_t ri = uint64_t(blockIdx.x - blk_begin[g]) * blockDim.x + threadIdx.x;
  __shared__ unsigned long long rmask[24];
  load_range_masks(rmask);
  if (ri >= sg.R) return;
  traverse_ray(ri, sg.rays, sg.nodes, sg.leafw, sg.bounds, sg.n, sg.k, sg.leaf_count, sg.hits,
               sg.t, rmask);
}

// LDS for the suspended frames: levels 2..k+1 (the walk saves a frame when it descends
// from it), kFrameWords words per thread each.
size_t frame_lds_bytes(uint32_t k, uint64_t lanes) {
  return size_t(k >= 1 ? k : 1) * kFrameWords * 4 * lanes;
}

// Work counters of persistent traversals in flight at once on one context (at most).
constexpr uint32_t kCounterSlots = 1024;

// Zeroes a persistent traversal's work counter on the launch's stream: a kernel rather than
// a memset, so the pair stays correct when a caller captures the call in a HIP graph.
__global__ void k_clear_counter(unsigned long long *ctr) { *ctr = 0ull; }

// The automatic lane schedule: persistent lanes refilled at 8 idle for a tree whose walked
// arrays (64-byte leaves, 4-byte bounds, 12-byte nodes) exceed the 256 MiB Infinity Cache,
// walked from HBM where a wave's lanes wait on leaf loads (soup_2048: 3.18 -> 2.33 ms when
// this schedule was introduced), and for a tree of k >= 4 levels (n = 2048 and up), whose
// walks are long and uneven (terrain_1024's tree, 15 MB: 67.8 visits and 14.9 node
// descents per ray; 3.96 -> 3.49 ms) when the launch holds at least 4 x the rays the
// device's resident waves walk at once (smaller launches fill fewer lanes on persistent
// lanes than one lane per ray does: 512 plain traversals of 1k-1M rays, 97 vs 119 ms);
// one lane per ray for the shallower cache-resident trees, bound by instruction issue,
// where the refills' partial-wave ray setups cost more than the idle lanes (sphere_256,
// k = 3: 0.77 vs 0.81 ms).
constexpr uint64_t kCacheBytes = 256ull << 20;
constexpr int kAutoRefill = 8;

int refill_for(vxl_ctx *ctx, const vxl_tree *tree, uint64_t R) {
  if (ctx->dda_refill >= 0) return ctx->dda_refill;
  const uint64_t walked = tree->leaf_count * 68 + tree->node_count * 12;
  if (walked > kCacheBytes) return kAutoRefill;
  if (tree->k < 4) return 0;
  if (!ctx->pl_min_rays) {
    int per_cu = 0;
    VXL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traverse_pl, 64,
                                                         frame_lds_bytes(4, 64)));
    if (per_cu < 1) per_cu = 1;
    ctx->pl_min_rays =
        4ull * uint64_t(ctx->cu_count > 0 ? ctx->cu_count : 256) * uint64_t(per_cu) * 64ull;
  }
  return R >= ctx->pl_min_rays ? kAutoRefill : 0;
}

}  // namespace

void launch_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                     uint32_t *d_hits, double *d_t, hipStream_t s, unsigned long long *counter) {
  if (R == 0) return;
  const size_t lds64 = frame_lds_bytes(tree->k, 64);
  const int refill = refill_for(ctx, tree, R);
  if (refill > 0) {
    // persistent lanes: as many one-wave blocks as the device holds at once, each launch
    // with a counter slot of its own (a ring, so launches on other streams keep theirs)
    int per_cu = 0;
    VXL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traverse_pl, 64, lds64));
    if (per_cu < 1) per_cu = 1;
    const uint64_t waves = (R + 63) / 64;
    uint64_t blocks = uint64_t(ctx->cu_count > 0 ? ctx->cu_count : 256) * uint64_t(per_cu);
    if (blocks > waves) blocks = waves;
    unsigned long long *ctr = counter;
    if (!ctr) {
      ctx->trav_ctr.reserve(kCounterSlots * 8 + 64);
      ctr = ctx->trav_ctr.as<unsigned long long>() + (ctx->trav_ctr_next++ % kCounterSlots);
    }
    hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(1), 0, s, ctr);
    hipLaunchKernelGGL(k_traverse_pl, dim3(uint32_t(blocks)), dim3(64), lds64, s, d_rays, R,
                       tree->nodes, reinterpret_cast<const uint32_t *>(tree->leaves),
                       tree->bounds, tree->n, tree->k, tree->leaf_count,
                       reinterpret_cast<uint4 *>(d_hits), ctr, uint32_t(refill));
    if (d_t) {
      const uint64_t tb = (R + 255) / 256;
      hipLaunchKernelGGL(k_hit_t, dim3(uint32_t(tb)), dim3(256), 0, s, d_rays, R,
                         reinterpret_cast<const uint4 *>(d_hits), d_t);
    }
    VXL_HIP(hipGetLastError());
    return;
  }
  launch_traverse_plain(ctx, tree, d_rays, R, d_hits, d_t, s);
}

void launch_traverse_plain(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                           uint32_t *d_hits, double *d_t, hipStream_t s) {
  if (R == 0) return;
  const int tb = ctx->dda_block;
  const uint64_t blocks = (R + tb - 1) / tb;
  if (blocks * uint64_t(tb) > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "traverse: ray batch exceeds the grid limit");
  hipLaunchKernelGGL(k_traverse, dim3(uint32_t(blocks)), dim3(tb), frame_lds_bytes(tree->k, tb), s,
                     d_rays, R, tree->nodes,
                     reinterpret_cast<const uint32_t *>(tree->leaves), tree->bounds, tree->n,
                     tree->k, tree->leaf_count, reinterpret_cast<uint4 *>(d_hits), d_t);
  VXL_HIP(hipGetLastError());
}

void traverse_stats(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                    uint64_t out[3]) {
  out[0] = out[1] = out[2] = 0;
  if (R == 0) return;
  // three sums past the ring of persistent-traversal counters (the call waits for them)
  ctx->trav_ctr.reserve(kCounterSlots * 8 + 64);
  unsigned long long *sums = ctx->trav_ctr.as<unsigned long long>() + kCounterSlots;
  VXL_HIP(hipMemsetAsync(sums, 0, 24, ctx->stream));
  const uint64_t blocks = (R + 63) / 64;
  hipLaunchKernelGGL(k_traverse_stats, dim3(uint32_t(blocks)), dim3(64),
                     frame_lds_bytes(tree->k, 64), ctx->stream, d_rays, R, tree->nodes,
                     reinterpret_cast<const uint32_t *>(tree->leaves), tree->bounds, tree->n,
                     tree->k, tree->leaf_count, sums);
  VXL_HIP(hipGetLastError());
  VXL_HIP(hipMemcpyAsync(ctx->host_pinned, sums, 24, hipMemcpyDeviceToHost, ctx->stream));
  wait_stream(ctx, ctx->stream);
  const uint64_t *h = static_cast<const uint64_t *>(ctx->host_pinned);
  out[0] = h[0];
  out[1] = h[1];
  out[2] = h[2];
}

void launch_render(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera &cam,
                   const uint32_t *d_colors, uint32_t ncolors, uint32_t *d_rgba, hipStream_t s) {
  if (cam.width == 0 || cam.height == 0) return;
  const dim3 grid((cam.width + 7) / 8, (cam.height + 7) / 8, 1);
  const bool skip = tree->leaf_count * 68 + tree->node_count * 12 > kCacheBytes;
  hipLaunchKernelGGL(skip ? k_render<true> : k_render<false>, grid, dim3(64),
                     frame_lds_bytes(tree->k, 64), s, cam, tree->nodes,
                     reinterpret_cast<const uint32_t *>(tree->leaves), tree->bounds, tree->n,
                     tree->k, tree->leaf_count, tree->materials, d_colors, ncolors, d_rgba);
  VXL_HIP(hipGetLastError());
}

void launch_traverse_batch(vxl_ctx *ctx, const RaySegment *segs, uint32_t S, hipStream_t s) {
  const uint32_t tb = uint32_t(ctx->dda_block);
  uint64_t blocks = 0, runs = 0;
  uint32_t kmax = 0;
  bool want_t = false;
  for (uint32_t g = 0; g < S; g++) {
    blocks += (segs[g].R + tb - 1) / tb;
    runs += (segs[g].R + kRayRun - 1) / kRayRun;
    kmax = segs[g].tree->k > kmax ? segs[g].tree->k : kmax;
    want_t = want_t || (segs[g].t && segs[g].R);
  }
  if (blocks == 0) return;
  if (blocks > 0xffffffffull || runs > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "traverse batch exceeds the grid limit");
  const size_t seg_bytes = (S * sizeof(TraverseSeg) + 255) / 256 * 256;
  const size_t blk_bytes = ((S + 1) * 4 + 255) / 256 * 256;
  const size_t run_bytes = ((S + 1) * 8 + 255) / 256 * 256;
  const size_t bytes = seg_bytes + blk_bytes + run_bytes;
  staging_begin(ctx, bytes);
  char *h = static_cast<char *>(staging_take(ctx, bytes));
  TraverseSeg *hs = reinterpret_cast<TraverseSeg *>(h);
  uint32_t *hb = reinterpret_cast<uint32_t *>(h + seg_bytes);
  uint64_t *hr = reinterpret_cast<uint64_t *>(h + seg_bytes + blk_bytes);
  blocks = runs = 0;
  for (uint32_t g = 0; g < S; g++) {
    const vxl_tree *t = segs[g].tree;
    hs[g] = TraverseSeg{t->nodes, reinterpret_cast<const uint32_t *>(t->leaves), t->bounds,
                        segs[g].rays, reinterpret_cast<uint4 *>(segs[g].hits), segs[g].t,
                        t->leaf_count, segs[g].R, t->n, t->k};
    hb[g] = uint32_t(blocks);
    hr[g] = runs;
    blocks += (segs[g].R + tb - 1) / tb;
    runs += (segs[g].R + kRayRun - 1) / kRayRun;
  }
  hb[S] = uint32_t(blocks);
  hr[S] = runs;
  ctx->trav_desc.reserve(bytes);
  char *d = static_cast<char *>(ctx->trav_desc.ptr);
  // the previous batch's kernel (possibly on another stream) reads the descriptors
  if (ctx->trav_done) VXL_HIP(hipStreamWaitEvent(s, ctx->trav_done, 0));
  else VXL_HIP(hipEventCreateWithFlags(&ctx->trav_done, hipEventDisableTiming));
  VXL_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
  staging_end(ctx, s);
  unsigned long long *ctr = nullptr;
  if (ctx->dda_batch_refill > 0) {
    ctx->trav_ctr.reserve(kCounterSlots * 8 + 64);
    ctr = ctx->trav_ctr.as<unsigned long long>() + (ctx->trav_ctr_next++ % kCounterSlots);
  }
  launch_traverse_segs(ctx, reinterpret_cast<const TraverseSeg *>(d),
                       reinterpret_cast<const uint32_t *>(d + seg_bytes),
                       reinterpret_cast<const uint64_t *>(d + seg_bytes + blk_bytes), S, blocks,
                       runs, kmax, want_t, ctr, s);
  VXL_HIP(hipEventRecord(ctx->trav_done, s));
}

uint64_t persistent_batch_min_rays(vxl_ctx *ctx) {
  if (!ctx->batch_pl_min_rays) {
    int per_cu = 0;
    VXL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traverse_batch_pl, 64,
                                                         frame_lds_bytes(4, 64)));
    if (per_cu < 1) per_cu = 1;
    ctx->batch_pl_min_rays =
        4ull * uint64_t(ctx->cu_count > 0 ? ctx->cu_count : 256) * uint64_t(per_cu) * 64ull;
  }
  return ctx->batch_pl_min_rays;
}

void launch_traverse_segs(vxl_ctx *ctx, const TraverseSeg *d_segs, const uint32_t *d_blk,
                          const uint64_t *d_runs, uint32_t S, uint64_t blocks, uint64_t runs,
                          uint32_t kmax, bool want_t, unsigned long long *counter,
                          hipStream_t s) {
  if (blocks == 0) return;
  if (blocks > 0xffffffffull || runs > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "traverse batch exceeds the grid limit");
  if (ctx->dda_batch_refill > 0) {
    if (!counter) throw Error(VXL_DEVICE_ERROR, "internal: persistent batch without a counter");
    const size_t lds64 = frame_lds_bytes(kmax, 64);
    int per_cu = 0;
    VXL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traverse_batch_pl, 64, lds64));
    if (per_cu < 1) per_cu = 1;
    uint64_t grid = uint64_t(ctx->cu_count > 0 ? ctx->cu_count : 256) * uint64_t(per_cu);
    if (grid > runs * (kRayRun / 64)) grid = runs * (kRayRun / 64);
    hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(k_traverse_batch_pl, dim3(uint32_t(grid)), dim3(64), lds64, s, d_segs,
                       d_runs, S, counter, uint32_t(ctx->dda_batch_refill));
    if (want_t)
      hipLaunchKernelGGL(k_hit_t_batch, dim3(uint32_t(runs)), dim3(kRayRun), 0, s, d_segs,
                         d_runs, S);
    VXL_HIP(hipGetLastError());
    return;
  }
  const uint32_t tb = uint32_t(ctx->dda_block);
  hipLaunchKernelGGL(k_traverse_batch, dim3(uint32_t(blocks)), dim3(tb), frame_lds_bytes(kmax, tb),
                     s, d_segs, d_blk, S);
  VXL_HIP(hipGetLastError());
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
unsigned long long *g_noalloc_ptr[kMaxDevices] = {};  // scratch's address on each device
std::atomic<uint32_t> g_noalloc_next[kMaxDevices];    // next slot per device
}  // namespace

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw Error(VXL_NO_DEVICE, "device index beyond the allocation-free path's table");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return;
  // the caller made `device` current; the symbol's address differs per device
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    throw Error(VXL_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                      hipGetErrorString(e));
  }
  g_noalloc_ptr[device] = static_cast<unsigned long long *>(p);
}

unsigned long long *noalloc_counter(int device) {
  unsigned long long *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (device >= 0 && device < kMaxDevices) base = g_noalloc_ptr[device];
  }
  if (!base) throw Error(VXL_DEVICE_ERROR, "allocation-free scratch not set up for this device");
  return base + (g_noalloc_next[device].fetch_add(1, std::memory_order_relaxed) % kNoAllocSlots);
}

}  // namespace vxl
