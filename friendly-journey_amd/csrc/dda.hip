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
// Frame storage: the active frame lives in registers. Each ancestor frame is kept in
// a register slot named by its level (the stack index of a frame is root_level - level,
// so "the frame at level L" is unique at any time); saving and restoring go through a
// switch on the level, so no frame is ever addressed dynamically.
//
// The hit's ray parameter t is the entry parameter of the f64 slab of the hit voxel for
// the ray widened to f64: max(0, max_a min((v_a - o_a)/d_a, (v_a + 1 - o_a)/d_a)) over
// axes with d_a != 0, which is the t_enter the f64 reference traversal reports for that
// voxel (voxel-core/src/layout.rs:176-183).
#include "vxl_internal.hpp"

namespace vxl {
namespace {

constexpr uint32_t kFull = (7u << 9) | (7u << 12) | (7u << 15);
constexpr uint32_t kMaxIter = 1u << 22;

// A suspended ancestor frame lives in LDS, one row per (level, field), one column per
// thread: node, packed cell, t_max x 3. Its t_delta is not kept (a function of the frame's
// cell size and the ray, recomputed with the same operations on restore), nor its t_entry
// (a restored frame steps at once, which overwrites it, or the walk ends).
constexpr int kFrameWords = 5;

__device__ __forceinline__ float axis_delta(float cs, float d, float rinv, bool big) {
  float q = cs * rinv;
  if (big) q = cs / d;
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
    if (r.d[a] == 0.0f) {
      if (r.o[a] < lo || r.o[a] > hi) miss = true;
    } else {
      float t1 = (lo - r.o[a]) * r.rinv[a];
      float t2 = (hi - r.o[a]) * r.rinv[a];
      if (t1 > t2) {
        const float s = t1;
        t1 = t2;
        t2 = s;
      }
      t_near = fmaxf(t_near, t1);
      t_far = fminf(t_far, t2);
    }
  }
  if (miss || t_near > t_far) return false;
  return t_far >= t_enter;
}

__device__ __forceinline__ void traverse_ray(
    uint64_t ri, const float *__restrict__ rays, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, uint4 *__restrict__ hits, double *__restrict__ tout) {
  Ray r;
  {
    const float2 a = *reinterpret_cast<const float2 *>(rays + 6 * ri);
    const float2 b = *reinterpret_cast<const float2 *>(rays + 6 * ri + 2);
    const float2 c = *reinterpret_cast<const float2 *>(rays + 6 * ri + 4);
    r.o[0] = a.x; r.o[1] = a.y; r.o[2] = b.x;
    r.d[0] = b.y; r.d[1] = c.x; r.d[2] = c.y;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    r.rinv[a] = 1.0f / r.d[a];
    r.big[a] = !(fabsf(r.d[a]) < 0x1p100f);
  }
  uint4 result = make_uint4(0u, 0u, 0u, 0u);
  bool hit = false;

  // grid clip, ray_aabb_f32 (mirror.rs:169-192) + t1 < 0 check (:273-277)
  const float nf = float(n);
  float t0 = -INFINITY, t1 = INFINITY;
  bool missed = leaf_count == 0;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (missed) break;
    if (r.d[a] == 0.0f) {
      if (r.o[a] < 0.0f || r.o[a] > nf) missed = true;
    } else {
      float ta = (0.0f - r.o[a]) * r.rinv[a];
      float tb = (nf - r.o[a]) * r.rinv[a];
      if (ta > tb) {
        const float s = ta;
        ta = tb;
        tb = s;
      }
      t0 = fmaxf(t0, ta);
      t1 = fminf(t1, tb);
      if (t0 > t1) missed = true;
    }
  }
  if (!missed && t1 < 0.0f) missed = true;

  if (!missed) {
    const uint32_t root_level = k == 0 ? 1u : k + 1;
    // active frame
    uint32_t level = root_level, node = 0;
    uint32_t ox = 0, oy = 0, oz = 0;
    uint32_t cx, cy, cz;
    float tm0, tm1, tm2, td0, td1, td2;
    float te = fmaxf(t0, 0.0f);
    uint32_t mlo = 0, mhi = 0, cbase = 0;  // cached node record of a node frame
    extern __shared__ uint32_t frames_lds[];
    uint32_t *const frames = frames_lds + threadIdx.x;  // frames[((L-2)*5 + f) * blockDim.x]
    const uint32_t fstride = blockDim.x;

#define VXL_MAKE_FRAME(LEVEL, ORGX, ORGY, ORGZ, TENTER)                                    \
  {                                                                                        \
    const uint32_t dim_ = (LEVEL) == 1 ? 8u : 4u;                                          \
    const uint32_t csu_ = frame_cell_size(LEVEL);                                          \
    const float cs_ = float(csu_);                                                         \
    const float ics_ = __uint_as_float((127u - (31u - __clz(csu_))) << 23); /* 2^-e */       \
    axis_enter(r, 0, float(ORGX), dim_, cs_, ics_, TENTER, cx, tm0, td0);                  \
    axis_enter(r, 1, float(ORGY), dim_, cs_, ics_, TENTER, cy, tm1, td1);                  \
    axis_enter(r, 2, float(ORGZ), dim_, cs_, ics_, TENTER, cz, tm2, td2);                  \
    te = (TENTER);                                                                         \
  }

    // Walker::step (mirror.rs:87-116) on the active frame; OK = false when exhausted. The
    // axis choice (lowest t_max, ties to the lowest axis) and the update are selects, so
    // lanes stepping along different axes do not split the wave.
#define VXL_STEP(OK)                                                                       \
  {                                                                                        \
    const uint32_t dim_ = level == 1 ? 8u : 4u;                                            \
    const bool a1_ = tm1 < tm0;                                                            \
    float best_ = a1_ ? tm1 : tm0;                                                         \
    const bool a2_ = tm2 < best_;                                                          \
    best_ = a2_ ? tm2 : best_;                                                             \
    const bool on0_ = !a1_ && !a2_, on1_ = a1_ && !a2_;                                    \
    const float da_ = a2_ ? r.d[2] : (a1_ ? r.d[1] : r.d[0]);                              \
    const uint32_t ca_ = a2_ ? cz : (a1_ ? cy : cx);                                       \
    OK = isfinite(best_) && ((da_ > 0.0f && ca_ + 1 < dim_) || (da_ < 0.0f && ca_ != 0));  \
    if (OK) {                                                                              \
      const uint32_t inc_ = da_ > 0.0f ? 1u : 0xFFFFFFFFu;                                 \
      cx += on0_ ? inc_ : 0u;                                                              \
      cy += on1_ ? inc_ : 0u;                                                              \
      cz += a2_ ? inc_ : 0u;                                                               \
      tm0 = on0_ ? tm0 + td0 : tm0;                                                        \
      tm1 = on1_ ? tm1 + td1 : tm1;                                                        \
      tm2 = a2_ ? tm2 + td2 : tm2;                                                         \
      te = best_;                                                                          \
    }                                                                                      \
  }

    VXL_MAKE_FRAME(root_level, 0u, 0u, 0u, te);
    if (level >= 2) {
      mlo = nodes[0];
      mhi = nodes[1];
      cbase = nodes[2];
    }
    bool done = false;
    uint32_t iter = 0;
    while (!done) {
      if (++iter > kMaxIter) {  // bounded: a legitimate walk never comes close
        result = make_uint4(0u, 0u, 0u, 0xFFFFFFFFu);
        break;
      }
      bool stepped;
      // the active cell: a voxel bit of the leaf (one word load) or a child bit of the node
      // record held in registers; one step site serves both kinds of frame
      uint32_t word, bitpos;
      if (level == 1) {
        const uint32_t m = dev::morton9(cx, cy, cz);
        word = leafw[uint64_t(node) * 16 + (m >> 5)];
        bitpos = m & 31;
      } else {
        const uint32_t bit = dev::child_bit(cx, cy, cz);
        word = bit < 32 ? mlo : mhi;
        bitpos = bit & 31;
      }
      const bool has = (word >> bitpos) & 1u;
      if (level == 1 && has) {
        result = make_uint4(ox + cx, oy + cy, oz + cz, 1u);
        hit = true;
        break;
      }
      if (level > 1 && has) {
        const uint32_t bit = dev::child_bit(cx, cy, cz);
        const uint32_t size = frame_cell_size(level);
        const uint32_t cox = ox + cx * size, coy = oy + cy * size, coz = oz + cz * size;
        const uint32_t below = bit < 32 ? __popc(mlo & ((1u << bit) - 1u))
                                        : __popc(mlo) + __popc(mhi & ((1u << (bit - 32)) - 1u));
        const uint32_t slot = cbase + below;
        const bool descend = level != 2 || leaf_reaches(bounds[slot], r, cox, coy, coz, te);
        if (descend) {
          {
            uint32_t *f = frames + (level - 2) * kFrameWords * fstride;
            f[0] = node;
            f[fstride] = cx | (cy << 8) | (cz << 16);
            f[2 * fstride] = __float_as_uint(tm0);
            f[3 * fstride] = __float_as_uint(tm1);
            f[4 * fstride] = __float_as_uint(tm2);
          }
          const float t_in = te;
          level -= 1;
          node = slot;
          ox = cox;
          oy = coy;
          oz = coz;
          VXL_MAKE_FRAME(level, ox, oy, oz, t_in);
          if (level >= 2) {
            mlo = nodes[3ull * node];
            mhi = nodes[3ull * node + 1];
            cbase = nodes[3ull * node + 2];
          }
          continue;
        }
      }
      VXL_STEP(stepped);
      // ascend (mirror.rs:155-165): pop parents until one steps to a new cell
      while (!stepped) {
        if (level == root_level) {
          done = true;
          break;
        }
        level += 1;
        const uint32_t size = frame_cell_size(level);
        {
          const uint32_t *f = frames + (level - 2) * kFrameWords * fstride;
          node = f[0];
          const uint32_t cell = f[fstride];
          cx = cell & 0xffu;
          cy = (cell >> 8) & 0xffu;
          cz = cell >> 16;
          tm0 = __uint_as_float(f[2 * fstride]);
          tm1 = __uint_as_float(f[3 * fstride]);
          tm2 = __uint_as_float(f[4 * fstride]);
          const float cs = float(size);
          td0 = axis_delta(cs, r.d[0], r.rinv[0], r.big[0]);
          td1 = axis_delta(cs, r.d[1], r.rinv[1], r.big[1]);
          td2 = axis_delta(cs, r.d[2], r.rinv[2], r.big[2]);
        }
        ox -= cx * size;
        oy -= cy * size;
        oz -= cz * size;
        mlo = nodes[3ull * node];
        mhi = nodes[3ull * node + 1];
        cbase = nodes[3ull * node + 2];
        VXL_STEP(stepped);
      }
    }
#undef VXL_MAKE_FRAME
#undef VXL_STEP
  }
  hits[ri] = result;
  if (tout) {
    double t = 0.0;
    if (hit) {
      const double v[3] = {double(result.x), double(result.y), double(result.z)};
      double tn = 0.0;
#pragma unroll
      for (int a = 0; a < 3; a++) {
        if (r.d[a] != 0.0f) {
          const double o = double(r.o[a]);
          const double inv = 1.0 / double(r.d[a]);
          const double ta = (v[a] - o) * inv, tb = (v[a] + 1.0 - o) * inv;
          tn = fmax(tn, fmin(ta, tb));
        }
      }
      t = tn;
    }
    tout[ri] = t;
  }
}

__global__ __launch_bounds__(256) void k_traverse(
    const float *__restrict__ rays, uint64_t R, const uint32_t *__restrict__ nodes,
    const uint32_t *__restrict__ leafw, const uint32_t *__restrict__ bounds, uint32_t n,
    uint32_t k, uint64_t leaf_count, uint4 *__restrict__ hits, double *__restrict__ tout) {
  const uint64_t ri = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ri >= R) return;
  traverse_ray(ri, rays, nodes, leafw, bounds, n, k, leaf_count, hits, tout);
}

// One segment of a batched traversal, as the kernel sees it.
struct TraverseSeg {
  const uint32_t *nodes, *leafw, *bounds;
  const float *rays;
  uint4 *hits;
  double *t;
  uint64_t leaf_count, R;
  uint32_t n, k;
};

// Segment g owns blocks [blk_begin[g], blk_begin[g+1]), so every block reads one tree.
__global__ __launch_bounds__(256) void k_traverse_batch(const TraverseSeg *__restrict__ seg,
                                                        const uint32_t *__restrict__ blk_begin,
                                                        uint32_t S) {
  const uint32_t g = dev::upper_index(blk_begin, S, blockIdx.x);
  const TraverseSeg &sg = seg[g];
  const uint64_t ri = uint64_t(blockIdx.x - blk_begin[g]) * blockDim.x + threadIdx.x;
  if (ri >= sg.R) return;
  traverse_ray(ri, sg.rays, sg.nodes, sg.leafw, sg.bounds, sg.n, sg.k, sg.leaf_count, sg.hits,
               sg.t);
}

// LDS for the suspended frames: levels 2..k+1 (the walk saves a frame when it descends
// from it), kFrameWords words per thread each.
size_t frame_lds_bytes(uint32_t k, uint64_t lanes) {
  return size_t(k >= 1 ? k : 1) * kFrameWords * 4 * lanes;
}

}  // namespace

void launch_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                     uint32_t *d_hits, double *d_t, hipStream_t s) {
  if (R == 0) return;
  const int tb = ctx->dda_block;
  const uint64_t blocks = (R + tb - 1) / tb;
  if (blocks > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "traverse: ray batch exceeds the grid limit");
  hipLaunchKernelGGL(k_traverse, dim3(uint32_t(blocks)), dim3(tb), frame_lds_bytes(tree->k, tb), s,
                     d_rays, R, tree->nodes,
                     reinterpret_cast<const uint32_t *>(tree->leaves), tree->bounds, tree->n,
                     tree->k, tree->leaf_count, reinterpret_cast<uint4 *>(d_hits), d_t);
  VXL_HIP(hipGetLastError());
}

void launch_traverse_batch(vxl_ctx *ctx, const RaySegment *segs, uint32_t S, hipStream_t s) {
  const uint32_t tb = uint32_t(ctx->dda_block);
  uint64_t blocks = 0;
  for (uint32_t g = 0; g < S; g++) blocks += (segs[g].R + tb - 1) / tb;
  if (blocks == 0) return;
  if (blocks > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "traverse batch exceeds the grid limit");
  const size_t seg_bytes = (S * sizeof(TraverseSeg) + 255) / 256 * 256;
  const size_t blk_bytes = ((S + 1) * 4 + 255) / 256 * 256;
  staging_begin(ctx, seg_bytes + blk_bytes);
  char *h = static_cast<char *>(staging_take(ctx, seg_bytes + blk_bytes));
  TraverseSeg *hs = reinterpret_cast<TraverseSeg *>(h);
  uint32_t *hb = reinterpret_cast<uint32_t *>(h + seg_bytes);
  blocks = 0;
  for (uint32_t g = 0; g < S; g++) {
    const vxl_tree *t = segs[g].tree;
    hs[g] = TraverseSeg{t->nodes, reinterpret_cast<const uint32_t *>(t->leaves), t->bounds,
                        segs[g].rays, reinterpret_cast<uint4 *>(segs[g].hits), segs[g].t,
                        t->leaf_count, segs[g].R, t->n, t->k};
    hb[g] = uint32_t(blocks);
    blocks += (segs[g].R + tb - 1) / tb;
  }
  hb[S] = uint32_t(blocks);
  ctx->trav_desc.reserve(seg_bytes + blk_bytes);
  char *d = static_cast<char *>(ctx->trav_desc.ptr);
  // the previous batch's kernel (possibly on another stream) reads the descriptors
  if (ctx->trav_done) VXL_HIP(hipStreamWaitEvent(s, ctx->trav_done, 0));
  else VXL_HIP(hipEventCreateWithFlags(&ctx->trav_done, hipEventDisableTiming));
  VXL_HIP(hipMemcpyAsync(d, h, seg_bytes + blk_bytes, hipMemcpyHostToDevice, s));
  staging_end(ctx, s);
  uint32_t kmax = 0;
  for (uint32_t g = 0; g < S; g++) kmax = segs[g].tree->k > kmax ? segs[g].tree->k : kmax;
  hipLaunchKernelGGL(k_traverse_batch, dim3(uint32_t(blocks)), dim3(tb),
                     frame_lds_bytes(kmax, tb), s,
                     reinterpret_cast<const TraverseSeg *>(d),
                     reinterpret_cast<const uint32_t *>(d + seg_bytes), S);
  VXL_HIP(hipGetLastError());
  VXL_HIP(hipEventRecord(ctx->trav_done, s));
}

}  // namespace vxl
