// Conservative triangle/box SAT voxelization into the dense leaf grid.
//
// Stage 1 (k_tri_setup / k_tri_setup_batch): world->grid transform
//   (voxelizer/src/core.rs:101-107), the epsilon-padded, clamped candidate box of
//   reference_cpu.rs:135-150, the candidate-voxel count and an 80-byte record per triangle.
// Stage 2: exclusive scan of the counts gives every (triangle, voxel) pair a global index.
//   Setup also leaves each block's total; one workgroup scans those and derives the SAT
//   plan from the grand total (k_setup_plan), all on the device with no host read-back.
// Stage 3 (k_offsets_units): per setup block, the pair offsets and the triangle holding the
//   first pair of each work unit. (Calls with over 16.7M triangles take a device-wide scan,
//   k_sat_plan and k_bstart instead.)
// Stage 4 (k_sat): a persistent grid of waves takes work units (ranges of consecutive
//   pairs): its own first unit, then more from atomic counters. A unit's triangles split
//   into segments of up to 12 consecutive pairs of one triangle; a lane walks one segment,
//   decoding its first pair to a voxel once and stepping through the candidate box in the
//   reference's scan order (x fastest, reference_cpu.rs:152-154). The plane test runs
//   first, survivors queue in LDS, and every 64 queued survivors run the remaining 12 axes
//   as a full wave. Each hit merges into a per-wave LDS table keyed by leaf-grid word; the
//   table is flushed with 64-bit atomic ORs into the leaf grid and the brick bitmap (bit per
//   8^3 brick).
// Many meshes share one launch: each record carries its mesh's leaf-grid base, so the
//   pair loop is the same for one mesh or hundreds.
// Owners (VoxelizeOpts::store_owner, reference_cpu.rs:163-169): with an owner array, each
//   hit also lowers owner[x + n*(y + n*z)] to its triangle index by an atomic min (a
//   relaxed read first: owners only fall, so a voxel already owned by a lower index
//   needs no atomic). The minimum over all covering triangles is order-free. A second
//   mode keeps owners per leaf voxel (owner[leaf * 512 + intra-brick Morton index], the
//   leaf found from a brick-code -> leaf map of an already built tree) and writes nothing
//   else: the owner pass of the material path, where a dense n^3 array would not fit.
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

constexpr uint32_t kLiteral = 1u;    // a grid vertex is not finite: every product is formed
constexpr uint32_t kWide = 2u;       // more than 2^22 candidates: 64-bit decode
constexpr uint32_t kPlaneHoist = 4u; // exact per-triangle plane normal and radius (below)
constexpr uint32_t kWindow = 8u;     // candidates: a window of W cells per column (below)

// The candidates a triangle enumerates are the cells of a box decoded x fastest: its
// clamped bounding box, or, for a windowed triangle, (W, ext_a, ext_b) -- W cells along
// the normal's largest axis c per (a, b) column, starting at floor(k0 + ka*a + kb*b).
struct __align__(16) TriRec {
  float g[9];         // grid-space vertices
  float nrm[3];       // e0 x e1 from the grid vertices (valid with kPlaneHoist)
  float rp;           // 0.5 |n|_1, summed as the reference sums it (valid with kPlaneHoist)
  uint32_t mnxy;      // bounding box min: x | y << 16
  uint32_t mnz_ex;    // bounding box min z | decoded box extent along its first axis << 16
  uint32_t exy;       // decoded box: extent of the first two axes multiplied
  uint32_t flags;     // kLiteral | kWide | kPlaneHoist | kWindow
  float rcp_ex, rcp_exy;  // 1/ex, 1/exy for the narrow decode
  uint32_t grid_base;  // mesh's first leaf-grid word / 512 (0 for a single mesh)
};
static_assert(sizeof(TriRec) == 80, "TriRec is one 80-byte record");
static_assert(offsetof(TriRec, nrm) == 36 && offsetof(TriRec, rp) == 48 &&
                  offsetof(TriRec, mnxy) == 52 && offsetof(TriRec, mnz_ex) == 56 &&
                  offsetof(TriRec, exy) == 60,
              "the SAT walk reads g, nrm and rp..exy as 16-byte blocks 0, 2 and 3");

// The window of a kWindow triangle, in a side array beside the records (written for
// windowed triangles only, so the others cost no bytes).
struct __align__(16) WinRec {
  float k0, ka, kb;   // window start along c
  uint32_t cinfo;     // c | W << 2 | bounding-box extent along c << 6
};

// True when g - c is exact in binary32 for every candidate centre c = m + 1/2, lo <= m <= hi.
// g is a multiple of u = 2^(E-23) (E its exponent); with E <= 22 the centres are multiples
// of u too, so g - c is a multiple of u, and any multiple of u below 2^24 u = 2^(E+1) in
// magnitude is a binary32 number. The largest |g - c| is at an end of the range.
__device__ __forceinline__ bool sub_exact(float g, int32_t lo, int32_t hi) {
  if (g == 0.0f) return true;
  const int32_t e = int32_t((__float_as_uint(g) >> 23) & 0xffu) - 127;
  if (e < -126 || e > 22) return false;  // subnormal, or ulp(g) above 1/2
  const double lim = ldexp(1.0, e + 1), gd = double(g);
  return fabs(gd - (double(lo) + 0.5)) < lim && fabs(gd - (double(hi) + 0.5)) < lim;
}

__device__ __forceinline__ int32_t sat_i32(float v) {
  // Rust `f32 as i32`: saturating, NaN -> 0
  if (v != v) return 0;
  if (v >= 2147483648.0f) return INT32_MAX;
  if (v <= -2147483648.0f) return INT32_MIN;
  return (int32_t)v;
}

__device__ __forceinline__ void tri_grid(const float *__restrict__ tri, const VoxelizeParams &p,
                                         float g[9]) {
  if (p.has_mat) {
    // Mat4::transform_point3 of the override: x_axis * x, then + y_axis * y, + z_axis * z,
    // + w_axis, per component (glam's column order; every product and sum rounded)
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
      for (int a = 0; a < 3; a++) {
        float s = p.m[a] * tri[3 * v];
        s = s + p.m[3 + a] * tri[3 * v + 1];
        s = s + p.m[6 + a] * tri[3 * v + 2];
        g[3 * v + a] = s + p.m[9 + a];
      }
    return;
  }
#pragma unroll
  for (int v = 0; v < 3; v++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      float s = p.inv * tri[3 * v + a];
      g[3 * v + a] = s + p.tx[a];
    }
}

// One pass per triangle: grid-space vertices, candidate box, candidate count, and the
// 64-byte record the SAT kernel reads. Triangles with no candidate keep count 0.
// *count receives the number of candidates this record decodes; returns how many fewer
// that is than the reference's count (the clamped bounding-box volume, as
// reference_cpu.rs:139-154 enumerates): 0 unless the triangle is windowed.
__device__ __forceinline__ uint64_t setup_one(const float *__restrict__ tri,
                                              const VoxelizeParams &p, uint32_t grid_base,
                                              uint64_t *__restrict__ count, uint64_t &decoded_out,
                                              TriRec *__restrict__ rec,
                                              WinRec *__restrict__ wrec) {
  TriRec r;
  tri_grid(tri, p, r.g);
  bool finite = true;
#pragma unroll
  for (int i = 0; i < 9; i++) finite = finite && isfinite(r.g[i]);
  uint32_t ext[3];
  int32_t r_mn[3];
  uint64_t c = 1;
  const float hi_clamp = (float)(p.n - 1);
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float lo = fminf(fminf(r.g[a], r.g[3 + a]), r.g[6 + a]) - p.eps;
    const float hi = fmaxf(fmaxf(r.g[a], r.g[3 + a]), r.g[6 + a]) + p.eps;
    const int32_t mn = sat_i32(fmaxf(floorf(lo), 0.0f));
    const int32_t mx = sat_i32(fminf(floorf(hi), hi_clamp));
    c = mx < mn ? 0 : c * uint64_t(int64_t(mx) - int64_t(mn) + 1);
    r_mn[a] = mx < mn ? 0 : mn;
    ext[a] = mx < mn ? 1u : uint32_t(mx - mn + 1);
  }
  *count = c;
  decoded_out = c;
  if (c == 0) return 0;  // never read
  r.mnxy = uint32_t(r_mn[0]) | (uint32_t(r_mn[1]) << 16);
  r.mnz_ex = uint32_t(r_mn[2]) | (ext[0] << 16);
  r.exy = ext[0] * ext[1];
  r.flags = finite ? 0u : kLiteral;
  // When every v_i = g_i - c is exact for all candidate centres, the reference's
  // centre-relative edges (v1 - v0, v2 - v1) round to the same values as g1 - g0 and
  // g2 - g1, so the plane normal and its radius are per-triangle constants; only
  // d = -(n . v0) depends on the voxel.
  bool hoist = finite;
#pragma unroll
  for (int v = 0; v < 3; v++)
#pragma unroll
    for (int a = 0; a < 3; a++)
      hoist = hoist && sub_exact(r.g[3 * v + a], r_mn[a], r_mn[a] + int32_t(ext[a]) - 1);
  {
    const float e0x = r.g[3] - r.g[0], e0y = r.g[4] - r.g[1], e0z = r.g[5] - r.g[2];
    const float e1x = r.g[6] - r.g[3], e1y = r.g[7] - r.g[4], e1z = r.g[8] - r.g[5];
    r.nrm[0] = e0y * e1z - e1y * e0z;
    r.nrm[1] = e0z * e1x - e1z * e0x;
    r.nrm[2] = e0x * e1y - e1x * e0y;
    float rp = 0.5f * fabsf(r.nrm[0]);
    rp = rp + 0.5f * fabsf(r.nrm[1]);
    rp = rp + 0.5f * fabsf(r.nrm[2]);
    r.rp = rp;
    hoist = hoist && isfinite(r.nrm[0]) && isfinite(r.nrm[1]) && isfinite(r.nrm[2]);
  }
  if (hoist) r.flags |= kPlaneHoist;
  uint64_t decoded = c;
  // Window (hoisted triangles whose box is at least twice the window): with c the axis of
  // the normal's largest component, a cell passes the plane test only if its centre lies
  // within R / |n_c| of the plane along c, where R bounds the test's radius plus the
  // rounding of |d| (relative 1e-4 of rp and of sum |n_i| |v0_i| over the box, orders of
  // magnitude above the few-ulp errors of the binary32 sums). The window of a column starts
  // at lo = floor(c0 - hw - 1/2), c0 the plane's c at the column centre as binary32 gives
  // it (error e) and hw = R / |n_c| + s, slack s > e; then lo is at or below the lowest
  // passing cell, and W >= 2 hw - s + e + 2 cells reach the highest one: W = ceil(2 hw +
  // 1.95) with e < 1/16 when s = 1/8. The start k0 + ka a + kb b takes about ten roundings
  // of values below mag (the sum of the magnitudes it adds), so e < mag 2^-20: s = 1/8 for
  // mag < 2^16, s = 1 for mag < 2^19, and no window beyond. Cells outside the bounding box
  // are skipped when decoded; the rest of the box is never enumerated.
  // (axis picks are selects: a runtime index into r or ext would put them in scratch)
  const float an0 = fabsf(r.nrm[0]), an1 = fabsf(r.nrm[1]), an2 = fabsf(r.nrm[2]);
  const bool c1 = an1 > an0 && an1 >= an2, c2 = !c1 && an2 > an0 && an2 > an1;
  const uint32_t ext_c = c1 ? ext[1] : (c2 ? ext[2] : ext[0]);
  if (hoist && ext_c >= 6) {
    const float nc = c1 ? r.nrm[1] : (c2 ? r.nrm[2] : r.nrm[0]);
    if (nc != 0.0f) {
      float S = 0.0f;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const float lo = (float(r_mn[i]) + 0.5f) - r.g[i];
        const float hi = (float(r_mn[i] + int32_t(ext[i]) - 1) + 0.5f) - r.g[i];
        S += fabsf(r.nrm[i]) * fmaxf(fabsf(lo), fabsf(hi));
      }
      const float R = r.rp * (1.0f + 1e-4f) + 1e-4f * S;
      // the other two axes in increasing order: a = first, b = second
      const bool c0 = !c1 && !c2;
      const uint32_t ext_a = c0 ? ext[1] : ext[0], ext_b = c2 ? ext[1] : ext[2];
      const float na = c0 ? r.nrm[1] : r.nrm[0], nb = c2 ? r.nrm[1] : r.nrm[2];
      const float ga = c0 ? r.g[1] : r.g[0], gb = c2 ? r.g[1] : r.g[2];
      const float gc = c1 ? r.g[1] : (c2 ? r.g[2] : r.g[0]);
      const int32_t ma = c0 ? r_mn[1] : r_mn[0], mb = c2 ? r_mn[1] : r_mn[2];
      const float ka = -na / nc, kb = -nb / nc;
      const float tc = (na * (0.5f - ga) + nb * (0.5f - gb)) / nc;
      const float mag = fabsf(gc) + fabsf(tc) + fabsf(ka) * float(ma + int32_t(ext_a)) +
                        fabsf(kb) * float(mb + int32_t(ext_b)) + R / fabsf(nc) + 2.0f;
      const float slack = mag < 65536.0f ? 0.125f : 1.0f;
      const float hw = R / fabsf(nc) + slack;
      const float wd = ceilf(2.0f * hw + 1.95f);
      if (wd <= 15.0f && mag < 524288.0f) {
        const uint32_t W = uint32_t(wd);
        const uint64_t decw = uint64_t(W) * ext_a * ext_b;
        if (2 * decw <= c) {
          const uint32_t ax = c1 ? 1u : (c2 ? 2u : 0u);
          WinRec w;
          w.ka = ka;
          w.kb = kb;
          w.k0 = gc - tc - hw - 0.5f;
          w.cinfo = ax | (W << 2) | (ext_c << 6);
          *wrec = w;
          r.mnz_ex = uint32_t(r_mn[2]) | (W << 16);
          r.exy = W * ext_a;
          r.flags |= kWindow;
          decoded = decw;
        }
      }
    }
  }
  // narrow decode: every local index < 2^22, so the float quotient is off by at most 1
  if (decoded >> 22) r.flags |= kWide;
  *count = decoded;
  decoded_out = decoded;
  r.rcp_ex = 1.0f / float(r.mnz_ex >> 16);
  r.rcp_exy = 1.0f / float(r.exy);
  r.grid_base = grid_base;
  *rec = r;
  return c - decoded;
}

// The cells the wave's windows skip (0 for a lane without a windowed triangle), added once
// to acc[slot].
__device__ __forceinline__ void note_windows(unsigned long long *acc, uint64_t v,
                                             uint32_t slot = 0) {
  if (__ballot(v != 0) == 0) return;  // wave-uniform: the common case
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += (uint64_t)__shfl_xor((long long)v, d, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(acc + slot, (unsigned long long)v);
}

constexpr int kSetupTB = 256;  // setup threads per block (one triangle each)

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) v += uint64_t(__shfl_xor((long long)v, d, 64));
  return v;
}

// Setup of one block's triangles: `tris`, `count`, `rec`, `wrec` point at the block's first
// triangle, count, record and window record; nt <= kSetupTB triangles are valid. The
// triangles come in through LDS with coalesced loads (lane k reads float k), and the
// finished 80-byte records go out through the same LDS so that the block stores them as
// consecutive 16-byte lanes instead of one 80-byte stride per lane. The block's candidate
// total goes to *block_sum. Returns the lane's skipped-cell count (setup_one).
__device__ __forceinline__ uint64_t setup_block(const float *__restrict__ tris, uint32_t nt,
                                                const VoxelizeParams &p, uint32_t grid_base,
                                                uint64_t *__restrict__ count,
                                                TriRec *__restrict__ rec,
                                                WinRec *__restrict__ wrec,
                                                uint64_t *__restrict__ block_sum) {
  constexpr uint32_t kStage = kSetupTB * sizeof(TriRec) / 16;
  static_assert(kSetupTB * 9 * 4 <= kStage * 16, "the triangles fit the record stage");
  __shared__ uint4 stage[kStage];
  float *tin = reinterpret_cast<float *>(stage);
  const uint32_t i = threadIdx.x;
  for (uint32_t k = i; k < nt * 9; k += kSetupTB) tin[k] = tris[k];
  __syncthreads();
  float tri[9];
  if (i < nt) {
#pragma unroll
    for (int k = 0; k < 9; k++) tri[k] = tin[9 * i + k];
  }
  __syncthreads();
  uint64_t skipped = 0, decoded = 0;
  // a triangle without candidates leaves its slot unwritten: its record is never read
  if (i < nt)
    skipped = setup_one(tri, p, grid_base, count + i, decoded,
                        reinterpret_cast<TriRec *>(stage) + i, wrec + i);
  __shared__ uint64_t wave_sums[kSetupTB / 64];
  decoded = wave_sum_u64(decoded);
  if ((i & 63) == 0) wave_sums[i >> 6] = decoded;
  __syncthreads();
  if (i == 0) {
    uint64_t b = 0;
#pragma unroll
    for (int w = 0; w < kSetupTB / 64; w++) b += wave_sums[w];
    *block_sum = b;
  }
  uint4 *dst = reinterpret_cast<uint4 *>(rec);
  for (uint32_t k = i; k < nt * (sizeof(TriRec) / 16); k += kSetupTB) dst[k] = stage[k];
  return skipped;
}

__global__ __launch_bounds__(kSetupTB) void k_tri_setup(
    const float *__restrict__ tris, uint64_t T, VoxelizeParams p, uint64_t *__restrict__ count,
    TriRec *__restrict__ rec, WinRec *__restrict__ wrec, uint64_t *__restrict__ block_sums,
    unsigned long long *__restrict__ skipped_acc) {
  const uint64_t t0 = uint64_t(blockIdx.x) * kSetupTB;
  const uint32_t nt = uint32_t(min(uint64_t(kSetupTB), T - t0));
  const uint64_t skipped = setup_block(tris + 9 * t0, nt, p, 0u, count + t0, rec + t0,
                                       wrec + t0, block_sums + blockIdx.x);
  note_windows(skipped_acc, skipped, blockIdx.x % kSkipSlots);
}

// Many meshes: mesh g owns setup blocks [sb_begin[g], sb_begin[g+1]) and record slots
// from sb_begin[g] * kSetupTB on; the slots past its triangle count hold count 0.
__global__ __launch_bounds__(kSetupTB) void k_tri_setup_batch(
    const BatchMesh *__restrict__ mesh,
    const uint32_t *__restrict__ sb_begin, uint32_t G, uint64_t *__restrict__ count,
    TriRec *__restrict__ rec, WinRec *__restrict__ wrec, uint64_t *__restrict__ block_sums,
    unsigned long long *__restrict__ ref_pairs) {
  const uint32_t g = dev::upper_index(sb_begin, G, blockIdx.x);
  const BatchMesh &m = mesh[g];
  const uint64_t tl0 = uint64_t(blockIdx.x - sb_begin[g]) * kSetupTB;
  const uint64_t slot0 = uint64_t(blockIdx.x) * kSetupTB;
  const uint32_t nt = tl0 < m.T ? uint32_t(min(uint64_t(kSetupTB), m.T - tl0)) : 0u;
  if (threadIdx.x >= nt) count[slot0 + threadIdx.x] = 0;
  const uint64_t skipped = setup_block(m.tris + 9 * tl0, nt, m.p, uint32_t(m.b1_begin),
                                       count + slot0, rec + slot0, wrec + slot0,
                                       block_sums + blockIdx.x);
  note_windows(ref_pairs + g, skipped);  // a block belongs to one mesh
}

constexpr uint64_t kChunksPerUnit = 16;  // at least 1024 pairs per work unit

// The SAT launch plan, computed on the device from the pair total so that no host
// read-back sits between the scan and the SAT kernel: work units of `cpu` chunks of 64
// pairs (at least kChunksPerUnit, about `want` units in all), and the work counter.
__global__ void k_sat_plan(const uint64_t *__restrict__ total, uint64_t want,
                           SatPlan *__restrict__ plan, uint32_t epoch) {
  const uint64_t P = *total;
  const uint64_t nchunks = (P + 63) / 64;
  uint64_t cpu = (nchunks + want - 1) / want;
  if (cpu < kChunksPerUnit) cpu = kChunksPerUnit;
  plan->P = P;
  unsigned long long skip = 0;
  for (int i = 0; i < kSkipSlots; i++) {
    skip += plan->skipped[epoch & 1][i];
    plan->skipped[(epoch + 1) & 1][i] = 0;
  }
  plan->P_ref = P + skip;
  plan->cpu = cpu;
  plan->nunits = (nchunks + cpu - 1) / cpu;
  for (int c = 0; c < kUnitCounters; c++) plan->counter[c][0] = 0;
}

// ustart[u] = triangle holding the first pair of unit u (triangles with no candidate never
// do).
__global__ void k_bstart(const uint64_t *__restrict__ coff, const uint64_t *__restrict__ ccount,
                         uint64_t Tc, const SatPlan *__restrict__ plan,
                         uint32_t *__restrict__ ustart) {
  uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= Tc) return;
  const uint64_t chunk = plan->cpu * 64, nunits = plan->nunits;
  uint64_t lo = coff[t], hi = lo + ccount[t];
  uint64_t b = (lo + chunk - 1) / chunk;
  for (; b * chunk < hi && b < nunits; b++) ustart[b] = uint32_t(t);
}

// ---- pair offsets and the plan from the setup blocks' sums -----------------------------
// The setup kernels leave each block's candidate total in block_sums. When there are at
// most kFusedBlocks of them, one workgroup scans those totals and derives the plan from
// the grand total (k_setup_plan), and one pass over the counts writes the pair offsets and
// the work-unit starts (k_offsets_units): three launches where a device-wide scan, the
// plan and the unit-start pass take six.
constexpr uint32_t kPlanThreads = 1024;
constexpr uint64_t kFusedBlocks = 65536;

// Exclusive scan of one value per thread over a block of NT threads; returns the total.
template <int NT>
__device__ __forceinline__ uint64_t block_excl_scan_u64(uint64_t v, uint64_t *excl) {
  __shared__ uint64_t wave_tot[NT / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = uint64_t(__shfl_up((long long)inc, d, 64));
    if (lane >= d) inc += o;
  }
  if (lane == 63) wave_tot[wave] = inc;
  __syncthreads();
  uint64_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) {
    const uint64_t t = wave_tot[w];
    if (w < wave) before += t;
    total += t;
  }
  *excl = before + inc - v;
  return total;
}

// One workgroup: block_sums[0..nb) -> their exclusive prefix sums in place, the grand total
// P to *total (= the offset past the last triangle), and the plan (k_sat_plan's fields).
__global__ __launch_bounds__(kPlanThreads) void k_setup_plan(uint64_t *__restrict__ block_sums,
                                                             uint64_t nb, uint64_t want,
                                                             SatPlan *__restrict__ plan,
                                                             uint32_t epoch,
                                                             uint64_t *__restrict__ total) {
  // each thread owns `per` consecutive sums, loaded kHeld at a time as independent loads
  // (one memory round trip per kHeld sums); up to kHeld of them stay in registers between
  // the two passes
  constexpr int kHeld = 16;
  const uint64_t per = (nb + kPlanThreads - 1) / kPlanThreads;
  const uint64_t b0 = min(uint64_t(threadIdx.x) * per, nb), b1 = min(b0 + per, nb);
  uint64_t held[kHeld];
  uint64_t s = 0;
  for (uint64_t b = b0; b < b1; b += kHeld) {
#pragma unroll
    for (int k = 0; k < kHeld; k++) held[k] = b + k < b1 ? block_sums[b + k] : 0;
#pragma unroll
    for (int k = 0; k < kHeld; k++) s += held[k];
  }
  uint64_t run = 0;
  const uint64_t P = block_excl_scan_u64<kPlanThreads>(s, &run);
  for (uint64_t b = b0; b < b1; b += kHeld) {
    if (per > kHeld) {  // more than one batch: reload (the held values are the last batch)
#pragma unroll
      for (int k = 0; k < kHeld; k++) held[k] = b + k < b1 ? block_sums[b + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < kHeld; k++) {
      if (b + k < b1) block_sums[b + k] = run;
      run += held[k];
    }
  }
  // the first wave: the skipped cells of this call's windows (one slot per lane), and the
  // other epoch's slots cleared for the next call
  static_assert(kSkipSlots == 64, "one skipped-cell slot per lane of the first wave");
  if (threadIdx.x < 64) {
    const uint64_t skip = wave_sum_u64(plan->skipped[epoch & 1][threadIdx.x]);
    plan->skipped[(epoch + 1) & 1][threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      *total = P;
      const uint64_t nchunks = (P + 63) / 64;
      uint64_t cpu = (nchunks + want - 1) / want;
      if (cpu < kChunksPerUnit) cpu = kChunksPerUnit;
      plan->P = P;
      plan->P_ref = P + skip;
      plan->cpu = cpu;
      plan->nunits = (nchunks + cpu - 1) / cpu;
      for (int c = 0; c < kUnitCounters; c++) plan->counter[c][0] = 0;
    }
  }
}

// kOffsetBlocks setup blocks' counts per workgroup, one setup block per wave: the
// wave-local exclusive scan of each lane's kOffsetItems consecutive counts plus the setup
// block's offset is each triangle's first pair (coff); a triangle whose pairs hold the
// first pair of work units is their start triangle (k_bstart's rule).
constexpr int kOffsetItems = kSetupTB / 64;  // counts per lane: one setup block per wave
constexpr int kOffsetBlocks = 4;              // setup blocks (waves) per workgroup
__global__ __launch_bounds__(64 * kOffsetBlocks) void k_offsets_units(
    const uint64_t *__restrict__ count, uint64_t *__restrict__ coff,
    const uint64_t *__restrict__ block_offsets, uint64_t nb, uint64_t Tc,
    const SatPlan *__restrict__ plan, uint32_t *__restrict__ ustart) {
  const uint64_t sb = uint64_t(blockIdx.x) * kOffsetBlocks + (threadIdx.x >> 6);
  if (sb >= nb) return;  // wave-uniform
  const int lane = threadIdx.x & 63;
  const uint64_t t0 = sb * kSetupTB + uint64_t(lane) * kOffsetItems;
  uint64_t c[kOffsetItems], s = 0;
#pragma unroll
  for (int k = 0; k < kOffsetItems; k++) {
    c[k] = t0 + k < Tc ? count[t0 + k] : 0;
    s += c[k];
  }
  uint64_t inc = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = uint64_t(__shfl_up((long long)inc, d, 64));
    if (lane >= d) inc += o;
  }
  uint64_t lo = block_offsets[sb] + inc - s;
  const uint64_t chunk = plan->cpu * 64, nunits = plan->nunits;
  // the first unit starting at or after the lane's first pair (one division per lane; the
  // lane's triangles are consecutive, so b only moves forward)
  uint64_t b = (lo + chunk - 1) / chunk, bstart = b * chunk;
#pragma unroll
  for (int k = 0; k < kOffsetItems; k++) {
    const uint64_t t = t0 + k;
    if (t < Tc) {
      coff[t] = lo;
      const uint64_t hi = lo + c[k];
      for (; bstart < hi && b < nunits; b++, bstart += chunk) ustart[b] = uint32_t(t);
    }
    lo += c[k];
  }
}

// ---- the SAT predicate, reference_cpu.rs:27-85 -----------------------------------------
// Split in two conjuncts evaluated in separate phases: the plane test (:76-82) and the
// rest (box faces :66-74 and the 9 edge axes :42-64). Their conjunction is the reference
// predicate; each conjunct uses the reference's operations in the reference's order.
struct Rel {
  float v0x, v0y, v0z, v1x, v1y, v1z, v2x, v2y, v2z;
};

__device__ __forceinline__ Rel rel_to(const float *g, float cx, float cy, float cz) {
  Rel r;
  r.v0x = g[0] - cx; r.v0y = g[1] - cy; r.v0z = g[2] - cz;
  r.v1x = g[3] - cx; r.v1y = g[4] - cy; r.v1z = g[5] - cz;
  r.v2x = g[6] - cx; r.v2y = g[7] - cy; r.v2z = g[8] - cz;
  return r;
}

template <bool kLiteral>
__device__ __forceinline__ bool plane_passes(const Rel &v) {
  const float h = 0.5f;
  const float e0x = v.v1x - v.v0x, e0y = v.v1y - v.v0y, e0z = v.v1z - v.v0z;
  const float e1x = v.v2x - v.v1x, e1y = v.v2y - v.v1y, e1z = v.v2z - v.v1z;
  const float nx = e0y * e1z - e1y * e0z;
  const float ny = e0z * e1x - e1z * e0x;
  const float nz = e0x * e1y - e1x * e0y;
  float dd = nx * v.v0x;
  dd = dd + ny * v.v0y;
  dd = dd + nz * v.v0z;
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
  return !(fabsf(s) > r);
}

// Finite vertices: the 12 axes as one straight-line conjunction (a wave rarely has all
// lanes fail the same early axis, so early exits only cost branches). In the reference's
// radius 0.5|ax| + 0.5|ay| + 0.5|az| one axis component is the constant 0; 0.5|0| = +0
// and adding +0 to a value that is >= +0 or NaN returns it unchanged, so each radius is
// the sum of two of the nine per-edge terms 0.5|e_c|, formed once.
__device__ __forceinline__ bool rest_passes_finite(const Rel &v) {
  const float h = 0.5f;
  const float v0x = v.v0x, v0y = v.v0y, v0z = v.v0z;
  const float v1x = v.v1x, v1y = v.v1y, v1z = v.v1z;
  const float v2x = v.v2x, v2y = v.v2y, v2z = v.v2z;
  bool ok = true;
  // box-face axes (:66-74)
  ok &= !(fminf(v0x, fminf(v1x, v2x)) > h || fmaxf(v0x, fmaxf(v1x, v2x)) < -h);
  ok &= !(fminf(v0y, fminf(v1y, v2y)) > h || fmaxf(v0y, fmaxf(v1y, v2y)) < -h);
  ok &= !(fminf(v0z, fminf(v1z, v2z)) > h || fmaxf(v0z, fmaxf(v1z, v2z)) < -h);
  const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
  const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
  const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;
  const float h0x = h * fabsf(e0x), h0y = h * fabsf(e0y), h0z = h * fabsf(e0z);
  const float h1x = h * fabsf(e1x), h1y = h * fabsf(e1y), h1z = h * fabsf(e1z);
  const float h2x = h * fabsf(e2x), h2y = h * fabsf(e2y), h2z = h * fabsf(e2z);
  // the 9 edge x box-axis tests (:42-64), p_i = v_i . axis without the zero component
#define VXL_AXIS2(P0, P1, P2, R)                                                      \
  {                                                                                   \
    const float p0 = (P0), p1 = (P1), p2 = (P2), r = (R);                             \
    ok &= !(fminf(p0, fminf(p1, p2)) > r || fmaxf(p0, fmaxf(p1, p2)) < -r);           \
  }
  // axes (0, -e.z, e.y)
  VXL_AXIS2(v0y * -e0z + v0z * e0y, v1y * -e0z + v1z * e0y, v2y * -e0z + v2z * e0y, h0z + h0y)
  VXL_AXIS2(v0y * -e1z + v0z * e1y, v1y * -e1z + v1z * e1y, v2y * -e1z + v2z * e1y, h1z + h1y)
  VXL_AXIS2(v0y * -e2z + v0z * e2y, v1y * -e2z + v1z * e2y, v2y * -e2z + v2z * e2y, h2z + h2y)
  // axes (e.z, 0, -e.x)
  VXL_AXIS2(v0x * e0z + v0z * -e0x, v1x * e0z + v1z * -e0x, v2x * e0z + v2z * -e0x, h0z + h0x)
  VXL_AXIS2(v0x * e1z + v0z * -e1x, v1x * e1z + v1z * -e1x, v2x * e1z + v2z * -e1x, h1z + h1x)
  VXL_AXIS2(v0x * e2z + v0z * -e2x, v1x * e2z + v1z * -e2x, v2x * e2z + v2z * -e2x, h2z + h2x)
  // axes (-e.y, e.x, 0)
  VXL_AXIS2(v0x * -e0y + v0y * e0x, v1x * -e0y + v1y * e0x, v2x * -e0y + v2y * e0x, h0y + h0x)
  VXL_AXIS2(v0x * -e1y + v0y * e1x, v1x * -e1y + v1y * e1x, v2x * -e1y + v2y * e1x, h1y + h1x)
  VXL_AXIS2(v0x * -e2y + v0y * e2x, v1x * -e2y + v1y * e2x, v2x * -e2y + v2y * e2x, h2y + h2x)
#undef VXL_AXIS2
  return ok;
}

template <bool kLiteral>
__device__ __forceinline__ bool rest_passes(const Rel &v) {
  const float h = 0.5f;
  const float v0x = v.v0x, v0y = v.v0y, v0z = v.v0z;
  const float v1x = v.v1x, v1y = v.v1y, v1z = v.v1z;
  const float v2x = v.v2x, v2y = v.v2y, v2z = v.v2z;
  // box-face axes (:66-74)
  if (fminf(v0x, fminf(v1x, v2x)) > h || fmaxf(v0x, fmaxf(v1x, v2x)) < -h) return false;
  if (fminf(v0y, fminf(v1y, v2y)) > h || fmaxf(v0y, fmaxf(v1y, v2y)) < -h) return false;
  if (fminf(v0z, fminf(v1z, v2z)) > h || fmaxf(v0z, fmaxf(v1z, v2z)) < -h) return false;
  const float e0x = v1x - v0x, e0y = v1y - v0y, e0z = v1z - v0z;
  const float e1x = v2x - v1x, e1y = v2y - v1y, e1z = v2z - v1z;
  const float e2x = v0x - v2x, e2y = v0y - v2y, e2z = v0z - v2z;
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

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)uint32_t(v), lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)uint32_t(v >> 32), lane);
  return (uint64_t(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = (uint32_t)__shfl((int)uint32_t(v), src, 64);
  uint32_t hi = (uint32_t)__shfl((int)uint32_t(v >> 32), src, 64);
  return (uint64_t(hi) << 32) | lo;
}

// Inclusive sum of v over lanes 0..lane of the wave, by row shifts within each row of 16
// lanes and the two row broadcasts (lane 15 to row 1, lane 47 to row 3; then lane 31 to
// rows 2 and 3). Lanes a shift or broadcast does not reach add 0. No lane addresses are
// formed, so nothing is held in registers between calls.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int x = int(v);
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return uint32_t(x);
}

// Low 32 bits of the product of the low 24 bits of a and b (one full-rate multiply), as
// an unsigned value: HIP's __umul24 returns int, which would make shifts of it arithmetic.
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) {
  return static_cast<uint32_t>(__umul24(a, b));
}

// quotient of a < 2^22 by b, from a float reciprocal, corrected to the exact value; every
// product is below 2^24, so the 24-bit multiply is exact
__device__ __forceinline__ uint32_t udiv_narrow(uint32_t a, uint32_t b, float rb) {
  uint32_t q = uint32_t(float(a) * rb);
  const int32_t r = int32_t(a - mul24(q, b));
  if (r < 0) q -= 1;
  else if (uint32_t(r) >= b) q += 1;
  return q;
}

constexpr int kWaves = 4;    // waves per block
constexpr int kQueue = 128;  // survivor queue entries per wave
constexpr int kTable = 128;  // leaf-word merge table slots per wave
constexpr int kTableBits = 7;
// The table is flushed once `flush_at` slots are in use (a launch argument, 1..kFlushMax):
// an insert adds at most 64 keys, so at most kFlushMax - 1 + 64 < kTable slots are ever in
// use and a probe always finds a free or matching slot.
constexpr int kFlushMax = 64;
constexpr unsigned long long kEmpty = ~0ull;
static_assert(kTable == 1 << kTableBits, "the slot hash yields kTableBits bits");
static_assert(kFlushMax - 1 + 64 < kTable, "a flush threshold of kFlushMax keeps a free slot");

// Merge-table key of a hit voxel: its mesh's grid base and its 4^3 sub-brick (one leaf-grid
// word) as packed coordinates, sx | sy << 11 | sz << 22 | grid_base << 33. Sub-brick
// coordinates are below 2048 (n <= 8192). Packing costs a few bit operations per hit;
// the Morton spreading to the leaf-grid word index (key_word) runs once per table entry
// at flush time instead.
__device__ __forceinline__ unsigned long long pack_key(uint32_t grid_base, uint32_t x, uint32_t y,
                                                       uint32_t z) {
  const uint32_t sx = x >> 2, sy = y >> 2, sz = z >> 2;
  const uint32_t lo = sx | (sy << 11) | (sz << 22);
  const uint32_t hi = (grid_base << 1) | (sz >> 10);
  return (uint64_t(hi) << 32) | lo;
}

// Leaf-grid word index of a packed key: the mesh's first word (grid_base * 512) + the
// brick's Morton code * 8 + the sub-brick's octant in the brick.
__device__ __forceinline__ unsigned long long key_word(unsigned long long key) {
  const uint32_t lo = uint32_t(key), hi = uint32_t(key >> 32);
  const uint32_t sx = lo & 0x7ffu, sy = (lo >> 11) & 0x7ffu, sz = (lo >> 22) | ((hi & 1u) << 10);
  const uint32_t code = dev::brick_code(sx >> 1, sy >> 1, sz >> 1);
  const uint32_t oct = (sx & 1u) | ((sy & 1u) << 1) | ((sz & 1u) << 2);
  return (uint64_t(hi >> 1) << 9) + uint64_t(code) * 8 + oct;
}

// Phase B on one survivor: box faces + edge axes. On a hit, returns the packed merge-table
// key of the voxel's leaf-grid word and the voxel's bit in that word (its 6-bit Morton
// index inside the 4^3 sub-brick).
__device__ __forceinline__ bool finish_pair(const TriRec *__restrict__ rec, uint32_t t, uint32_t xy,
                                            uint32_t z, unsigned long long &key,
                                            unsigned long long &bit) {
  const TriRec r = rec[t];
  const uint32_t x = xy & 0xffffu, y = xy >> 16;
  const Rel v = rel_to(r.g, float(x) + 0.5f, float(y) + 0.5f, float(z) + 0.5f);
  const bool hit = (r.flags & kLiteral) ? rest_passes<true>(v) : rest_passes_finite(v);
  key = pack_key(r.grid_base, x, y, z);
  bit = 1ull << dev::child_bit(x, y, z);
  return hit;
}

// Merges hit bits into the wave's LDS table (open addressing, linear probing).
// Returns the number of slots this call newly occupied (wave-uniform).
__device__ __forceinline__ uint32_t table_insert(unsigned long long *tk, unsigned long long *tv,
                                                 bool hit, unsigned long long key,
                                                 unsigned long long bit) {
  bool fresh = false;
  if (hit) {
    // Fibonacci hash of the key folded to 24 bits: the top kTableBits of
    // (h * floor(2^24 / golden ratio)) mod 2^24, from one full-rate 24-bit multiply
    const uint32_t lo = uint32_t(key), h = lo ^ (lo >> 15) ^ uint32_t(key >> 32);
    uint32_t slot = (mul24(h, 0x9E3779u) >> (24 - kTableBits)) & (kTable - 1);
    for (;;) {
      const unsigned long long old = atomicCAS(&tk[slot], kEmpty, key);
      if (old == kEmpty || old == key) {
        atomicOr(&tv[slot], bit);
        fresh = old == kEmpty;
        break;
      }
      slot = (slot + 1) & (kTable - 1);
    }
  }
  __builtin_amdgcn_wave_barrier();
  return __popcll(__ballot(fresh));
}

// ORs `bits` into *p unless a coherent read already shows them all. During the SAT bits
// only ever go from 0 to 1, so a read can only under-report what is set: skipping is
// exact, and it keeps hot words (dense meshes on small grids hit a few thousand leaf words
// and a handful of bitmap words from every wave) off the atomic units.
__device__ __forceinline__ void or_if_missing(unsigned long long *p, unsigned long long bits) {
  const unsigned long long cur =
      __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((cur & bits) != bits) atomicOr(p, bits);
}

// Owner modes of k_sat: none, a dense n^3 array, or per leaf voxel of a built tree.
constexpr int kOwnNone = 0, kOwnDense = 1, kOwnLeaf = 2;

// Lowers the owner of voxel (x, y, z) to triangle t (the reference keeps the lowest index).
template <int kOwn>
__device__ __forceinline__ void owner_min(uint32_t *__restrict__ owner,
                                          const uint32_t *__restrict__ leaf_map, uint32_t n,
                                          uint32_t t, uint32_t xy, uint32_t z) {
  const uint32_t x = xy & 0xffffu, y = xy >> 16;
  uint64_t lin;
  if (kOwn == kOwnDense) {
    lin = uint64_t(x) + uint64_t(n) * (uint64_t(y) + uint64_t(n) * uint64_t(z));
  } else {
    const uint32_t leaf = leaf_map[dev::brick_code(x >> 3, y >> 3, z >> 3)];
    lin = uint64_t(leaf) * 512 + dev::morton9(x, y, z);
  }
  const uint32_t cur = __hip_atomic_load(owner + lin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (t < cur) atomicMin(owner + lin, t);
}

// Writes the table to the leaf grid and the brick bitmap, then empties it.
__device__ __forceinline__ void table_flush(unsigned long long *tk, unsigned long long *tv, int lane,
                                            unsigned long long *__restrict__ leafgrid,
                                            unsigned long long *__restrict__ brickbits) {
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int i = 0; i < kTable / 64; i++) {
    const int slot = i * 64 + lane;
    const unsigned long long key = tk[slot];
    if (key != kEmpty) {
      const unsigned long long word = key_word(key);
      or_if_missing(&leafgrid[word], tv[slot]);
      const uint64_t code = word >> 3;
      or_if_missing(&brickbits[code >> 6], 1ull << (code & 63));
      tk[slot] = kEmpty;
      tv[slot] = 0ull;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// At most kSegPairs pairs of one triangle per segment (a triangle's pairs in a unit split
// into that many segments of nearly equal size: 8 to 16 measured, 12 best on the terrain,
// whose boxes hold 9, 18 or 27 cells), and the largest pair range whose per-lane counts
// and segment numbers are taken in 32 bits.
constexpr uint32_t kSegPairs = 12;
constexpr uint64_t kSpanPairs = uint64_t(1) << 30;

// Each wave walks units of consecutive pairs [u0, u1): its own, then any left over from the
// counters (below). The unit's triangles are loaded 64 at a time, one per lane, with their
// pair ranges clipped
// to the unit; each range splits into segments of at most kSegPairs pairs, and the wave
// walks the segments 64 at a time, one per lane: the lane finds its segment's triangle
// by a binary search over the lanes' running segment counts, loads the triangle's record
// once, decodes the segment's first pair to its cell (two divisions) and then steps
// through the box cell by cell (first axis fastest, the reference's scan order) with
// adds. Phase A evaluates the plane test and queues survivors in LDS; every 64 queued
// survivors run phase B as a full wave. Hits are merged per leaf word in an LDS table and
// flushed to HBM in bulk. So the cost of finding a pair's triangle and cell is paid once
// per segment, not once per pair.
//
// The kernel is held to 64 VGPRs (8 waves per SIMD, the most a SIMD holds; the LDS of 8
// blocks, 8 x 16 KiB, fits a CU). Unit-level values (unit, pair range, first triangle)
// stay wave-uniform in scalar registers. The window decode is a branch that calls without
// windowed triangles never take.
//
// kOwn: also lower each hit voxel's owner to the triangle index (single-mesh calls, where
// the record index is the triangle index); kOwnLeaf writes owners only.
template <int kOwn>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(8))) void k_sat(
    const TriRec *__restrict__ rec, const WinRec *__restrict__ wrec,
    const uint64_t *__restrict__ coff,
    const uint32_t *__restrict__ ustart, uint64_t T, SatPlan *__restrict__ plan,
    unsigned long long *__restrict__ leafgrid, unsigned long long *__restrict__ brickbits,
    uint32_t *__restrict__ owner, const uint32_t *__restrict__ leaf_map,
    uint32_t own_n, uint32_t flush_at) {
  __shared__ uint4 q_p[kWaves][kQueue];  // (t, x | y << 16, z, -)
  __shared__ unsigned long long t_key[kWaves][kTable], t_val[kWaves][kTable];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t P = plan->P, chunks_per_unit = plan->cpu, nunits = plan->nunits;
  // the counter of this workgroup's class (every class has workgroups: grids hold far more
  // than kUnitCounters of them, and a smaller grid uses as many classes as it has)
  const uint32_t nctr = gridDim.x < uint32_t(kUnitCounters) ? gridDim.x : uint32_t(kUnitCounters);
  const uint32_t cls = blockIdx.x % nctr;
  unsigned int *next_unit = &plan->counter[cls][0];
  uint4 *qp2 = q_p[wave];
  unsigned long long *tk = t_key[wave], *tv = t_val[wave];
  // Every wave's first unit is its own (unit = the wave's index in the grid); units past
  // the grid's waves are handed out by atomic counters, and only when there are any. The
  // atomics on one counter address serialize: when all 8192 waves took their first unit
  // from one counter, the last ones started tens of microseconds late (terrain SAT 0.419 ->
  // 0.347 ms without it, the 16k-triangle sphere 0.136 -> 0.047 ms).
  // A wave without a unit leaves before touching LDS (the grid is sized for the device).
  const uint32_t nwaves = gridDim.x * kWaves;
  uint32_t unit = blockIdx.x * kWaves + uint32_t(wave);
  if (uint64_t(unit) >= nunits) return;  // wave-uniform
#pragma unroll
  for (int i = 0; i < kTable / 64; i++) {
    tk[i * 64 + lane] = kEmpty;
    tv[i * 64 + lane] = 0ull;
  }
  __builtin_amdgcn_wave_barrier();
  uint32_t qn = 0, used = 0;
  const unsigned long long below = (1ull << lane) - 1ull;

  for (bool first = true;; first = false) {
  if (!first) {
    if (uint64_t(nwaves) >= nunits) break;  // no units past the waves' own
    uint32_t k = 0;
    if (lane == 0) k = atomicAdd(next_unit, 1u);
    k = __builtin_amdgcn_readfirstlane(k);  // lane 0's value, in a scalar register
    const uint64_t u = uint64_t(nwaves) + cls + uint64_t(nctr) * k;
    if (u >= nunits) break;  // wave-uniform exit: every wave ends here
    unit = uint32_t(u);
  }
  const uint64_t u0 = uint64_t(unit) * chunks_per_unit * 64;
  const uint64_t u1 = min(u0 + chunks_per_unit * 64, P);
  uint32_t tb = ustart[unit];  // the triangle holding pair u0 (wave-uniform)
  // The unit's pairs [u0, u1) in batches of the 64 triangles tb .. tb + 63, clipped to spans
  // of at most kSpanPairs pairs so that per-lane counts stay 32-bit.
  for (uint64_t pos = u0; pos < u1;) {
    const uint64_t span_end = min(u1, pos + kSpanPairs);
    const uint64_t idx = uint64_t(tb) + uint64_t(lane);
    const uint64_t s = idx < T ? coff[idx] : P;  // coff[T] == P
    const uint64_t e = idx < T ? coff[idx + 1] : P;
    const uint64_t a = max(s, pos), b = min(e, span_end);
    const uint32_t len = b > a ? uint32_t(b - a) : 0u;  // this lane's triangle's pairs here
    // the first of them as the triangle's local pair index: 0 but for lane 0, whose
    // triangle holds pos (every later triangle starts at or after the end of lane 0's)
    const uint64_t skip0 = readlane_u64(a - s, 0);
    // ceil(len / kSegPairs) segments of nearly equal size: the first len % nseg of them one
    // pair longer than len / nseg
    const uint32_t nseg = (len + kSegPairs - 1) / kSegPairs;
    const uint32_t sq = nseg ? len / nseg : 0u;
    const uint32_t sr = len - sq * nseg;
    const uint32_t incl = wave_incl_scan(nseg);  // segment counts summed over lanes <= lane
    const uint32_t total = uint32_t(__builtin_amdgcn_readlane(int(incl), 63));
    for (uint32_t gb = 0; gb < total; gb += 64) {  // wave-uniform: 64 segments per pass
      const uint32_t g = gb + uint32_t(lane);
      // the lane holding segment g: the number of lanes whose inclusive count is <= g
      uint32_t k = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const uint32_t v = uint32_t(__shfl(int(incl), int(k) + step - 1, 64));
        if (v <= g) k += uint32_t(step);
      }
      k = k > 63 ? 63 : k;  // only lanes past the last segment (g >= total) reach 64
      const uint32_t first = uint32_t(__shfl(int(incl - nseg), int(k), 64));
      const uint32_t sqk = uint32_t(__shfl(int(sq), int(k), 64));
      const uint32_t srk = uint32_t(__shfl(int(sr), int(k), 64));
      const uint64_t skipk = k == 0 ? skip0 : 0;
      const uint32_t si = g - first;  // the segment's index in lane k's pairs
      const uint32_t soff = si * sqk + min(si, srk);
      uint32_t cnt = g < total ? sqk + (si < srk ? 1u : 0u) : 0u;
      const uint32_t t = tb + k;
      // the segment's first pair decoded in the triangle's box (first axis fastest)
      uint32_t dx = 0, dy = 0, dz = 0, yrow = 0;
      if (cnt) {
        const TriRec &r = rec[t];
        const uint32_t ex = r.mnz_ex >> 16, exy = r.exy;
        const uint64_t l = skipk + soff;
        if (__builtin_expect((r.flags & kWide) == 0, 1)) {
          const uint32_t l32 = uint32_t(l);  // < 2^22 without kWide
          dz = udiv_narrow(l32, exy, r.rcp_exy);
          const uint32_t rem = l32 - mul24(dz, exy);
          dy = udiv_narrow(rem, ex, r.rcp_ex);
          dx = rem - mul24(dy, ex);
        } else {
          const uint64_t zz = l / exy, rem = l - zz * exy;
          dz = uint32_t(zz);
          dy = uint32_t(rem / ex);
          dx = uint32_t(rem - uint64_t(dy) * ex);
        }
        yrow = dy * ex;
      }
      // walk the segment: one pair per lane per step, survivors of the plane test queued.
      // The record's fields are read again at each step (cache hits) rather than held in
      // registers across the survivor pass below, which needs them.
      uint32_t tr = t;
      while (__ballot(cnt != 0)) {  // wave-uniform
        bool surv = false;
        uint32_t x = 0, y = 0, z = 0;
        if (cnt) {
          asm volatile("" : "+v"(tr));  // the loads below stay in the loop
          // the record's 16-byte blocks 0, 2 and 3 and its flags, loaded together
          const TriRec &r = rec[tr];
          const float4 q0 = reinterpret_cast<const float4 *>(&r)[0];  // g[0..3]
          const float4 q2 = reinterpret_cast<const float4 *>(&r)[2];  // g[8], nrm[0..2]
          const uint4 q3 = reinterpret_cast<const uint4 *>(&r)[3];    // rp, mnxy, mnz_ex, exy
          const uint32_t flags = r.flags;
          const uint32_t mnxy = q3.y, mnz_ex = q3.z, exy = q3.w;
          const uint32_t ex = mnz_ex >> 16;
          bool inside = true;
          if (__builtin_expect((flags & kWindow) != 0, 0)) {
            // (w, la, lb) -> the cell w of the window of column (a, b) along c
            const WinRec wf = wrec[tr];
            const uint32_t cinfo = wf.cinfo;
            const uint32_t ax = cinfo & 3u, ec = cinfo >> 6;
            const uint32_t m0 = mnxy & 0xffffu, m1 = mnxy >> 16, m2 = mnz_ex & 0xffffu;
            const uint32_t ma = ax == 0 ? m1 : m0, mb = ax == 2 ? m1 : m2;
            const uint32_t mc = ax == 0 ? m0 : (ax == 1 ? m1 : m2);
            const uint32_t ca = ma + dy, cb = mb + dz;
            const int32_t cc =
                int32_t(floorf(wf.k0 + wf.ka * float(ca) + wf.kb * float(cb))) + int32_t(dx);
            inside = cc >= int32_t(mc) && cc < int32_t(mc + ec);
            const uint32_t ucc = uint32_t(cc);
            x = ax == 0 ? ucc : ca;
            y = ax == 1 ? ucc : (ax == 0 ? ca : cb);
            z = ax == 2 ? ucc : cb;
          } else {
            x = dx + (mnxy & 0xffffu);
            y = dy + (mnxy >> 16);
            z = dz + (mnz_ex & 0xffffu);
          }
          const float cx = float(x) + 0.5f, cy = float(y) + 0.5f, cz = float(z) + 0.5f;
          if (!inside) {
            surv = false;
          } else if (flags & kPlaneHoist) {
            // the plane test of reference_cpu.rs:76-82 with per-triangle n and r:
            // |normal.dot(ZERO) + d| = |n . v0| for a finite normal
            const float v0x = q0.x - cx, v0y = q0.y - cy, v0z = q0.z - cz;
            float dd = q2.y * v0x;
            dd = dd + q2.z * v0y;
            dd = dd + q2.w * v0z;
            surv = !(fabsf(dd) > __uint_as_float(q3.x));
          } else {
            const Rel v = rel_to(r.g, cx, cy, cz);
            surv = (flags & kLiteral) ? plane_passes<true>(v) : plane_passes<false>(v);
          }
          // next cell of the box, first axis fastest
          cnt -= 1;
          dx += 1;
          if (dx == ex) {
            dx = 0;
            dy += 1;
            yrow += ex;
            if (yrow == exy) {
              dy = 0;
              yrow = 0;
              dz += 1;
            }
          }
        }
        // queue survivors
        const unsigned long long sm = __ballot(surv);
        if (surv) {
          const uint32_t qp = qn + __popcll(sm & below);
          qp2[qp] = make_uint4(tr, x | (y << 16), z, 0u);
        }
        qn += __popcll(sm);
        if (qn >= 64) {
          __builtin_amdgcn_wave_barrier();
          qn -= 64;
          const uint32_t qe = qn + lane;
          unsigned long long key, bit;
          const uint4 e = qp2[qe];
          const bool hit = finish_pair(rec, e.x, e.y, e.z, key, bit);
          if (kOwn != kOwnNone && hit)
            owner_min<kOwn>(owner, leaf_map, own_n, e.x, e.y, e.z);
          if (kOwn != kOwnLeaf) {
            used += table_insert(tk, tv, hit, key, bit);
            if (used >= flush_at) {
              table_flush(tk, tv, lane, leafgrid, brickbits);
              used = 0;
            }
          }
        }
      }
    }
    // the next batch starts at the first triangle past this one's pairs
    uint64_t ib = idx;
    asm volatile("" : "+v"(ib));  // a second load, not the first one's value kept live
    const uint64_t eb = ib < T ? coff[ib + 1] : P;
    const uint64_t e63 = readlane_u64(eb, 63);
    if (e63 < span_end) {
      pos = e63;
      tb += 64;
    } else {
      tb += uint32_t(__popcll(__ballot(eb <= span_end)));
      pos = span_end;
    }
  }
  }
  __builtin_amdgcn_wave_barrier();
  {
    unsigned long long key = 0, bit = 0;
    bool hit = false;
    const uint4 e = qp2[lane];
    if (uint32_t(lane) < qn) hit = finish_pair(rec, e.x, e.y, e.z, key, bit);
    if (kOwn != kOwnNone && hit)
      owner_min<kOwn>(owner, leaf_map, own_n, e.x, e.y, e.z);
    if (kOwn != kOwnLeaf) table_insert(tk, tv, hit, key, bit);
  }
  if (kOwn != kOwnLeaf) table_flush(tk, tv, lane, leafgrid, brickbits);
}

}  // namespace

namespace {

// Scan of the Tp per-slot candidate counts in ctx->tri_count, the device-side plan, and
// the SAT launch into `leafgrid` / `brickbits`, all enqueued without a host round trip. The
// persistent grid is sized for the device (waves that find no unit exit at once).
// The Tp window records follow the Tp triangle records in ctx->rec.
WinRec *win_records(vxl_ctx *ctx, uint64_t Tp) {
  return reinterpret_cast<WinRec *>(ctx->rec.as<char>() + Tp * sizeof(TriRec));
}

uint32_t merge_words(const vxl_ctx *ctx) {
  const int w = ctx->sat_merge_words;
  return uint32_t(w < 1 ? 1 : (w > kFlushMax ? kFlushMax : w));
}

uint64_t plan_want(vxl_ctx *ctx, uint64_t *max_waves) {
  *max_waves = uint64_t(ctx->cu_count > 0 ? ctx->cu_count : 256) * ctx->sat_waves_per_cu;
  return *max_waves * uint64_t(ctx->sat_units_per_wave);
}

static_assert(sizeof(SatPlan) <= kPlanBytes, "the unit start table follows the plan");

// The plan area (SatPlan, then the unit starts), sized before setup writes into it; a new
// area starts with both skipped-cell slots clear.
SatPlan *plan_area(vxl_ctx *ctx) {
  uint64_t mw = 0;
  const uint64_t want = plan_want(ctx, &mw);
  const size_t had = ctx->bstart.bytes;
  ctx->bstart.reserve(kPlanBytes + (want + 64) * 4);
  SatPlan *plan = ctx->bstart.as<SatPlan>();
  if (ctx->bstart.bytes != had)
    VXL_HIP(hipMemsetAsync(plan->skipped, 0, sizeof(plan->skipped), ctx->stream));
  return plan;
}

void sat_after_setup(vxl_ctx *ctx, uint64_t Tp, unsigned long long *leafgrid,
                     unsigned long long *brickbits, uint32_t epoch, uint32_t *owner = nullptr,
                     const uint32_t *leaf_map = nullptr, uint32_t own_n = 0) {
  hipStream_t s = ctx->stream;
  const int tb = 256;
  uint64_t *count = ctx->tri_count.as<uint64_t>();
  TriRec *rec = ctx->rec.as<TriRec>();
  const WinRec *wrec = win_records(ctx, Tp);
  uint64_t *coff = ctx->coff.as<uint64_t>();
  uint64_t max_waves = 0;
  const uint64_t want = plan_want(ctx, &max_waves);
  SatPlan *plan = plan_area(ctx);
  uint32_t *ustart = reinterpret_cast<uint32_t *>(ctx->bstart.as<char>() + kPlanBytes);
  const uint64_t nb = (Tp + kSetupTB - 1) / kSetupTB;
  if (nb <= kFusedBlocks) {
    uint64_t *sums = ctx->setup_sums.as<uint64_t>();
    hipLaunchKernelGGL(k_setup_plan, dim3(1), dim3(kPlanThreads), 0, s, sums, nb, want, plan,
                       epoch, coff + Tp);
    if (nb)
      hipLaunchKernelGGL(k_offsets_units, dim3(uint32_t((nb + kOffsetBlocks - 1) / kOffsetBlocks)),
                         dim3(64 * kOffsetBlocks), 0, s, count, coff, sums, nb, Tp, plan, ustart);
  } else {
    scan_exclusive_u64(count, coff, Tp, coff + Tp, ctx->scan_tmp, s);
    hipLaunchKernelGGL(k_sat_plan, dim3(1), dim3(1), 0, s, coff + Tp, want, plan, epoch);
    hipLaunchKernelGGL(k_bstart, dim3((Tp + tb - 1) / tb), dim3(tb), 0, s, coff, count, Tp, plan,
                       ustart);
  }
  mark(ctx, 1);
  const uint64_t blocks = (max_waves + kWaves - 1) / kWaves;
  const dim3 g(static_cast<uint32_t>(blocks), 1, 1), b(kWaves * 64, 1, 1);
  if (owner && leaf_map)
    hipLaunchKernelGGL((k_sat<kOwnLeaf>), g, b, 0, s, rec, wrec, coff, ustart, Tp, plan, leafgrid,
                       brickbits, owner, leaf_map, own_n, merge_words(ctx));
  else if (owner)
    hipLaunchKernelGGL((k_sat<kOwnDense>), g, b, 0, s, rec, wrec, coff, ustart, Tp, plan, leafgrid,
                       brickbits, owner, nullptr, own_n, merge_words(ctx));
  else
    hipLaunchKernelGGL((k_sat<kOwnNone>), g, b, 0, s, rec, wrec, coff, ustart, Tp, plan, leafgrid,
                       brickbits, nullptr, nullptr, 0u, merge_words(ctx));
  mark(ctx, 2);
  VXL_HIP(hipGetLastError());
}

void reserve_records(vxl_ctx *ctx, uint64_t Tp) {
  ctx->tri_count.reserve(Tp * 8);
  ctx->setup_sums.reserve(((Tp + kSetupTB - 1) / kSetupTB + 1) * 8);
  ctx->rec.reserve(Tp * (sizeof(TriRec) + sizeof(WinRec)));
  ctx->coff.reserve(Tp * 8 + 8);
}

}  // namespace

void run_sat(vxl_ctx *ctx, const float *d_tris, uint64_t T, const VoxelizeParams &p,
             uint32_t *owner, const uint32_t *leaf_map) {
  reserve_records(ctx, T);
  SatPlan *plan = plan_area(ctx);
  const uint32_t e = ++ctx->sat_epoch;
  hipLaunchKernelGGL(k_tri_setup, dim3((T + kSetupTB - 1) / kSetupTB), dim3(kSetupTB), 0,
                     ctx->stream, d_tris, T, p, ctx->tri_count.as<uint64_t>(),
                     ctx->rec.as<TriRec>(), win_records(ctx, T), ctx->setup_sums.as<uint64_t>(),
                     plan->skipped[e & 1]);
  sat_after_setup(ctx, T, ctx->leafgrid.as<unsigned long long>(),
                  ctx->bitmap.as<unsigned long long>(), e, owner, leaf_map, p.n);
}

const uint64_t *sat_pairs_dev(vxl_ctx *ctx) { return &ctx->bstart.as<SatPlan>()->P_ref; }

namespace {
__global__ void k_plan_rearm(SatPlan *__restrict__ plan) {
  for (int c = 0; c < kUnitCounters; c++) plan->counter[c][0] = 0;
}
}  // namespace

void sat_owner_pass(vxl_ctx *ctx, uint64_t T, uint32_t *owner, const uint32_t *leaf_map,
                    uint32_t own_n) {
  hipStream_t s = ctx->stream;
  SatPlan *plan = ctx->bstart.as<SatPlan>();
  uint32_t *ustart = reinterpret_cast<uint32_t *>(ctx->bstart.as<char>() + kPlanBytes);
  hipLaunchKernelGGL(k_plan_rearm, dim3(1), dim3(1), 0, s, plan);
  uint64_t max_waves = 0;
  (void)plan_want(ctx, &max_waves);
  const dim3 g(static_cast<uint32_t>((max_waves + kWaves - 1) / kWaves), 1, 1), b(kWaves * 64, 1, 1);
  hipLaunchKernelGGL((k_sat<kOwnLeaf>), g, b, 0, s, ctx->rec.as<TriRec>(), win_records(ctx, T),
                     ctx->coff.as<uint64_t>(), ustart, T, plan,
                     ctx->leafgrid.as<unsigned long long>(), ctx->bitmap.as<unsigned long long>(),
                     owner, leaf_map, own_n, merge_words(ctx));
  VXL_HIP(hipGetLastError());
}

void run_sat_batch(vxl_ctx *ctx, const BatchMesh *d_mesh, const uint32_t *d_sb_begin,
                   uint32_t G, uint32_t setup_blocks, unsigned long long *leafgrid,
                   unsigned long long *brickbits, unsigned long long *d_ref_pairs) {
  const uint64_t Tp = uint64_t(setup_blocks) * kSetupTB;
  reserve_records(ctx, Tp);
  SatPlan *plan = plan_area(ctx);
  const uint32_t e = ++ctx->sat_epoch;
  if (setup_blocks)
    hipLaunchKernelGGL(k_tri_setup_batch, dim3(setup_blocks), dim3(kSetupTB), 0, ctx->stream,
                       d_mesh, d_sb_begin, G, ctx->tri_count.as<uint64_t>(),
                       ctx->rec.as<TriRec>(), win_records(ctx, Tp),
                       ctx->setup_sums.as<uint64_t>(), d_ref_pairs);
  sat_after_setup(ctx, Tp, leafgrid, brickbits, e);
}

uint64_t setup_slots_per_block() { return kSetupTB; }

namespace {
// hash_color (reference_cpu.rs:91-99): three steps of the 32-bit LCG x -> 1664525 x +
// 1013904223 from the owner index, one byte each for r, g, b, and alpha 255.
__device__ __forceinline__ uint32_t hash_color(uint32_t id) {
  uint32_t x = id * 1664525u + 1013904223u;
  const uint32_t r = x & 0xffu;
  x = x * 1664525u + 1013904223u;
  const uint32_t g = x & 0xffu;
  x = x * 1664525u + 1013904223u;
  const uint32_t b = x & 0xffu;
  return r | (g << 8) | (b << 16) | (255u << 24);
}

// color[i] = hash_color(owner[i]) where a triangle owns voxel i, else 0
// (reference_cpu.rs:173-186; voxelize.wgsl:264-281).
__global__ void k_owner_color(const uint32_t *__restrict__ owner, uint32_t *__restrict__ color,
                              uint64_t N) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t o = owner[i];
    color[i] = o != 0xFFFFFFFFu ? hash_color(o) : 0u;
  }
}
}  // namespace

namespace {
__global__ void k_fill_u64(unsigned long long *__restrict__ p, uint64_t count,
                           unsigned long long v) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
}  // namespace

void fill_u64(unsigned long long *p, uint64_t count, unsigned long long v, hipStream_t s) {
  if (count == 0) return;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_fill_u64, dim3(uint32_t(blocks)), dim3(256), 0, s, p, count, v);
  VXL_HIP(hipGetLastError());
}

void owner_colors(vxl_ctx *ctx, const uint32_t *owner, uint32_t *color, uint64_t N) {
  if (N == 0) return;
  uint64_t blocks = (N + 255) / 256;  // grid-stride kernel: at most 2^22 blocks
  if (blocks > (1ull << 22)) blocks = 1ull << 22;
  hipLaunchKernelGGL(k_owner_color, dim3(uint32_t(blocks)), dim3(256), 0, ctx->stream, owner,
                     color, N);
  VXL_HIP(hipGetLastError());
}

}  // namespace vxl
