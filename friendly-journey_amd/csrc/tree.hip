// Sparse-tree construction on the device from the dense leaf grid + brick bitmap.
//
// The reference builds the tree on the host: an n^3 scan for occupied 8^3 bricks
// (voxel-core/src/sparse.rs:197-271), a Morton sort (sparse.rs:281-305), a bottom-up
// 4^3 OR-reduce (build_parents/build_levels, sparse.rs:128-163) and a recursive
// parent-first DFS re-serialisation (emit_subtree/from_sparse, school_b.rs:68-151).
//
// Here the same arrays come out of a bitmap pyramid, with no sort:
//  * B1 has one bit per brick Morton code. Word j of B1 holds the 64 children of the
//    level-2 node whose code is j, so its bits ARE that node's child mask; its rank among
//    non-zero words is the node's School-A index; the exclusive prefix popcount is its
//    child_base (the Morton-sorted index of its first leaf). This is build_parents on the
//    sorted code list, computed by scans.
//  * B_L (L >= 2) has one bit per level-L node code, set when word j of B_{L-1} is
//    non-zero; the wave ballot over 64 consecutive words produces one B_L word.
//  * School-B positions follow from subtree sizes: desc(v) = sum over children c of
//    (1 + desc(c)), 0 on level 2; pos(c_j) = blockstart(P) + j and
//    blockstart(c_j) = blockstart(P) + m(P) + sum_{i<j} desc(c_i), with the root at 0 and
//    its block at 1 -- exactly the order emit_subtree appends nodes in. Level-2 nodes
//    keep the School-A leaf base as child_base (school_b.rs:75-79).
//  * The leaf gather copies each occupied brick's 8 words out of the leaf grid into the
//    Morton-sorted leaf array, computes the packed LeafBounds (leaf.rs:185-213), and
//    writes zeros back, so the grid and B1 are all-zero again when the build ends.
#include <algorithm>

#include "vxl_internal.hpp"

namespace vxl {
namespace {

constexpr int kTB = 256;
constexpr uint32_t kFull = (7u << 9) | (7u << 12) | (7u << 15);  // LeafBounds::FULL.pack()

inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Per word j of B_{L-1}: popcount and non-zero flag; ballot-pack the flags into B_L.
__global__ __launch_bounds__(kTB) void k_level(const unsigned long long *__restrict__ bin,
                                                 uint64_t win, uint32_t *__restrict__ pop,
                                                 uint32_t *__restrict__ nz,
                                                 unsigned long long *__restrict__ bout) {
  uint64_t j = uint64_t(blockIdx.x) * kTB + threadIdx.x;
  unsigned long long w = j < win ? bin[j] : 0ull;
  if (j < win) {
    pop[j] = __popcll(w);
    nz[j] = w != 0ull;
  }
  unsigned long long b = __ballot(w != 0ull);
  if (bout && (threadIdx.x & 63) == 0 && j < win) bout[j >> 6] = b;
}

// School-A node arrays of one level from its bitmap words.
__global__ __launch_bounds__(kTB) void k_emit_a(const unsigned long long *__restrict__ bin,
                                                  uint64_t win, const uint32_t *__restrict__ cb,
                                                  const uint32_t *__restrict__ ni,
                                                  unsigned long long *__restrict__ a_mask,
                                                  uint32_t *__restrict__ a_base,
                                                  uint32_t *__restrict__ a_code) {
  uint64_t j = uint64_t(blockIdx.x) * kTB + threadIdx.x;
  if (j >= win) return;
  unsigned long long w = bin[j];
  if (!w) return;
  uint32_t i = ni[j];
  a_mask[i] = w;
  a_base[i] = cb[j];
  a_code[i] = uint32_t(j);
}

// Coordinates 0..3 along `axis` occupied in a 4^3 sub-brick word, as a 4-bit mask
// (bit = x0|y0<<1|z0<<2|x1<<3|y1<<4|z1<<5).
__device__ __forceinline__ uint32_t word_coords(unsigned long long w, int axis) {
  const unsigned long long lo0[3] = {0x5555555555555555ull, 0x3333333333333333ull,
                                     0x0F0F0F0F0F0F0F0Full};
  const unsigned long long lo1[3] = {0x00FF00FF00FF00FFull, 0x0000FFFF0000FFFFull,
                                     0x00000000FFFFFFFFull};
  const unsigned long long m0 = lo0[axis] & lo1[axis], m1 = ~lo0[axis] & lo1[axis];
  const unsigned long long m2 = lo0[axis] & ~lo1[axis], m3 = ~lo0[axis] & ~lo1[axis];
  return ((w & m0) ? 1u : 0u) | ((w & m1) ? 2u : 0u) | ((w & m2) ? 4u : 0u) |
         ((w & m3) ? 8u : 0u);
}

// Coordinate masks of a 4^3 sub-brick word (bit = x0|y0<<1|z0<<2|x1<<3|y1<<4|z1<<5).
__device__ __forceinline__ void word_extent(unsigned long long w, int axis, uint32_t &mn,
                                            uint32_t &mx) {
  const unsigned long long lo0[3] = {0x5555555555555555ull, 0x3333333333333333ull,
                                     0x0F0F0F0F0F0F0F0Full};
  const unsigned long long lo1[3] = {0x00FF00FF00FF00FFull, 0x0000FFFF0000FFFFull,
                                     0x00000000FFFFFFFFull};
  unsigned long long m[4];
  m[0] = lo0[axis] & lo1[axis];
  m[1] = ~lo0[axis] & lo1[axis];
  m[2] = lo0[axis] & ~lo1[axis];
  m[3] = ~lo0[axis] & ~lo1[axis];
  mn = 7;
  mx = 0;
#pragma unroll
  for (int v = 3; v >= 0; v--)
    if (w & m[v]) mn = v;
#pragma unroll
  for (int v = 0; v < 4; v++)
    if (w & m[v]) mx = v;
}

// One wave per level-2 node: 8 children at a time, 8 lanes (one per u64 word) per child.
// In a batch group (n2_beg non-null) the stored brick code is relative to the mesh's region.
// With store false the node's grid words are read and cleared but nothing is written.
__device__ __forceinline__ void gather_node(
    uint64_t node, uint64_t n2, const unsigned long long *__restrict__ a_mask,
    const uint32_t *__restrict__ a_base, const uint32_t *__restrict__ a_code,
    unsigned long long *__restrict__ leafgrid, unsigned long long *__restrict__ brickbits,
    unsigned long long *__restrict__ leaves, uint32_t *__restrict__ bounds,
    unsigned long long *__restrict__ codes, const uint64_t *__restrict__ n2_beg, uint32_t G,
    const BatchMesh *__restrict__ mesh, bool store = true) {
  if (node >= n2) return;
  const int lane = threadIdx.x & 63;
  const unsigned long long mask = a_mask[node];
  const uint32_t base = a_base[node];
  const uint32_t j = a_code[node];
  const uint64_t code_bias = n2_beg ? mesh[dev::upper_index(n2_beg, G, node)].b1_begin << 6 : 0;
  const uint32_t nch = __popcll(mask);
  // the position of the mask's s-th set bit, held by lane s: lane j whose bit is set sends j
  // to the lane of its rank (the set bits below it); the other lanes send to lane 63, which
  // is no rank's lane unless all 64 bits are set (and then every lane sends its own rank)
  const bool own = (mask >> lane) & 1ull;
  const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                                  __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
  const int bitpos = __builtin_amdgcn_ds_permute(own ? int(rank) * 4 : 63 * 4, lane);
  const int sub = lane >> 3, w = lane & 7;
  for (uint32_t g = 0; g < nch; g += 8) {
    const uint32_t s = g + sub;
    const bool act = s < nch;
    const uint32_t bit = uint32_t(__shfl(bitpos, act ? int(s) : 0, 64));
    const uint64_t code = (uint64_t(j) << 6) | bit;
    unsigned long long v = 0;
    if (act) {
      v = leafgrid[code * 8 + w];
      leafgrid[code * 8 + w] = 0ull;
      if (store) leaves[(uint64_t(base) + s) * 8 + w] = v;
    }
    // LeafBounds over the 8 words of this leaf (8-lane groups): each lane marks the
    // coordinates its 4^3 sub-brick word occupies along x, y, z (one byte each, shifted to
    // the word's octant), then three xor-swizzle rounds OR the marks and add the counts.
    uint32_t cnt = __popcll(v);
    uint32_t occ = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) occ |= (word_coords(v, a) << ((w >> a) & 1) * 4) << (8 * a);
    cnt += (uint32_t)__builtin_amdgcn_ds_swizzle((int)cnt, 0x041f);  // lane ^ 1
    occ |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)occ, 0x041f);
    cnt += (uint32_t)__builtin_amdgcn_ds_swizzle((int)cnt, 0x081f);  // lane ^ 2
    occ |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)occ, 0x081f);
    cnt += (uint32_t)__builtin_amdgcn_ds_swizzle((int)cnt, 0x101f);  // lane ^ 4
    occ |= (uint32_t)__builtin_amdgcn_ds_swizzle((int)occ, 0x101f);
    uint32_t mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const uint32_t m = (occ >> (8 * a)) & 0xffu;
      mn[a] = m ? uint32_t(__ffs(m) - 1) : 7u;
      mx[a] = m ? 31u - uint32_t(__clz(m)) : 0u;
    }
    if (store && act && w == 0) {
      uint32_t b = (cnt > 64 || cnt == 0)
                       ? kFull
                       : (mn[0] | (mn[1] << 3) | (mn[2] << 6) | (mx[0] << 9) | (mx[1] << 12) |
                          (mx[2] << 15));
      bounds[base + s] = b;
      codes[base + s] = code - code_bias;
    }
  }
  if (lane == 0) brickbits[j] = 0ull;  // B1 back to all-zero
}

__global__ __launch_bounds__(kTB) void k_leaf_gather(
    uint64_t n2, const unsigned long long *__restrict__ a_mask, const uint32_t *__restrict__ a_base,
    const uint32_t *__restrict__ a_code, unsigned long long *__restrict__ leafgrid,
    unsigned long long *__restrict__ brickbits, unsigned long long *__restrict__ leaves,
    uint32_t *__restrict__ bounds, unsigned long long *__restrict__ codes,
    const uint64_t *__restrict__ n2_beg, uint32_t G, const BatchMesh *__restrict__ mesh) {
  gather_node((uint64_t(blockIdx.x) * kTB + threadIdx.x) >> 6, n2, a_mask, a_base, a_code,
              leafgrid, brickbits, leaves, bounds, codes, n2_beg, G, mesh);
}

// desc_L[i] = m_i + sum of desc over its children (level L-1), from the exclusive scan E.
__global__ void k_desc(uint64_t nl, const unsigned long long *__restrict__ a_mask,
                       const uint32_t *__restrict__ a_base, const uint32_t *__restrict__ e_child,
                       uint32_t *__restrict__ desc) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  uint32_t m = __popcll(a_mask[i]);
  uint32_t b = a_base[i];
  desc[i] = m + (e_child ? (e_child[b + m] - e_child[b]) : 0u);
}

// Root of the School-B buffer.
__global__ void k_root(const unsigned long long *__restrict__ a_mask,
                       const uint32_t *__restrict__ a_base, int root_is_level2,
                       uint32_t *__restrict__ nodes_b, uint32_t *__restrict__ pos,
                       uint32_t *__restrict__ bs) {
  unsigned long long m = a_mask[0];
  nodes_b[0] = uint32_t(m);
  nodes_b[1] = uint32_t(m >> 32);
  nodes_b[2] = root_is_level2 ? a_base[0] : 1u;
  pos[0] = 0;
  bs[0] = 1;
}

// Top-down placement of level-L nodes under their level-(L+1) parents.
__global__ void k_place(uint64_t nl, int is_level2, const unsigned long long *__restrict__ a_mask,
                        const uint32_t *__restrict__ a_base, const uint32_t *__restrict__ a_code,
                        const uint32_t *__restrict__ e_self, const uint32_t *__restrict__ ni_parent,
                        const unsigned long long *__restrict__ p_mask,
                        const uint32_t *__restrict__ p_base, const uint32_t *__restrict__ p_bs,
                        uint32_t *__restrict__ pos, uint32_t *__restrict__ bs,
                        uint32_t *__restrict__ nodes_b) {
  uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const uint32_t q = a_code[c];
  const uint32_t p = ni_parent[q >> 6];
  const uint32_t pb = p_base[p];
  const uint32_t jj = uint32_t(c) - pb;
  const uint32_t mp = __popcll(p_mask[p]);
  const uint32_t pbs = p_bs[p];
  const uint32_t mypos = pbs + jj;
  const uint32_t mybs = pbs + mp + (e_self ? (e_self[c] - e_self[pb]) : 0u);
  pos[c] = mypos;
  bs[c] = mybs;
  const unsigned long long m = a_mask[c];
  nodes_b[3ull * mypos + 0] = uint32_t(m);
  nodes_b[3ull * mypos + 1] = uint32_t(m >> 32);
  nodes_b[3ull * mypos + 2] = is_level2 ? a_base[c] : mybs;
}

// School-B entry c of one tree once its node array exists: c < n2 places level-2 node c
// under its level-3 parent (k_place's rule, child_base = the School-A leaf base); the rest
// are the nodes of levels 3..k+1 (n3, n4, n5 of them, in that order), whose positions the
// upper-levels kernel computed. Entries past the node array's capacity are never written.
__device__ __forceinline__ void place_entry(uint64_t c, uint64_t n2, const LevelPtrs &lp,
                                            const uint32_t *__restrict__ upper, uint64_t cap,
                                            uint32_t *__restrict__ nodes_b) {
  const uint32_t n3 = upper[0], n4 = upper[1], n5 = upper[2];
  if (c >= n2 + n3 + n4 + n5) return;
  uint32_t mypos, child_base;
  unsigned long long m;
  if (c < n2) {
    const uint32_t q = lp.a_code[2][c];
    const uint32_t p = lp.ni[3][q >> 6];
    mypos = lp.bs[3][p] + (uint32_t(c) - lp.a_base[3][p]);
    m = lp.a_mask[2][c];
    child_base = lp.a_base[2][c];
  } else {
    uint32_t i = uint32_t(c - n2);
    int L = 3;
    if (i >= n3) {
      i -= n3;
      L = 4;
      if (i >= n4) {
        i -= n4;
        L = 5;
      }
    }
    mypos = lp.pos[L][i];
    m = lp.a_mask[L][i];
    child_base = lp.bs[L][i];
  }
  if (mypos >= cap) return;
  nodes_b[3ull * mypos + 0] = uint32_t(m);
  nodes_b[3ull * mypos + 1] = uint32_t(m >> 32);
  nodes_b[3ull * mypos + 2] = child_base;
}

// The leaf gather (blocks [0, gather_blocks), one wave per level-2 node) and the School-B
// node entries (the blocks after them, one thread per entry) of one tree with k >= 2, in
// one launch: both read only what the kernels before them wrote.
__global__ __launch_bounds__(kTB) void k_gather_place(
    uint64_t n2, uint32_t gather_blocks, LevelPtrs lp, const uint32_t *__restrict__ upper,
    uint64_t cap, unsigned long long *__restrict__ leafgrid,
    unsigned long long *__restrict__ brickbits, unsigned long long *__restrict__ leaves,
    uint32_t *__restrict__ bounds, unsigned long long *__restrict__ codes,
    uint32_t *__restrict__ nodes_b) {
  if (blockIdx.x < gather_blocks) {
    gather_node((uint64_t(blockIdx.x) * kTB + threadIdx.x) >> 6, n2, lp.a_mask[2], lp.a_base[2],
                lp.a_code[2], leafgrid, brickbits, leaves, bounds, codes, nullptr, 0u, nullptr);
    return;
  }
  place_entry(uint64_t(blockIdx.x - gather_blocks) * kTB + threadIdx.x, n2, lp, upper, cap,
              nodes_b);
}

// x-major dense words from the leaf grid: word i covers x0..x0+31 of row (y, z).
__global__ void k_grid_to_words(uint32_t n, uint64_t nwords,
                                const unsigned long long *__restrict__ leafgrid,
                                uint32_t *__restrict__ words) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  uint64_t lin = i * 32;
  uint32_t out = 0;
  for (int b = 0; b < 32; b++) {
    uint64_t v = lin + b;
    uint32_t x = uint32_t(v % n);
    uint64_t yz = v / n;
    uint32_t y = uint32_t(yz % n), z = uint32_t(yz / n);
    if (z >= n) break;
    uint32_t code = dev::brick_code(x >> 3, y >> 3, z >> 3);
    uint32_t m = dev::morton9(x, y, z);
    if ((leafgrid[uint64_t(code) * 8 + (m >> 6)] >> (m & 63)) & 1ull) out |= 1u << b;
  }
  words[i] = out;
}

// Leaf grid + brick bitmap from x-major words: one thread per sub-brick word.
__global__ void k_words_to_grid(uint32_t n, uint64_t ngrid, const uint32_t *__restrict__ words,
                                unsigned long long *__restrict__ leafgrid,
                                unsigned long long *__restrict__ brickbits) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= ngrid) return;
  uint32_t code = uint32_t(i >> 3), w = uint32_t(i & 7);
  uint32_t bx = dev::compact3_10(code), by = dev::compact3_10(code >> 1),
           bz = dev::compact3_10(code >> 2);
  uint32_t x0 = bx * 8 + (w & 1) * 4, y0 = by * 8 + ((w >> 1) & 1) * 4,
           z0 = bz * 8 + ((w >> 2) & 1) * 4;
  unsigned long long v = 0;
  for (uint32_t dz = 0; dz < 4; dz++)
    for (uint32_t dy = 0; dy < 4; dy++) {
      uint64_t lin = uint64_t(x0) + uint64_t(n) * (uint64_t(y0 + dy) + uint64_t(n) * (z0 + dz));
      uint32_t row = (words[lin >> 5] >> (lin & 31)) & 0xFu;
      for (uint32_t dx = 0; dx < 4; dx++)
        if ((row >> dx) & 1u) v |= 1ull << dev::child_bit(dx, dy, dz);
    }
  leafgrid[i] = v;
  if (v) atomicOr(&brickbits[code >> 6], 1ull << (code & 63));
}

// ---- fast path (n <= 2048): level 2 multi-block, upper levels in one workgroup ----------
constexpr int kSmallThreads = 1024;
constexpr uint64_t kSmallWords = 4096;  // largest upper-level input handled by one workgroup

// Level-2 pass: packed (popcount | nonzero << 32) per B1 word, ballot-packed B2.
__global__ __launch_bounds__(kTB) void k_level2(const unsigned long long *__restrict__ bin,
                                                  uint64_t win, uint64_t *__restrict__ pn,
                                                  unsigned long long *__restrict__ bout) {
  const uint64_t j = uint64_t(blockIdx.x) * kTB + threadIdx.x;
  const unsigned long long w = j < win ? bin[j] : 0ull;
  if (j < win) pn[j] = uint64_t(__popcll(w)) | (uint64_t(w != 0ull) << 32);
  const unsigned long long b = __ballot(w != 0ull);
  if (bout && (threadIdx.x & 63) == 0 && j < win) bout[j >> 6] = b;
}

__global__ __launch_bounds__(kTB) void k_emit_a2(const unsigned long long *__restrict__ bin,
                                                   uint64_t win, const uint64_t *__restrict__ pn,
                                                   unsigned long long *__restrict__ a_mask,
                                                   uint32_t *__restrict__ a_base,
                                                   uint32_t *__restrict__ a_code) {
  const uint64_t j = uint64_t(blockIdx.x) * kTB + threadIdx.x;
  if (j >= win) return;
  const unsigned long long w = bin[j];
  if (!w) return;
  const uint64_t v = pn[j];
  const uint32_t i = uint32_t(v >> 32);
  a_mask[i] = w;
  a_base[i] = uint32_t(v);
  a_code[i] = uint32_t(j);
}

// Exclusive block scan of one u32 per thread (kSmallThreads threads); returns the total.
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t *excl, uint32_t *sh) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = (uint32_t)__shfl_up((int)inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kSmallThreads / 64; w++) {
    const uint32_t t = sh[w];
    if (w < wave) before += t;
    total += t;
  }
  *excl = before + inc - v;
  __syncthreads();
  return total;
}

// Levels first..top (each with at most kSmallWords input words) in one workgroup:
// popcounts, flags, scans, School-A arrays and the next bitmap, level after level.
__device__ __forceinline__ void small_levels(const LevelPtrs &lp, int first, int top,
                                             const unsigned long long *__restrict__ bin_first,
                                             uint32_t *__restrict__ tot) {
  __shared__ unsigned long long bcur[kSmallWords];
  __shared__ unsigned long long bnext[kSmallWords / 64];
  __shared__ uint32_t sh[kSmallThreads / 64];
  const int tid = threadIdx.x;
  for (uint64_t j = tid; j < lp.win[first]; j += kSmallThreads) bcur[j] = bin_first[j];
  __syncthreads();
  for (int L = first; L <= top; L++) {
    const uint64_t win = lp.win[L];
    const int ipt = int((win + kSmallThreads - 1) / kSmallThreads);
    for (uint64_t j = tid; j < kSmallWords / 64; j += kSmallThreads) bnext[j] = 0ull;
    uint32_t spop = 0, snz = 0;
    for (int i = 0; i < ipt; i++) {
      const uint64_t j = uint64_t(tid) * ipt + i;
      const unsigned long long w = j < win ? bcur[j] : 0ull;
      spop += __popcll(w);
      snz += w != 0ull;
    }
    uint32_t epop, enz;
    const uint32_t tpop = block_scan_u32(spop, &epop, sh);
    const uint32_t tnz = block_scan_u32(snz, &enz, sh);
    for (int i = 0; i < ipt; i++) {
      const uint64_t j = uint64_t(tid) * ipt + i;
      if (j >= win) break;
      const unsigned long long w = bcur[j];
      lp.cb[L][j] = epop;
      lp.ni[L][j] = enz;
      if (w) {
        lp.a_mask[L][enz] = w;
        lp.a_base[L][enz] = epop;
        lp.a_code[L][enz] = uint32_t(j);
        atomicOr(&bnext[j >> 6], 1ull << (j & 63));
      }
      epop += __popcll(w);
      enz += w != 0ull;
    }
    if (tid == 0) {
      lp.cb[L][win] = tpop;
      lp.ni[L][win] = tnz;
      tot[2 * L] = tpop;
      tot[2 * L + 1] = tnz;
    }
    __syncthreads();
    const uint64_t wnext = (win + 63) / 64;
    for (uint64_t j = tid; j < wnext; j += kSmallThreads) bcur[j] = bnext[j];
    __syncthreads();
  }
}

// One workgroup per mesh of a batch group; its upper bitmap words start at b1_begin / 64.
__global__ __launch_bounds__(kSmallThreads) void k_small_levels_batch(
    const BatchMesh *__restrict__ mesh, const LevelPtrs *__restrict__ lps,
    const unsigned long long *__restrict__ b2) {
  const BatchMesh &m = mesh[blockIdx.x];
  const LevelPtrs &lp = lps[blockIdx.x];
  small_levels(lp, 3, int(m.k) + 1, b2 + m.b1_begin / 64, lp.tot);
}

// Subtree sizes (levels 3..k+1), the root, and School-B placement of levels k..3, in one
// workgroup; the level-2 placement follows as a multi-block pass. With nodes_b null only the
// positions (pos, bs) are computed.
__device__ __forceinline__ void upper_place(const LevelPtrs &lp, int k,
                                            const uint32_t *__restrict__ tot,
                                            uint32_t *__restrict__ nodes_b) {
  __shared__ uint32_t sh[kSmallThreads / 64];
  const int tid = threadIdx.x;
  for (int L = 3; L <= k + 1; L++) {
    const uint32_t nl = tot[2 * L + 1];
    const int ipt = int((nl + kSmallThreads - 1) / kSmallThreads);
    uint32_t local = 0;
    for (int i = 0; i < ipt; i++) {
      const uint32_t c = uint32_t(tid) * ipt + i;
      if (c >= nl) break;
      const uint32_t m = __popcll(lp.a_mask[L][c]);
      const uint32_t b = lp.a_base[L][c];
      const uint32_t d = m + (L > 3 ? lp.desc[L - 1][b + m] - lp.desc[L - 1][b] : 0u);
      lp.pos[L][c] = d;  // staged desc
      local += d;
    }
    uint32_t excl;
    const uint32_t total = block_scan_u32(local, &excl, sh);
    for (int i = 0; i < ipt; i++) {
      const uint32_t c = uint32_t(tid) * ipt + i;
      if (c >= nl) break;
      const uint32_t d = lp.pos[L][c];
      lp.desc[L][c] = excl;  // E_L: exclusive scan of desc
      excl += d;
    }
    if (tid == 0) lp.desc[L][nl] = total;
    __syncthreads();
  }
  if (tid == 0) {
    if (nodes_b) {
      const unsigned long long m = lp.a_mask[k + 1][0];
      nodes_b[0] = uint32_t(m);
      nodes_b[1] = uint32_t(m >> 32);
      nodes_b[2] = 1u;
    }
    lp.pos[k + 1][0] = 0;
    lp.bs[k + 1][0] = 1;
  }
  __syncthreads();
  for (int L = k; L >= 3; L--) {
    const uint32_t nl = tot[2 * L + 1];
    for (uint32_t c = tid; c < nl; c += kSmallThreads) {
      const uint32_t q = lp.a_code[L][c];
      const uint32_t p = lp.ni[L + 1][q >> 6];
      const uint32_t pb = lp.a_base[L + 1][p];
      const uint32_t mp = __popcll(lp.a_mask[L + 1][p]);
      const uint32_t pbs = lp.bs[L + 1][p];
      const uint32_t mypos = pbs + (c - pb);
      const uint32_t mybs = pbs + mp + (lp.desc[L][c] - lp.desc[L][pb]);
      lp.pos[L][c] = mypos;
      lp.bs[L][c] = mybs;
      if (nodes_b) {
        const unsigned long long m = lp.a_mask[L][c];
        nodes_b[3ull * mypos + 0] = uint32_t(m);
        nodes_b[3ull * mypos + 1] = uint32_t(m >> 32);
        nodes_b[3ull * mypos + 2] = mybs;
      }
    }
    __syncthreads();
  }
}

// ---- batch groups -------------------------------------------------------------------
// Group-wide arrays: pn2 is the packed scan over every mesh's B1 region in order, so a
// level-2 node's global index and leaf base are group-wide; per mesh they start at
// n2_beg[g] and leaf_beg[g]. Upper levels are per mesh (one workgroup each).
__global__ void k_batch_totals(const BatchMesh *__restrict__ mesh, const LevelPtrs *__restrict__ lps,
                               uint32_t G, const uint64_t *__restrict__ pn2,
                               const uint64_t *__restrict__ coff,
                               const unsigned long long *__restrict__ skipped,
                               BatchTotals *__restrict__ out, uint64_t *__restrict__ n2_beg) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const BatchMesh &m = mesh[g];
  const uint64_t lo = pn2[m.b1_begin], hi = pn2[m.b1_begin + m.w1];
  BatchTotals t;
  t.leaves = uint32_t(hi) - uint32_t(lo);
  t.n2 = uint32_t(hi >> 32) - uint32_t(lo >> 32);
  t.leaf_begin = uint32_t(lo);
  t.n2_begin = lo >> 32;
  t.pairs = coff[m.slot_end] - coff[m.slot_begin] + skipped[g];  // the reference's count
  for (int L = 3; L <= 5; L++)
    t.nodes[L - 3] = int(L) <= int(m.k) + 1 ? lps[g].tot[2 * L + 1] : 0u;
  out[g] = t;
  n2_beg[g] = t.n2_begin;
  if (g == G - 1) n2_beg[G] = t.n2_begin + t.n2;
}

__global__ void k_copy_totals(const BatchTotals *__restrict__ in, uint32_t G,
                              BatchTotals *__restrict__ out) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) out[g] = in[g];
}

__global__ __launch_bounds__(kSmallThreads) void k_upper_place_batch(
    const BatchMesh *__restrict__ mesh, const LevelPtrs *__restrict__ lps,
    const BatchTotals *__restrict__ tot, uint32_t *const *__restrict__ nodes_ptr) {
  const uint32_t g = blockIdx.x;
  if (tot[g].leaves == 0) return;
  upper_place(lps[g], int(mesh[g].k), lps[g].tot, nodes_ptr[g]);
}

// Level-2 placement under level-3 parents: level-2 node c of the group (nothing for a mesh
// whose node array is null: no tree placed).
__device__ __forceinline__ void place2_one(uint64_t c, uint32_t G, const uint64_t *__restrict__ n2_beg,
                                           const BatchMesh *__restrict__ mesh,
                                           const LevelPtrs *__restrict__ lps,
                                           const BatchTotals *__restrict__ tot,
                                           uint32_t *const *__restrict__ nodes_ptr,
                                           const unsigned long long *__restrict__ a_mask,
                                           const uint32_t *__restrict__ a_base,
                                           const uint32_t *__restrict__ a_code) {
  const uint32_t g = dev::upper_index(n2_beg, G, c);
  uint32_t *nodes_b = nodes_ptr[g];
  if (!nodes_b) return;
  const LevelPtrs &lp = lps[g];
  const uint32_t cl = uint32_t(c - n2_beg[g]);
  const uint32_t q = uint32_t(a_code[c] - mesh[g].b1_begin);
  const uint32_t p = lp.ni[3][q >> 6];
  const uint32_t pb = lp.a_base[3][p];
  const uint32_t mypos = lp.bs[3][p] + (cl - pb);
  const unsigned long long m = a_mask[c];
  nodes_b[3ull * mypos + 0] = uint32_t(m);
  nodes_b[3ull * mypos + 1] = uint32_t(m >> 32);
  nodes_b[3ull * mypos + 2] = uint32_t(a_base[c] - tot[g].leaf_begin);
}

// For every level-2 node of the group.
__global__ void k_place2_batch(uint64_t N2, uint32_t G, const uint64_t *__restrict__ n2_beg,
                               const BatchMesh *__restrict__ mesh,
                               const LevelPtrs *__restrict__ lps,
                               const BatchTotals *__restrict__ tot,
                               uint32_t *const *__restrict__ nodes_ptr,
                               const unsigned long long *__restrict__ a_mask,
                               const uint32_t *__restrict__ a_base,
                               const uint32_t *__restrict__ a_code) {
  const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c < N2) place2_one(c, G, n2_beg, mesh, lps, tot, nodes_ptr, a_mask, a_base, a_code);
}

struct Level {
  uint64_t win = 0;        // words of the input bitmap B_{L-1}
  unsigned long long *bin = nullptr;
  unsigned long long *bout = nullptr;  // B_L (null at the root level)
  uint32_t *cb = nullptr;  // child base per word (exclusive scan of popcounts) + total
  uint32_t *ni = nullptr;  // node index per word (exclusive scan of flags) + total
  unsigned long long *a_mask = nullptr;
  uint32_t *a_base = nullptr, *a_code = nullptr;
  uint32_t *desc = nullptr;  // desc, then in-place exclusive scan E (+ total)
  uint32_t *pos = nullptr, *bs = nullptr;
  uint64_t nodes = 0, children = 0;
};

template <class T>
T *carve(char *&p, uint64_t count) {
  T *r = reinterpret_cast<T *>(p);
  p += ((count * sizeof(T) + 255) / 256) * 256;
  return r;
}

}  // namespace

size_t bitmap_level1_words(uint32_t n) {
  uint64_t nb = uint64_t(n / 8);
  return (size_t)div_up(nb * nb * nb, 64);
}

void prepare_grid(vxl_ctx *ctx, uint32_t n) {
  const uint64_t nb = uint64_t(n / 8);
  const uint64_t bricks = nb * nb * nb;
  const size_t grid_bytes = size_t(bricks) * 64;
  const size_t w1 = bitmap_level1_words(n);
  // ctx->bitmap holds B1 only (all-zero between calls); the upper levels live in
  // levelscratch and are fully rewritten by every build.
  const size_t words = w1;
  bool fresh = false;
  if (ctx->leafgrid.bytes < grid_bytes) {
    ctx->leafgrid.reserve(grid_bytes);
    fresh = true;
  }
  if (ctx->bitmap.bytes < words * 8 + 256) {
    ctx->bitmap.reserve(words * 8 + 256);
    fresh = true;
  }
  if (fresh || ctx->leafgrid_dirty) {
    VXL_HIP(hipMemsetAsync(ctx->leafgrid.ptr, 0, ctx->leafgrid.bytes, ctx->stream));
    VXL_HIP(hipMemsetAsync(ctx->bitmap.ptr, 0, ctx->bitmap.bytes, ctx->stream));
    ctx->leafgrid_dirty = false;
  }
}

void grid_to_words(vxl_ctx *ctx, uint32_t n, uint32_t *d_words) {
  const uint64_t nn = n;
  const uint64_t nwords = div_up(nn * nn * nn, 32);
  hipLaunchKernelGGL(k_grid_to_words, dim3(div_up(nwords, kTB)), dim3(kTB), 0, ctx->stream, n,
                     nwords, ctx->leafgrid.as<unsigned long long>(), d_words);
  VXL_HIP(hipGetLastError());
  const uint64_t nb = nn / 8;
  VXL_HIP(hipMemsetAsync(ctx->leafgrid.ptr, 0, size_t(nb * nb * nb * 64), ctx->stream));
  VXL_HIP(hipMemsetAsync(ctx->bitmap.ptr, 0, bitmap_level1_words(n) * 8, ctx->stream));
}

void words_to_grid(vxl_ctx *ctx, uint32_t n, const uint32_t *d_words) {
  const uint64_t nb = uint64_t(n / 8);
  const uint64_t ngrid = nb * nb * nb * 8;
  hipLaunchKernelGGL(k_words_to_grid, dim3(div_up(ngrid, kTB)), dim3(kTB), 0, ctx->stream, n,
                     ngrid, d_words, ctx->leafgrid.as<unsigned long long>(),
                     ctx->bitmap.as<unsigned long long>());
  VXL_HIP(hipGetLastError());
}

static void *dmalloc(size_t bytes) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(VXL_OOM, std::string("tree allocation of ") + std::to_string(bytes) +
                             " bytes failed: " + hipGetErrorString(e));
  }
  return p;
}

// ---- tree memory: a context-owned slab arena ------------------------------------------
// Trees are carved from slabs of at least kSlabBytes with a bump pointer; a slab whose
// trees are all freed is reused from its start. This keeps device allocation off the
// critical path of each voxelize call (a fresh stream-ordered allocation per tree cost
// tens of microseconds of idle GPU). Callers free trees only after the context's stream
// has drained (vxl_tree_free synchronises first), so reuse never races a reader.
constexpr size_t kSlabBytes = size_t(256) << 20;

void *arena_alloc(vxl_ctx *ctx, size_t bytes, int *slab_out) {
  bytes = (bytes + 255) / 256 * 256;
  for (size_t i = 0; i < ctx->slabs.size(); i++) {
    Slab &sl = ctx->slabs[i];
    if (sl.base && sl.bytes - sl.used >= bytes) {
      void *p = sl.base + sl.used;
      sl.used += bytes;
      sl.live++;
      *slab_out = int(i);
      return p;
    }
  }
  Slab sl;
  sl.bytes = bytes > kSlabBytes ? (bytes + (size_t(2) << 20) - 1) / (size_t(2) << 20) * (size_t(2) << 20)
                                : kSlabBytes;
  sl.base = static_cast<char *>(dmalloc(sl.bytes));
  sl.used = bytes;
  sl.live = 1;
  ctx->slabs.push_back(sl);
  *slab_out = int(ctx->slabs.size() - 1);
  return sl.base;
}

void arena_release(vxl_ctx *ctx, int slab) {
  if (!ctx || slab < 0 || size_t(slab) >= ctx->slabs.size()) return;
  Slab &sl = ctx->slabs[slab];
  if (sl.live && --sl.live == 0) {
    sl.used = 0;
    // keep at most two empty slabs; give larger idle ones back
    size_t empty = 0;
    for (const Slab &o : ctx->slabs) empty += (o.base && o.live == 0) ? 1 : 0;
    if (empty > 2) {
      (void)hipFree(sl.base);
      sl.base = nullptr;
      sl.bytes = 0;
    }
  }
}

void arena_destroy(vxl_ctx *ctx) {
  for (Slab &sl : ctx->slabs)
    if (sl.base) (void)hipFree(sl.base);
  ctx->slabs.clear();
}

void tree_free_device(vxl_tree *t) {
  if (t->shared) {
    if (--t->shared->refs == 0) {
      if (t->ctx) arena_release(t->ctx, t->shared->slab);
      delete t->shared;
    }
    t->shared = nullptr;
  } else if (t->slab >= 0) {
    if (t->ctx) arena_release(t->ctx, t->slab);
    t->slab = -1;
  } else {
    if (t->nodes) (void)hipFree(t->nodes);
    if (t->leaves) (void)hipFree(t->leaves);
    if (t->bounds) (void)hipFree(t->bounds);
    if (t->codes) (void)hipFree(t->codes);
  }
  if (t->materials) (void)hipFree(t->materials);
  t->nodes = nullptr;
  t->leaves = nullptr;
  t->bounds = nullptr;
  t->codes = nullptr;
  t->materials = nullptr;
}

// One arena block for a tree's four arrays: leaves, bounds, codes, then room for node_cap
// School-B entries, so a build can allocate before it knows its exact node count (it sets
// node_count and bytes once the count is in).
static vxl_tree *alloc_tree(vxl_ctx *ctx, uint32_t n, uint32_t k, uint64_t node_cap,
                            uint64_t leaves) {
  vxl_tree *t = new vxl_tree();
  t->ctx = ctx;
  t->n = n;
  t->k = k;
  t->leaf_count = leaves;
  t->node_count = node_cap;
  auto up = [](uint64_t b) { return ((b ? b : 16) + 255) / 256 * 256; };
  const uint64_t bl = up(leaves * 64), bb = up(leaves * 4), bc = up(leaves * 8), bn = up(node_cap * 12);
  char *c = nullptr;
  try {
    c = static_cast<char *>(arena_alloc(ctx, bl + bb + bc + bn, &t->slab));
  } catch (...) {
    delete t;
    throw;
  }
  t->leaves = reinterpret_cast<uint64_t *>(c);
  t->bounds = reinterpret_cast<uint32_t *>(c + bl);
  t->codes = reinterpret_cast<uint64_t *>(c + bl + bb);
  t->nodes = reinterpret_cast<uint32_t *>(c + bl + bb + bc);
  t->bytes = node_cap * 12 + leaves * 76;
  return t;
}

// Sizes a build needs before the tree is allocated, written by the device straight into
// mapped fine-grained host memory (no copy call): leaves, level-2 nodes, the SAT pair count
// (two 32-bit halves) and the node counts of levels 3..5. Each is one 64-bit word, the value
// in the low 48 bits and the build's tag (1..65535) in the high 16. The host zeroes the words
// before it queues the size kernel; the device stores each word with one relaxed
// system-scope atomic store, with no release fence (a system-scope fence writes the L2 back
// first, about 4 us here); the host takes the set once every word carries the build's tag.
// A word of an earlier build, or a set only partly written, is never taken.
enum : int { kSzLeaves, kSzLevel2, kSzPairsLo, kSzPairsHi, kSzLevel3, kSzWords = kSzLevel3 + 3 };
struct BuildSizes {
  uint64_t w[kSzWords];
};
constexpr uint64_t kSzValue = (uint64_t(1) << 48) - 1;

__device__ __forceinline__ void put_size(BuildSizes *out, int i, uint64_t v, uint64_t tag) {
  __hip_atomic_store(&out->w[i], v | (tag << 48), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- single-tree level 2 without a device-wide scan ------------------------------------
// k_level2_sums leaves each block's sum of the packed (popcount | nonzero << 32) values of
// its B1 words (a block's popcounts sum to at most 256 * 64, so the halves never carry);
// k_upper_emit's first block reduces those sums to the sizes the host reads before the
// upper levels; its other blocks rebuild each word's packed offset from a block-local scan
// plus the sum of the block sums before the block.
__device__ __forceinline__ uint64_t packed_word(unsigned long long w) {
  return uint64_t(__popcll(w)) | (uint64_t(w != 0ull) << 32);
}

// Exclusive scan of one u64 per thread over a block of NT threads; returns the total.
template <int NT>
__device__ __forceinline__ uint64_t block_scan_u64(uint64_t v, uint64_t *excl) {
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
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kTB) void k_level2_sums(const unsigned long long *__restrict__ bin,
                                                       uint64_t win,
                                                       uint64_t *__restrict__ block_sums,
                                                       unsigned long long *__restrict__ bout) {
  const uint64_t j = uint64_t(blockIdx.x) * kTB + threadIdx.x;
  const unsigned long long w = j < win ? bin[j] : 0ull;
  const unsigned long long b = __ballot(w != 0ull);
  if (bout && (threadIdx.x & 63) == 0 && j < win) bout[j >> 6] = b;
  uint64_t excl;
  const uint64_t total = block_scan_u64<kTB>(packed_word(w), &excl);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Exclusive scan of one u64 per lane over one wave; returns the wave's total.
__device__ __forceinline__ uint64_t wave_scan_u64(uint64_t v, uint64_t *excl) {
  const int lane = threadIdx.x & 63;
  uint64_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t o = uint64_t(__shfl_up((long long)inc, d, 64));
    if (lane >= d) inc += o;
  }
  *excl = inc - v;
  return uint64_t(__shfl((long long)inc, 63, 64));
}

// The upper levels of one tree with k = top - 1 <= 4, in one workgroup: the total of the
// level-2 block sums (at most kSmallThreads of them) gives the leaf and level-2 node counts,
// which go to the host-mapped BuildSizes at once (the host allocates the tree with them);
// then level 3 (at most kSmallWords input words, kW per thread) with one packed scan,
// levels 4 and 5 (at most 64 input words) in wave 0, their node counts (to the host and to
// `upper` for the node-entry pass), and the School-B positions of levels 3..top.
// Everything the positions read stays in LDS (level 3's child bases are its exclusive
// subtree-size scan E_3, since a level-3 node's subtree size is its child count), so the
// chain has no dependent global loads after the first.
__device__ __forceinline__ void upper_block(const LevelPtrs &lp, int top,
                                            const unsigned long long *__restrict__ b2,
                                            const uint64_t *__restrict__ block_sums, uint32_t nb2,
                                            const uint64_t *__restrict__ pairs,
                                            BuildSizes *__restrict__ out, uint32_t tag,
                                            uint32_t *__restrict__ upper) {
  constexpr int kW = int(kSmallWords / kSmallThreads);
  static_assert(64 % kW == 0, "a thread's level-3 words share one level-4 input word");
  __shared__ uint32_t s_base3[kSmallWords + 1];  // a_base of level 3 (= E_3); [n3] = n2
  __shared__ uint16_t s_code3[kSmallWords];
  __shared__ unsigned long long s_w4[64], s_mask4[64];
  __shared__ uint32_t s_base4[64], s_ni4[64], s_bs4[64], s_n[2];
  __shared__ unsigned long long s_mask5;
  __shared__ uint64_t s_pairs;
  const uint32_t tid = threadIdx.x;
  const uint64_t win3 = top >= 3 ? lp.win[3] : 0;
  // every global read of the kernel, issued together
  if (tid == 0) s_pairs = pairs ? *pairs : 0;
  const uint64_t v = tid < nb2 ? block_sums[tid] : 0;
  unsigned long long w[kW];
#pragma unroll
  for (int i = 0; i < kW; i++) {
    const uint64_t j = uint64_t(tid) * kW + i;
    w[i] = j < win3 ? b2[j] : 0ull;
  }
  if (tid < 64) s_w4[tid] = 0ull;
  uint64_t excl;
  const uint64_t pn2_total = block_scan_u64<kSmallThreads>(v, &excl);
  if (tid == 0) {  // s_pairs was written before the scan's barriers
    put_size(out, kSzLeaves, pn2_total & 0xffffffffull, tag);
    put_size(out, kSzLevel2, pn2_total >> 32, tag);
    put_size(out, kSzPairsLo, s_pairs & 0xffffffffull, tag);
    put_size(out, kSzPairsHi, s_pairs >> 32, tag);
  }
  uint64_t packed = 0;
#pragma unroll
  for (int i = 0; i < kW; i++) packed += packed_word(w[i]);
  uint64_t e3;
  const uint64_t t3 = block_scan_u64<kSmallThreads>(packed, &e3);  // zero when top < 3
  const uint32_t n2 = uint32_t(t3), n3 = uint32_t(t3 >> 32);
  // level 3: per input word its node index (read by the level-2 placement), per node its
  // mask, child base and code; the level-4 input bits
  uint32_t epop = uint32_t(e3), enz = uint32_t(e3 >> 32);
  unsigned long long bits = 0;
#pragma unroll
  for (int i = 0; i < kW; i++) {
    const uint32_t j = tid * kW + i;
    if (j < win3) lp.ni[3][j] = enz;
    if (w[i]) {
      lp.a_mask[3][enz] = w[i];
      lp.a_base[3][enz] = epop;
      s_base3[enz] = epop;
      s_code3[enz] = uint16_t(j);
      bits |= 1ull << (j & 63);
    }
    epop += __popcll(w[i]);
    enz += w[i] != 0ull;
  }
  if (bits) atomicOr(&s_w4[(tid * kW) >> 6], bits);
  if (tid == 0) s_base3[n3] = n2;
  __syncthreads();
  // levels 4 and 5 in wave 0: lane j holds level-4 input word j
  if (top >= 4 && tid < 64) {
    const uint64_t win4 = lp.win[4];
    const unsigned long long x = tid < win4 ? s_w4[tid] : 0ull;
    uint64_t e4;
    const uint64_t t4 = wave_scan_u64(packed_word(x), &e4);
    const uint32_t i4 = uint32_t(e4 >> 32);
    s_ni4[tid] = i4;
    if (x) {
      s_mask4[i4] = x;
      s_base4[i4] = uint32_t(e4);
      lp.a_mask[4][i4] = x;
      lp.a_base[4][i4] = uint32_t(e4);
      lp.a_code[4][i4] = tid;
    }
    const unsigned long long x5 = __ballot(x != 0ull);  // the one level-5 input word
    if (tid == 0) {
      s_n[0] = uint32_t(t4 >> 32);
      s_n[1] = (top >= 5 && x5) ? 1u : 0u;
      s_mask5 = x5;
      if (top >= 5 && x5) {
        lp.a_mask[5][0] = x5;
        lp.a_base[5][0] = 0;
        lp.a_code[5][0] = 0;
      }
    }
  }
  __syncthreads();
  const uint32_t n4 = top >= 4 ? s_n[0] : 0u, n5 = top >= 5 ? s_n[1] : 0u;
  if (tid == 0) {
    put_size(out, kSzLevel3, n3, tag);
    put_size(out, kSzLevel3 + 1, n4, tag);
    put_size(out, kSzLevel3 + 2, n5, tag);
    upper[0] = n3;
    upper[1] = n4;
    upper[2] = n5;
  }
  if (top < 3) return;
  // positions. k = 4: the root (level 5) at 0 with its block at 1; level-4 node c, the
  // root's child c, at 1 + c with its block after the root's block and the subtrees of the
  // children before it; subtree size of a level-4 node = its child count + the level-3
  // subtree sizes under it, E_3[b + m] - E_3[b]
  if (top == 5 && tid < 64) {
    uint32_t d = 0;
    if (tid < n4) {
      const uint32_t m = __popcll(s_mask4[tid]), b = s_base4[tid];
      d = m + s_base3[b + m] - s_base3[b];
    }
    uint64_t e4;
    wave_scan_u64(d, &e4);
    if (tid < n4) {
      const uint32_t bs = 1u + uint32_t(__popcll(s_mask5)) + uint32_t(e4);
      s_bs4[tid] = bs;
      lp.pos[4][tid] = 1u + tid;
      lp.bs[4][tid] = bs;
    }
    if (tid == 0) {
      lp.pos[5][0] = 0;
      lp.bs[5][0] = 1;
    }
  }
  if (top == 4 && tid == 0) {
    lp.pos[4][0] = 0;
    lp.bs[4][0] = 1;
  }
  __syncthreads();
  // level 3: under its level-4 parent (k = 4), as the root's children (k = 3), or the root
  const uint32_t mp4 = top == 4 ? uint32_t(__popcll(s_w4[0])) : 0u;  // root mask of k = 3
  for (uint32_t c = tid; c < n3; c += kSmallThreads) {
    uint32_t pos, bs;
    if (top == 5) {
      const uint32_t p = s_ni4[s_code3[c] >> 6];
      const uint32_t pb = s_base4[p], pbs = s_bs4[p];
      pos = pbs + (c - pb);
      bs = pbs + uint32_t(__popcll(s_mask4[p])) + (s_base3[c] - s_base3[pb]);
    } else if (top == 4) {
      pos = 1u + c;
      bs = 1u + mp4 + s_base3[c];
    } else {
      pos = 0;
      bs = 1;
    }
    lp.pos[3][c] = pos;
    lp.bs[3][c] = bs;
  }
}

// Level-2 School-A arrays of B1 words [e * kSmallThreads, (e + 1) * kSmallThreads): the
// block's offset is the sum of the level-2 block sums (kTB words each) before its first
// word, then a block-local scan of the packed words.
__device__ __forceinline__ void emit_block(uint32_t e, const unsigned long long *__restrict__ b1,
                                           uint64_t w1, const uint64_t *__restrict__ block_sums,
                                           unsigned long long *__restrict__ a_mask,
                                           uint32_t *__restrict__ a_base,
                                           uint32_t *__restrict__ a_code) {
  static_assert(kSmallThreads % kTB == 0, "an emit block covers whole level-2 blocks");
  const uint32_t tid = threadIdx.x;
  const uint64_t j = uint64_t(e) * kSmallThreads + tid;
  const unsigned long long w = j < w1 ? b1[j] : 0ull;
  const uint32_t before = e * uint32_t(kSmallThreads / kTB);  // at most nb2 <= kSmallThreads
  uint64_t unused, excl;
  const uint64_t off = block_scan_u64<kSmallThreads>(tid < before ? block_sums[tid] : 0ull, &unused);
  block_scan_u64<kSmallThreads>(packed_word(w), &excl);
  if (!w) return;
  const uint64_t v = off + excl;
  const uint32_t i = uint32_t(v >> 32);
  a_mask[i] = w;
  a_base[i] = uint32_t(v);
  a_code[i] = uint32_t(j);
}

// Block 0: the upper levels and the sizes (upper_block); blocks 1..: the level-2 School-A
// arrays (emit_block). Both read only what k_level2_sums wrote.
__global__ __launch_bounds__(kSmallThreads) void k_upper_emit(
    LevelPtrs lp, int top, const unsigned long long *__restrict__ b1,
    const unsigned long long *__restrict__ b2, const uint64_t *__restrict__ block_sums,
    uint32_t nb2, const uint64_t *__restrict__ pairs, BuildSizes *__restrict__ out, uint32_t tag,
    uint32_t *__restrict__ upper) {
  if (blockIdx.x == 0)
    upper_block(lp, top, b2, block_sums, nb2, pairs, out, tag, upper);
  else
    emit_block(blockIdx.x - 1, b1, lp.win[2], block_sums, lp.a_mask[2], lp.a_base[2],
               lp.a_code[2]);
}

// Waits until size words [first, first + count) carry `tag`, so the host goes on while the
// kernel that writes them runs. Every 256 polls the event recorded after those kernels is
// checked too: a failed launch ends the wait with its error, and finished kernels whose
// sizes are still missing are an error rather than an endless wait.
static void wait_sizes(const BuildSizes *hs, uint64_t tag, hipEvent_t ev, int first, int count,
                       uint64_t *v) {
  const volatile uint64_t *w = hs->w;
  auto all_in = [&] {
    for (int j = first; j < first + count; j++)
      if ((w[j] >> 48) != tag) return false;
    return true;
  };
  for (uint32_t i = 1;; i++) {
    if (all_in()) break;
    if ((i & 255u) == 0) {
      const hipError_t e = hipEventQuery(ev);
      if (e == hipSuccess) {
        if (all_in()) break;
        throw Error(VXL_DEVICE_ERROR, "tree build: the size kernel finished without its sizes");
      }
      if (e != hipErrorNotReady) {
        (void)hipGetLastError();
        throw Error(VXL_DEVICE_ERROR, std::string("tree build failed: ") + hipGetErrorString(e));
      }
    }
  }
  for (int j = first; j < first + count; j++) v[j] = w[j] & kSzValue;
}

// n <= 2048 (k <= 4): every level above 2 has at most 4096 input words.
static vxl_tree *build_tree_fast(vxl_ctx *ctx, uint32_t n, uint32_t k, const uint64_t *pairs_dev,
                                 uint64_t *pairs) {
  hipStream_t s = ctx->stream;
  const int top = int(k + 1 > 2 ? k + 1 : 2);
  LevelPtrs lp;
  std::memset(&lp, 0, sizeof(lp));
  uint64_t w = bitmap_level1_words(n);
  for (int L = 2; L <= top; L++) {
    lp.win[L] = w;
    w = div_up(w, 64);
  }
  const uint64_t w1 = lp.win[2];
  size_t scratch = 4096 + ((w1 + 1) * 8 + 255) / 256 * 256 + (lp.win[3] * 8 + 255) / 256 * 256;
  for (int L = 2; L <= top; L++) scratch += 10 * (((lp.win[L] + 1) * 8 + 255) / 256) * 256;
  ctx->levelscratch.reserve(scratch);
  char *p = static_cast<char *>(ctx->levelscratch.ptr);
  const uint32_t nb2 = uint32_t(div_up(w1, kTB));  // level-2 blocks
  if (nb2 > uint32_t(kSmallThreads))  // n <= 2048: (256^3 / 64) / 256 = 1024 blocks at most
    throw Error(VXL_INVALID_ARG, "tree build: level-2 block count exceeds one workgroup");
  uint64_t *l2sums = carve<uint64_t>(p, nb2 + 1);
  unsigned long long *b2 = top >= 3 ? carve<unsigned long long>(p, lp.win[3]) : nullptr;
  for (int L = 2; L <= top; L++) {
    const uint64_t e = lp.win[L] + 1;
    lp.cb[L] = carve<uint32_t>(p, e);
    lp.ni[L] = carve<uint32_t>(p, e);
    lp.a_mask[L] = carve<unsigned long long>(p, e);
    lp.a_base[L] = carve<uint32_t>(p, e);
    lp.a_code[L] = carve<uint32_t>(p, e);
    lp.desc[L] = carve<uint32_t>(p, e);
    lp.pos[L] = carve<uint32_t>(p, e);
    lp.bs[L] = carve<uint32_t>(p, e);
  }
  uint32_t *upper = carve<uint32_t>(p, 4);  // node counts of levels 3..5, on the device
  unsigned long long *b1 = ctx->bitmap.as<unsigned long long>();
  // 1. level 2: per-block sums of the packed popcount/flag values of B1, B2 by ballot
  hipLaunchKernelGGL(k_level2_sums, dim3(nb2), dim3(kTB), 0, s, b1, w1, l2sums, b2);
  // 2. the sizes, upper levels and positions (block 0) beside the level-2 School-A arrays
  BuildSizes *hs = static_cast<BuildSizes *>(ctx->host_pinned);
  const uint32_t tag = ctx->size_seq++ % 65535u + 1u;
  // no size word holds this build's tag before its size kernel writes it (other read-backs
  // share the area, so it is cleared here, before the kernel is queued)
  for (int j = 0; j < kSzWords; j++) reinterpret_cast<volatile uint64_t *>(hs->w)[j] = 0ull;
  const uint32_t emit_blocks = uint32_t(div_up(w1, kSmallThreads));
  hipLaunchKernelGGL(k_upper_emit, dim3(1 + emit_blocks), dim3(kSmallThreads), 0, s, lp, top, b1,
                     b2, l2sums, nb2, pairs_dev, static_cast<BuildSizes *>(ctx->host_pinned_dev),
                     tag, upper);
  VXL_HIP(hipGetLastError());
  if (!ctx->ev_wait) VXL_HIP(hipEventCreateWithFlags(&ctx->ev_wait, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ctx->ev_wait, s));
  // 3. leaves, level-2 nodes and pairs: the first words the kernel writes, and all the
  //    allocation needs (the node array gets room for the most nodes n2 level-2 nodes can
  //    have above them, at most one per input word of each level)
  uint64_t sz[kSzWords];
  wait_sizes(hs, tag, ctx->ev_wait, kSzLeaves, kSzLevel3 - kSzLeaves, sz);
  const uint64_t leaves = sz[kSzLeaves];
  const uint64_t n2 = sz[kSzLevel2];
  if (pairs) *pairs = sz[kSzPairsLo] | (sz[kSzPairsHi] << 32);
  // the kernels that follow index per-level scratch and the new tree by these sizes: refuse
  // any that the grid cannot hold rather than launch them out of bounds
  const uint64_t bricks = uint64_t(n / 8) * (n / 8) * (n / 8);
  if (!(leaves <= bricks && n2 <= w1 && n2 <= leaves && (leaves == 0) == (n2 == 0)))
    throw Error(VXL_DEVICE_ERROR, "tree build: the sizes read back are inconsistent");
  uint64_t cap = 0;
  if (k >= 1 && leaves) {
    cap = n2;
    for (int L = 3; L <= int(k) + 1; L++) cap += std::min<uint64_t>(n2, lp.win[L]);
  }
  vxl_tree *t = alloc_tree(ctx, n, k, cap, leaves);
  if (leaves == 0) return t;
  try {
    // 4. leaf gather (clears the grid and B1) and the School-B node entries
    const uint32_t gather_blocks = uint32_t(div_up(n2 * 64, kTB));
    if (k >= 2) {
      hipLaunchKernelGGL(k_gather_place, dim3(gather_blocks + div_up(cap, kTB)), dim3(kTB), 0, s,
                         n2, gather_blocks, lp, upper, cap,
                         ctx->leafgrid.as<unsigned long long>(), b1,
                         reinterpret_cast<unsigned long long *>(t->leaves), t->bounds,
                         reinterpret_cast<unsigned long long *>(t->codes), t->nodes);
    } else {
      hipLaunchKernelGGL(k_leaf_gather, dim3(gather_blocks), dim3(kTB), 0, s, n2, lp.a_mask[2],
                         lp.a_base[2], lp.a_code[2], ctx->leafgrid.as<unsigned long long>(), b1,
                         reinterpret_cast<unsigned long long *>(t->leaves), t->bounds,
                         reinterpret_cast<unsigned long long *>(t->codes), nullptr, 0u, nullptr);
      if (k == 1)
        hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, s, lp.a_mask[2], lp.a_base[2], 1,
                           t->nodes, lp.pos[2], lp.bs[2]);
    }
    VXL_HIP(hipGetLastError());
    // 5. the node counts of levels 3..5, written by the size kernel after the first words
    wait_sizes(hs, tag, ctx->ev_wait, kSzLevel3, kSzWords - kSzLevel3, sz);
    uint64_t level_nodes[16] = {0};
    level_nodes[2] = n2;
    for (int L = 3; L <= top; L++) level_nodes[L] = sz[kSzLevel3 + L - 3];
    uint64_t nodes = 0;
    bool sane = true;
    for (int L = 3; L <= top; L++) sane = sane && level_nodes[L] <= lp.win[L] && level_nodes[L] <= n2;
    if (k >= 1)
      for (int L = 2; L <= int(k) + 1; L++) nodes += level_nodes[L];
    if (!sane || nodes > cap)
      throw Error(VXL_DEVICE_ERROR, "tree build: the node counts read back are inconsistent");
    t->node_count = nodes;
    t->bytes = nodes * 12 + leaves * 76;
    if (k >= 1)
      for (int L = 2; L <= int(k) + 1; L++) t->level_nodes[L] = level_nodes[L];
  } catch (...) {
    tree_free_device(t);
    delete t;
    throw;
  }
  return t;
}

vxl_tree *tree_copy_arrays(vxl_ctx *ctx, uint32_t n, uint32_t k, uint64_t leaves,
                           uint64_t nodes, const uint64_t *level_nodes, const void *src_nodes,
                           const void *src_leaves, const void *src_bounds,
                           const void *src_codes, hipStream_t s) {
  if (!leaves) return empty_tree_device(ctx, n, k);
  vxl_tree *t = alloc_tree(ctx, n, k, nodes, leaves);
  try {
    VXL_HIP(hipMemcpyAsync(t->nodes, src_nodes, nodes * 12, hipMemcpyDeviceToDevice, s));
    VXL_HIP(hipMemcpyAsync(t->leaves, src_leaves, leaves * 64, hipMemcpyDeviceToDevice, s));
    VXL_HIP(hipMemcpyAsync(t->bounds, src_bounds, leaves * 4, hipMemcpyDeviceToDevice, s));
    VXL_HIP(hipMemcpyAsync(t->codes, src_codes, leaves * 8, hipMemcpyDeviceToDevice, s));
  } catch (...) {
    tree_free_device(t);
    delete t;
    throw;
  }
  for (int L = 0; L < 16; L++) t->level_nodes[L] = level_nodes[L];
  return t;
}

vxl_tree *empty_tree_device(vxl_ctx *ctx, uint32_t n, uint32_t k) {
  return alloc_tree(ctx, n, k, 0, 0);
}

void build_trees_batch(vxl_ctx *ctx, const std::vector<BatchMesh> &meshes,
                       const BatchMesh *d_mesh, uint64_t W1, vxl_tree **trees, uint64_t *pairs,
                       const unsigned long long *d_ref_pairs) {
  hipStream_t s = ctx->stream;
  const uint32_t G = uint32_t(meshes.size());
  auto up = [](uint64_t b) { return (b + 255) / 256 * 256; };
  // scratch: group arrays, then per-mesh upper levels
  size_t need = up((W1 + 1) * 8) + up(W1 / 64 * 8 + 8) + up(G * sizeof(BatchTotals)) +
                up((G + 1) * 8) + up(G * 8) + up(G * sizeof(LevelPtrs));
  std::vector<LevelPtrs> lps(G);
  for (uint32_t g = 0; g < G; g++) {
    const BatchMesh &m = meshes[g];
    std::memset(&lps[g], 0, sizeof(LevelPtrs));
    uint64_t w = m.w1;
    for (int L = 2; L <= int(m.k) + 1; L++) {
      lps[g].win[L] = w;
      w = div_up(w, 64);
      if (L >= 3) need += 8 * up((lps[g].win[L] + 1) * 8);
    }
    need += up(64 * 4);
  }
  ctx->batch_scratch.reserve(need);
  ctx->batch_l2.reserve(up(W1 * 8 + 8) + 2 * up(W1 * 4 + 4));
  char *p = static_cast<char *>(ctx->batch_scratch.ptr);
  uint64_t *pn2 = carve<uint64_t>(p, W1 + 1);
  unsigned long long *b2 = carve<unsigned long long>(p, W1 / 64 + 1);
  BatchTotals *d_tot = carve<BatchTotals>(p, G);
  uint64_t *n2_beg = carve<uint64_t>(p, G + 1);
  uint32_t **nodes_ptr = carve<uint32_t *>(p, G);
  LevelPtrs *d_lps = carve<LevelPtrs>(p, G);
  for (uint32_t g = 0; g < G; g++) {
    LevelPtrs &lp = lps[g];
    for (int L = 3; L <= int(meshes[g].k) + 1; L++) {
      const uint64_t e = lp.win[L] + 1;
      lp.cb[L] = carve<uint32_t>(p, e * 2);
      lp.ni[L] = carve<uint32_t>(p, e * 2);
      lp.a_mask[L] = carve<unsigned long long>(p, e);
      lp.a_base[L] = carve<uint32_t>(p, e * 2);
      lp.a_code[L] = carve<uint32_t>(p, e * 2);
      lp.desc[L] = carve<uint32_t>(p, e * 2);
      lp.pos[L] = carve<uint32_t>(p, e * 2);
      lp.bs[L] = carve<uint32_t>(p, e * 2);
    }
    lp.tot = carve<uint32_t>(p, 64);
  }
  char *q = static_cast<char *>(ctx->batch_l2.ptr);
  unsigned long long *a_mask2 = carve<unsigned long long>(q, W1 + 1);
  uint32_t *a_base2 = carve<uint32_t>(q, W1 + 1);
  uint32_t *a_code2 = carve<uint32_t>(q, W1 + 1);

  LevelPtrs *h_lps = static_cast<LevelPtrs *>(staging_take(ctx, G * sizeof(LevelPtrs)));
  std::memcpy(h_lps, lps.data(), G * sizeof(LevelPtrs));
  VXL_HIP(hipMemcpyAsync(d_lps, h_lps, G * sizeof(LevelPtrs), hipMemcpyHostToDevice, s));

  unsigned long long *b1 = ctx->pool_b1.as<unsigned long long>();
  // 1. level 2 over the whole group, then the per-mesh upper levels
  hipLaunchKernelGGL(k_level2, dim3(div_up(W1, kTB)), dim3(kTB), 0, s, b1, W1, pn2, b2);
  scan_exclusive_u64(pn2, pn2, W1, pn2 + W1, ctx->scan_tmp, s);
  hipLaunchKernelGGL(k_small_levels_batch, dim3(G), dim3(kSmallThreads), 0, s, d_mesh, d_lps, b2);
  hipLaunchKernelGGL(k_batch_totals, dim3(div_up(G, 64)), dim3(64), 0, s, d_mesh, d_lps, G, pn2,
                     ctx->coff.as<uint64_t>(), d_ref_pairs, d_tot, n2_beg);
  VXL_HIP(hipGetLastError());
  // 2. per-mesh sizes (the one host round trip of the group): k_batch_totals also wrote
  //    them into mapped pinned staging
  BatchTotals *h_tot = static_cast<BatchTotals *>(staging_take(ctx, G * sizeof(BatchTotals)));
  hipLaunchKernelGGL(k_copy_totals, dim3(div_up(G, 64)), dim3(64), 0, s, d_tot, G,
                     static_cast<BatchTotals *>(staging_dev(ctx, h_tot)));
  wait_stream(ctx, s);
  // 3. one allocation for the group's trees: node arrays, then leaves / bounds / codes of
  //    every mesh in group order (a mesh's leaves start at its leaf_begin)
  uint64_t total_leaves = 0, node_bytes = 0, N2 = 0, live = 0;
  std::vector<uint64_t> nodes(G, 0), node_off(G, 0);
  for (uint32_t g = 0; g < G; g++) {
    const BatchTotals &t = h_tot[g];
    if (pairs) pairs[g] = t.pairs;
    total_leaves += t.leaves;
    N2 += t.n2;
    if (!t.leaves) continue;
    live++;
    nodes[g] = t.n2;
    for (int L = 3; L <= int(meshes[g].k) + 1; L++) nodes[g] += t.nodes[L - 3];
    node_off[g] = node_bytes;
    node_bytes += up(nodes[g] * 12);
  }
  SharedBlock *blk = nullptr;
  char *base = nullptr;
  if (live) {
    const uint64_t bytes = node_bytes + up(total_leaves * 64) + up(total_leaves * 4) +
                           up(total_leaves * 8);
    int slab = -1;
    void *mem = arena_alloc(ctx, bytes, &slab);
    blk = new SharedBlock();
    blk->base = mem;
    blk->slab = slab;
    base = static_cast<char *>(mem);
  }
  char *leaves_base = base + node_bytes;
  char *bounds_base = leaves_base + up(total_leaves * 64);
  char *codes_base = bounds_base + up(total_leaves * 4);
  uint32_t **h_nodes = static_cast<uint32_t **>(staging_take(ctx, G * sizeof(uint32_t *)));
  for (uint32_t g = 0; g < G; g++) {
    const BatchTotals &t = h_tot[g];
    const BatchMesh &m = meshes[g];
    h_nodes[g] = nullptr;
    if (!t.leaves) {
      trees[g] = alloc_tree(ctx, m.n, m.k, 0, 0);
      continue;
    }
    vxl_tree *tr = new vxl_tree();
    tr->ctx = ctx;
    tr->n = m.n;
    tr->k = m.k;
    tr->leaf_count = t.leaves;
    tr->node_count = nodes[g];
    tr->level_nodes[2] = t.n2;
    for (int L = 3; L <= int(m.k) + 1; L++) tr->level_nodes[L] = t.nodes[L - 3];
    tr->nodes = reinterpret_cast<uint32_t *>(base + node_off[g]);
    tr->leaves = reinterpret_cast<uint64_t *>(leaves_base) + t.leaf_begin * 8;
    tr->bounds = reinterpret_cast<uint32_t *>(bounds_base) + t.leaf_begin;
    tr->codes = reinterpret_cast<uint64_t *>(codes_base) + t.leaf_begin;
    tr->bytes = nodes[g] * 12 + t.leaves * 76;
    tr->shared = blk;
    blk->refs++;
    trees[g] = tr;
    h_nodes[g] = tr->nodes;
  }
  if (!live) return;
  VXL_HIP(hipMemcpyAsync(nodes_ptr, h_nodes, G * sizeof(uint32_t *), hipMemcpyHostToDevice, s));
  // 4. level-2 School-A arrays, leaf gather (clears the pooled grid and B1), placement
  hipLaunchKernelGGL(k_emit_a2, dim3(div_up(W1, kTB)), dim3(kTB), 0, s, b1, W1, pn2, a_mask2,
                     a_base2, a_code2);
  hipLaunchKernelGGL(k_leaf_gather, dim3(div_up(N2 * 64, kTB)), dim3(kTB), 0, s, N2, a_mask2,
                     a_base2, a_code2, ctx->pool_grid.as<unsigned long long>(), b1,
                     reinterpret_cast<unsigned long long *>(leaves_base),
                     reinterpret_cast<uint32_t *>(bounds_base),
                     reinterpret_cast<unsigned long long *>(codes_base), n2_beg, G, d_mesh);
  hipLaunchKernelGGL(k_upper_place_batch, dim3(G), dim3(kSmallThreads), 0, s, d_mesh, d_lps,
                     d_tot, nodes_ptr);
  hipLaunchKernelGGL(k_place2_batch, dim3(div_up(N2, kTB)), dim3(kTB), 0, s, N2, G, n2_beg,
                     d_mesh, d_lps, d_tot, nodes_ptr, a_mask2, a_base2, a_code2);
  VXL_HIP(hipGetLastError());
}

// ---- batch groups of a graph: the sizes stay on the device -------------------------
namespace {

// Places a group's trees in its arena region: mesh node arrays one after another (each
// 256-byte aligned, as the host-sized path lays them out), leaves / bounds / codes at each
// mesh's group-wide leaf_begin. When the group's trees do not fit the region, st->ok = 0
// and nothing is placed: every node pointer and segment tree is null, so the gather only
// clears the pooled grid and the traversal reports misses. One wave.
__global__ __launch_bounds__(64) void k_graph_place(
    uint32_t G, const BatchMesh *__restrict__ mesh, const BatchTotals *__restrict__ tot,
    char *node_region, uint64_t node_cap, char *leaves_base, char *bounds_base,
    char *codes_base, uint64_t leaf_cap, uint32_t **__restrict__ nodes_ptr,
    uint64_t *__restrict__ node_off, TraverseSeg *__restrict__ segs,
    GraphState *__restrict__ st) {
  const uint32_t lane = threadIdx.x;
  uint64_t carry = 0;
  for (uint32_t g0 = 0; g0 < G; g0 += 64) {
    const uint32_t g = g0 + lane;
    uint64_t bytes = 0;
    if (g < G && tot[g].leaves) {
      uint64_t nn = tot[g].n2;
      for (int L = 3; L <= int(mesh[g].k) + 1; L++) nn += tot[g].nodes[L - 3];
      bytes = (nn * 12 + 255) / 256 * 256;
    }
    uint64_t excl;
    const uint64_t sum = wave_scan_u64(bytes, &excl);
    if (g < G) node_off[g] = carry + excl;
    carry += sum;
  }
  const uint64_t leaves = G ? tot[G - 1].leaf_begin + tot[G - 1].leaves : 0;
  const bool ok = carry <= node_cap && leaves <= leaf_cap;
  for (uint32_t g = lane; g < G; g += 64) {
    const BatchTotals &t = tot[g];
    const bool live = ok && t.leaves != 0;
    uint32_t *np = live ? reinterpret_cast<uint32_t *>(node_region + node_off[g]) : nullptr;
    nodes_ptr[g] = np;
    TraverseSeg &sg = segs[g];
    sg.nodes = np;
    sg.leafw = live ? reinterpret_cast<const uint32_t *>(leaves_base + t.leaf_begin * 64) : nullptr;
    sg.bounds = live ? reinterpret_cast<const uint32_t *>(bounds_base) + t.leaf_begin : nullptr;
    sg.leaf_count = live ? t.leaves : 0;
  }
  if (lane == 0) {
    st->ok = ok ? 1u : 0u;
    st->node_bytes = carry;
    st->leaves = leaves;
  }
}

// k_leaf_gather over a level-2 node count that is on the device: a fixed grid whose waves
// stride over the nodes. The grid words are cleared whether or not the trees fit the arena,
// so the pooled grid is all-zero for the next replay either way.
__global__ __launch_bounds__(kTB) void k_leaf_gather_graph(
    const uint64_t *__restrict__ n2_beg, uint32_t G, const GraphState *__restrict__ st,
    const unsigned long long *__restrict__ a_mask, const uint32_t *__restrict__ a_base,
    const uint32_t *__restrict__ a_code, unsigned long long *__restrict__ leafgrid,
    unsigned long long *__restrict__ brickbits, unsigned long long *__restrict__ leaves,
    uint32_t *__restrict__ bounds, unsigned long long *__restrict__ codes,
    const BatchMesh *__restrict__ mesh) {
  const uint64_t n2 = n2_beg[G];
  const bool store = st->ok != 0;
  const uint64_t waves = uint64_t(gridDim.x) * (kTB / 64);
  for (uint64_t node = (uint64_t(blockIdx.x) * kTB + threadIdx.x) >> 6; node < n2; node += waves)
    gather_node(node, n2, a_mask, a_base, a_code, leafgrid, brickbits, leaves, bounds, codes,
                n2_beg, G, mesh, store);
}

// k_place2_batch over a level-2 node count that is on the device (a fixed, striding grid).
__global__ __launch_bounds__(kTB) void k_place2_graph(
    uint32_t G, const uint64_t *__restrict__ n2_beg, const BatchMesh *__restrict__ mesh,
    const LevelPtrs *__restrict__ lps, const BatchTotals *__restrict__ tot,
    uint32_t *const *__restrict__ nodes_ptr, const unsigned long long *__restrict__ a_mask,
    const uint32_t *__restrict__ a_base, const uint32_t *__restrict__ a_code) {
  const uint64_t N2 = n2_beg[G];
  const uint64_t stride = uint64_t(gridDim.x) * kTB;
  for (uint64_t c = uint64_t(blockIdx.x) * kTB + threadIdx.x; c < N2; c += stride)
    place2_one(c, G, n2_beg, mesh, lps, tot, nodes_ptr, a_mask, a_base, a_code);
}

uint64_t up256(uint64_t b) { return (b + 255) / 256 * 256; }

// The upper-level arrays of one mesh (levels 3..k+1; level 2 is group-wide), as
// build_trees_batch carves them.
uint64_t mesh_levels(const BatchMesh &m, char **p, LevelPtrs *lp) {
  uint64_t bytes = 0, w = m.w1;
  if (lp) std::memset(lp, 0, sizeof(LevelPtrs));
  for (int L = 2; L <= int(m.k) + 1; L++) {
    const uint64_t win = w;
    w = div_up(w, 64);
    if (lp) lp->win[L] = win;
    if (L < 3) continue;
    const uint64_t e = win + 1;
    bytes += 8 * up256(e * 8);
    if (lp) {
      lp->cb[L] = carve<uint32_t>(*p, e * 2);
      lp->ni[L] = carve<uint32_t>(*p, e * 2);
      lp->a_mask[L] = carve<unsigned long long>(*p, e);
      lp->a_base[L] = carve<uint32_t>(*p, e * 2);
      lp->a_code[L] = carve<uint32_t>(*p, e * 2);
      lp->desc[L] = carve<uint32_t>(*p, e * 2);
      lp->pos[L] = carve<uint32_t>(*p, e * 2);
      lp->bs[L] = carve<uint32_t>(*p, e * 2);
    }
  }
  if (lp) lp->tot = carve<uint32_t>(*p, 64);
  return bytes + up256(64 * 4);
}

}  // namespace

size_t group_scratch_bytes(const std::vector<BatchMesh> &meshes, uint64_t W1) {
  size_t need = up256((W1 + 1) * 8) + up256((W1 / 64 + 1) * 8) + up256((W1 + 1) * 8) +
                2 * up256((W1 + 1) * 4);
  for (const BatchMesh &m : meshes) need += mesh_levels(m, nullptr, nullptr);
  return need;
}

GroupScratch carve_group_scratch(const std::vector<BatchMesh> &meshes, uint64_t W1, char *p,
                                 LevelPtrs *lps) {
  GroupScratch gs;
  gs.pn2 = carve<uint64_t>(p, W1 + 1);
  gs.b2 = carve<unsigned long long>(p, W1 / 64 + 1);
  gs.a_mask2 = carve<unsigned long long>(p, W1 + 1);
  gs.a_base2 = carve<uint32_t>(p, W1 + 1);
  gs.a_code2 = carve<uint32_t>(p, W1 + 1);
  for (size_t g = 0; g < meshes.size(); g++) mesh_levels(meshes[g], &p, &lps[g]);
  return gs;
}

void enqueue_trees_graph(vxl_ctx *ctx, const GraphGroup &gg, const GroupScratch &gs,
                         TraverseSeg *segs, hipStream_t s) {
  const uint32_t G = gg.G;
  const uint64_t W1 = gg.W1;
  unsigned long long *b1 = ctx->pool_b1.as<unsigned long long>();
  const uint32_t cus = uint32_t(ctx->cu_count > 0 ? ctx->cu_count : 256);
  // 1. level 2 over the whole group, the per-mesh upper levels, the per-mesh sizes
  hipLaunchKernelGGL(k_level2, dim3(div_up(W1, kTB)), dim3(kTB), 0, s, b1, W1, gs.pn2, gs.b2);
  scan_exclusive_u64(gs.pn2, gs.pn2, W1, gs.pn2 + W1, ctx->scan_tmp, s);
  hipLaunchKernelGGL(k_small_levels_batch, dim3(G), dim3(kSmallThreads), 0, s, gg.d_mesh,
                     gg.d_lps, gs.b2);
  hipLaunchKernelGGL(k_batch_totals, dim3(div_up(G, 64)), dim3(64), 0, s, gg.d_mesh, gg.d_lps, G,
                     gs.pn2, ctx->coff.as<uint64_t>(), gg.d_ref, gg.d_tot, gg.n2_beg);
  // 2. placement in the arena region, decided on the device
  hipLaunchKernelGGL(k_graph_place, dim3(1), dim3(64), 0, s, G, gg.d_mesh, gg.d_tot,
                     gg.node_region, gg.node_cap, gg.leaves_base, gg.bounds_base, gg.codes_base,
                     gg.leaf_cap, gg.nodes_ptr, gg.node_off, segs, gg.state);
  // 3. level-2 School-A arrays, leaf gather (clears the pooled grid and B1), placement
  hipLaunchKernelGGL(k_emit_a2, dim3(div_up(W1, kTB)), dim3(kTB), 0, s, b1, W1, gs.pn2,
                     gs.a_mask2, gs.a_base2, gs.a_code2);
  hipLaunchKernelGGL(k_leaf_gather_graph, dim3(cus * 8), dim3(kTB), 0, s, gg.n2_beg, G,
                     gg.state, gs.a_mask2, gs.a_base2, gs.a_code2,
                     ctx->pool_grid.as<unsigned long long>(), b1,
                     reinterpret_cast<unsigned long long *>(gg.leaves_base),
                     reinterpret_cast<uint32_t *>(gg.bounds_base),
                     reinterpret_cast<unsigned long long *>(gg.codes_base), gg.d_mesh);
  hipLaunchKernelGGL(k_upper_place_batch, dim3(G), dim3(kSmallThreads), 0, s, gg.d_mesh,
                     gg.d_lps, gg.d_tot, gg.nodes_ptr);
  hipLaunchKernelGGL(k_place2_graph, dim3(cus * 4), dim3(kTB), 0, s, G, gg.n2_beg, gg.d_mesh,
                     gg.d_lps, gg.d_tot, gg.nodes_ptr, gs.a_mask2, gs.a_base2, gs.a_code2);
  VXL_HIP(hipGetLastError());
}

vxl_tree *build_tree_from_grid(vxl_ctx *ctx, uint32_t n, uint32_t k, const uint64_t *pairs_dev,
                               uint64_t *pairs) {
  if (k <= 4) return build_tree_fast(ctx, n, k, pairs_dev, pairs);
  if (pairs) {
    *pairs = 0;
    if (pairs_dev) {
      VXL_HIP(hipMemcpyAsync(ctx->host_pinned, pairs_dev, 8, hipMemcpyDeviceToHost, ctx->stream));
      wait_stream(ctx, ctx->stream);
      *pairs = *static_cast<uint64_t *>(ctx->host_pinned);
    }
  }
  hipStream_t s = ctx->stream;
  const uint32_t top = k + 1 > 2 ? k + 1 : 2;  // node levels 2..top (level 2 always built)
  Level lv[16];
  // B1 is ctx->bitmap (zero between calls); B2, B3, ... are carved from levelscratch.
  uint64_t w = bitmap_level1_words(n);
  size_t scratch = 0;
  for (uint32_t L = 2; L <= top; L++) {
    lv[L].win = w;
    w = div_up(w, 64);
    const uint64_t e = lv[L].win + 1;
    scratch += 10 * (((e * 8) + 255) / 256) * 256 + ((w * 8 + 255) / 256) * 256;
  }
  ctx->levelscratch.reserve(scratch + 4096);
  char *p = static_cast<char *>(ctx->levelscratch.ptr);
  lv[2].bin = ctx->bitmap.as<unsigned long long>();
  for (uint32_t L = 2; L <= top; L++) {
    if (L < top) {
      lv[L].bout = carve<unsigned long long>(p, lv[L + 1].win);
      lv[L + 1].bin = lv[L].bout;
    }
  }
  for (uint32_t L = 2; L <= top; L++) {
    const uint64_t e = lv[L].win + 1;
    lv[L].cb = carve<uint32_t>(p, e);
    lv[L].ni = carve<uint32_t>(p, e);
    lv[L].a_mask = carve<unsigned long long>(p, e);
    lv[L].a_base = carve<uint32_t>(p, e);
    lv[L].a_code = carve<uint32_t>(p, e);
    lv[L].desc = carve<uint32_t>(p, e);
    lv[L].pos = carve<uint32_t>(p, e);
    lv[L].bs = carve<uint32_t>(p, e);
  }
  // 1. pyramid + scans
  for (uint32_t L = 2; L <= top; L++) {
    Level &l = lv[L];
    hipLaunchKernelGGL(k_level, dim3(div_up(l.win, kTB)), dim3(kTB), 0, s, l.bin, l.win, l.cb,
                       l.ni, l.bout);
    scan_exclusive_u32(l.cb, l.cb, l.win, l.cb + l.win, ctx->scan_tmp, s);
    scan_exclusive_u32(l.ni, l.ni, l.win, l.ni + l.win, ctx->scan_tmp, s);
  }
  // 2. totals to the host (one sync)
  uint32_t *host = static_cast<uint32_t *>(ctx->host_pinned);
  for (uint32_t L = 2; L <= top; L++) {
    VXL_HIP(hipMemcpyAsync(host + 2 * L, lv[L].cb + lv[L].win, 4, hipMemcpyDeviceToHost, s));
    VXL_HIP(hipMemcpyAsync(host + 2 * L + 1, lv[L].ni + lv[L].win, 4, hipMemcpyDeviceToHost, s));
  }
  wait_stream(ctx, s);
  for (uint32_t L = 2; L <= top; L++) {
    lv[L].children = host[2 * L];
    lv[L].nodes = host[2 * L + 1];
  }
  const uint64_t leaves = lv[2].children;
  uint64_t nodes_total = 0;
  if (k >= 1)
    for (uint32_t L = 2; L <= k + 1; L++) nodes_total += lv[L].nodes;

  vxl_tree *t = new vxl_tree();
  t->ctx = ctx;
  t->n = n;
  t->k = k;
  t->leaf_count = leaves;
  t->node_count = leaves ? nodes_total : 0;
  if (k >= 1 && leaves)
    for (uint32_t L = 2; L <= k + 1; L++) t->level_nodes[L] = lv[L].nodes;
  try {
    t->nodes = static_cast<uint32_t *>(dmalloc(t->node_count * 12));
    t->leaves = static_cast<uint64_t *>(dmalloc(leaves * 64));
    t->bounds = static_cast<uint32_t *>(dmalloc(leaves * 4));
    t->codes = static_cast<uint64_t *>(dmalloc(leaves * 8));
  } catch (...) {
    tree_free_device(t);
    delete t;
    throw;
  }
  t->bytes = t->node_count * 12 + leaves * 76;
  if (leaves == 0) return t;

  // 3. School-A node arrays per level
  for (uint32_t L = 2; L <= top; L++) {
    Level &l = lv[L];
    hipLaunchKernelGGL(k_emit_a, dim3(div_up(l.win, kTB)), dim3(kTB), 0, s, l.bin, l.win, l.cb,
                       l.ni, l.a_mask, l.a_base, l.a_code);
  }
  // 4. leaf gather (clears the leaf grid and B1)
  {
    const uint64_t n2 = lv[2].nodes;
    hipLaunchKernelGGL(k_leaf_gather, dim3(div_up(n2 * 64, kTB)), dim3(kTB), 0, s, n2,
                       lv[2].a_mask, lv[2].a_base, lv[2].a_code,
                       ctx->leafgrid.as<unsigned long long>(), lv[2].bin,
                       reinterpret_cast<unsigned long long *>(t->leaves), t->bounds,
                       reinterpret_cast<unsigned long long *>(t->codes), nullptr, 0u, nullptr);
  }
  if (k >= 1) {
    // 5. subtree sizes bottom-up: desc, then E = exclusive scan of desc (in place)
    for (uint32_t L = 2; L <= k + 1; L++) {
      Level &l = lv[L];
      const uint32_t *e_child = L > 2 ? lv[L - 1].desc : nullptr;
      hipLaunchKernelGGL(k_desc, dim3(div_up(l.nodes, kTB)), dim3(kTB), 0, s, l.nodes, l.a_mask,
                         l.a_base, e_child, l.desc);
      if (L == 2) {
        // level-2 nodes have no node descendants: desc = 0
        VXL_HIP(hipMemsetAsync(l.desc, 0, (l.nodes + 1) * 4, s));
      } else {
        scan_exclusive_u32(l.desc, l.desc, l.nodes, l.desc + l.nodes, ctx->scan_tmp, s);
      }
    }
    // 6. root, then top-down placement
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, s, lv[k + 1].a_mask, lv[k + 1].a_base,
                       k + 1 == 2 ? 1 : 0, t->nodes, lv[k + 1].pos, lv[k + 1].bs);
    for (uint32_t L = k; L >= 2; L--) {
      Level &l = lv[L];
      Level &pl = lv[L + 1];
      hipLaunchKernelGGL(k_place, dim3(div_up(l.nodes, kTB)), dim3(kTB), 0, s, l.nodes,
                         L == 2 ? 1 : 0, l.a_mask, l.a_base, l.a_code, L > 2 ? l.desc : nullptr,
                         pl.ni, pl.a_mask, pl.a_base, pl.bs, l.pos, l.bs, t->nodes);
    }
  }
  VXL_HIP(hipGetLastError());
  return t;
}

// ---- incremental edits (SparseTree::set_voxel, voxel-core/src/sparse.rs:384-436) ---------
namespace {

// LeafBounds of one leaf from its 8 words (leaf.rs:185-213), by one thread: the same marks
// and counts k_leaf_gather forms across 8 lanes.
__device__ uint32_t leaf_bounds_of(const unsigned long long *__restrict__ words) {
  uint32_t cnt = 0, occ = 0;
#pragma unroll
  for (int w = 0; w < 8; w++) {
    const unsigned long long v = words[w];
    cnt += __popcll(v);
#pragma unroll
    for (int a = 0; a < 3; a++) occ |= (word_coords(v, a) << ((w >> a) & 1) * 4) << (8 * a);
  }
  uint32_t mn[3], mx[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const uint32_t m = (occ >> (8 * a)) & 0xffu;
    mn[a] = m ? uint32_t(__ffs(m) - 1) : 7u;
    mx[a] = m ? 31u - uint32_t(__clz(m)) : 0u;
  }
  return (cnt > 64 || cnt == 0) ? kFull
                                : (mn[0] | (mn[1] << 3) | (mn[2] << 6) | (mx[0] << 9) |
                                   (mx[1] << 12) | (mx[2] << 15));
}

// Leaf index of brick `code` among the L ascending codes, or L when absent.
__device__ __forceinline__ uint64_t find_leaf(const unsigned long long *__restrict__ codes,
                                              uint64_t L, uint64_t code) {
  uint64_t lo = 0, hi = L;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (codes[mid] < code) lo = mid + 1;
    else hi = mid;
  }
  return (lo < L && codes[lo] == code) ? lo : L;
}

struct EditFlags {
  unsigned long long changed;  // voxels whose state changed
  unsigned int topology;       // a brick appears or empties
  unsigned int pad;
};

// One thread per (distinct, in-bounds) edit: an edit in an existing brick flips its bit in
// place (a newly set voxel's material slot reads 0 again, sparse.rs:405-411); a set in a
// missing brick is left for the rebuild (pending) and marks a topology change.
__global__ void k_edit_patch(const vxl_voxel_edit *__restrict__ edits, uint64_t E,
                             unsigned long long *__restrict__ leaves,
                             const unsigned long long *__restrict__ codes, uint64_t L,
                             uint16_t *__restrict__ mats, uint32_t *__restrict__ touched,
                             EditFlags *__restrict__ flags) {
  const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool changed = false;
  if (e < E) {
    const vxl_voxel_edit ed = edits[e];
    const uint64_t code = dev::brick_code(ed.x >> 3, ed.y >> 3, ed.z >> 3);
    const uint32_t m = dev::morton9(ed.x, ed.y, ed.z);
    const uint64_t leaf = find_leaf(codes, L, code);
    uint32_t t = 0xFFFFFFFFu;
    if (leaf < L) {
      unsigned long long *w = leaves + leaf * 8 + (m >> 6);
      const unsigned long long bit = 1ull << (m & 63);
      if (ed.occupied) {
        changed = !(atomicOr(w, bit) & bit);
        if (changed && mats) mats[leaf * 512 + m] = 0;
      } else {
        changed = (atomicAnd(w, ~bit) & bit) != 0;
      }
      if (changed) t = uint32_t(leaf);
    } else if (ed.occupied) {
      changed = true;
      t = 0xFFFFFFFEu;  // pending: a new brick
      atomicOr(&flags->topology, 1u);
    }
    touched[e] = t;
  }
  const unsigned long long cm = __ballot(changed);
  if ((threadIdx.x & 63) == 0 && cm) atomicAdd(&flags->changed, (unsigned long long)__popcll(cm));
}

// Bounds of every patched leaf (after all patches); an emptied leaf is a topology change.
__global__ void k_edit_bounds(const uint32_t *__restrict__ touched, uint64_t E,
                              const unsigned long long *__restrict__ leaves,
                              uint32_t *__restrict__ bounds, EditFlags *__restrict__ flags) {
  const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const uint32_t t = touched[e];
  if (t >= 0xFFFFFFFEu) return;
  const unsigned long long *w = leaves + uint64_t(t) * 8;
  bool empty = true;
#pragma unroll
  for (int i = 0; i < 8; i++) empty = empty && w[i] == 0;
  if (empty) atomicOr(&flags->topology, 1u);
  bounds[t] = leaf_bounds_of(w);
}

// The tree's (patched) leaves back into the dense leaf grid; a non-empty leaf sets its
// brick's bitmap bit.
__global__ void k_leaves_to_grid(const unsigned long long *__restrict__ leaves,
                                 const unsigned long long *__restrict__ codes, uint64_t L,
                                 unsigned long long *__restrict__ leafgrid,
                                 unsigned long long *__restrict__ b1) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const uint64_t code = codes[i];
  unsigned long long any = 0;
#pragma unroll
  for (int w = 0; w < 8; w++) {
    const unsigned long long v = leaves[i * 8 + w];
    leafgrid[code * 8 + w] = v;
    any |= v;
  }
  if (any) atomicOr(&b1[code >> 6], 1ull << (code & 63));
}

__global__ void k_edit_new_bricks(const vxl_voxel_edit *__restrict__ edits,
                                  const uint32_t *__restrict__ touched, uint64_t E,
                                  unsigned long long *__restrict__ leafgrid,
                                  unsigned long long *__restrict__ b1) {
  const uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= E || touched[e] != 0xFFFFFFFEu) return;
  const vxl_voxel_edit ed = edits[e];
  const uint64_t code = dev::brick_code(ed.x >> 3, ed.y >> 3, ed.z >> 3);
  const uint32_t m = dev::morton9(ed.x, ed.y, ed.z);
  atomicOr(&leafgrid[code * 8 + (m >> 6)], 1ull << (m & 63));
  atomicOr(&b1[code >> 6], 1ull << (code & 63));
}

// Materials of the rebuilt tree: a surviving brick keeps its 512 slots, a new one reads 0.
__global__ void k_remap_materials(const unsigned long long *__restrict__ new_codes, uint64_t L2,
                                  const unsigned long long *__restrict__ old_codes, uint64_t L1,
                                  const uint16_t *__restrict__ old_mats,
                                  uint16_t *__restrict__ new_mats) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < L2 * 512;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t leaf = i >> 9;
    const uint64_t j = find_leaf(old_codes, L1, new_codes[leaf]);
    new_mats[i] = j < L1 ? old_mats[j * 512 + (i & 511)] : uint16_t(0);
  }
}

}  // namespace

EditResult edit_tree(vxl_ctx *ctx, vxl_tree *t, const vxl_voxel_edit *d_edits, uint64_t E) {
  EditResult res{0, false};
  if (E == 0) return res;
  hipStream_t s = ctx->stream;
  ctx->edit_scratch.reserve(E * 4 + 256);
  uint32_t *touched = reinterpret_cast<uint32_t *>(ctx->edit_scratch.as<char>() + 256);
  // flags in the mapped readback area (past what the tree build uses of it)
  constexpr size_t kFlagSlot = 3328;
  EditFlags *dflags =
      reinterpret_cast<EditFlags *>(static_cast<char *>(ctx->host_pinned_dev) + kFlagSlot);
  volatile EditFlags *hflags =
      reinterpret_cast<volatile EditFlags *>(static_cast<char *>(ctx->host_pinned) + kFlagSlot);
  hflags->changed = 0;
  hflags->topology = 0;
  const uint32_t blocks = uint32_t(div_up(E, kTB));
  hipLaunchKernelGGL(k_edit_patch, dim3(blocks), dim3(kTB), 0, s, d_edits, E,
                     reinterpret_cast<unsigned long long *>(t->leaves),
                     reinterpret_cast<const unsigned long long *>(t->codes), t->leaf_count,
                     t->materials, touched, dflags);
  hipLaunchKernelGGL(k_edit_bounds, dim3(blocks), dim3(kTB), 0, s, touched, E,
                     reinterpret_cast<const unsigned long long *>(t->leaves), t->bounds, dflags);
  VXL_HIP(hipGetLastError());
  wait_stream(ctx, s);
  res.changed = hflags->changed;
  res.topology = hflags->topology != 0;
  if (!res.topology) return res;  // every change was in place (Edit::Leaf)
  // a brick appears or empties (Edit::Topology): rebuild from the patched leaves plus the
  // new bricks' voxels, as SparseTree rebuilds its levels from the spliced codes
  prepare_grid(ctx, t->n);
  unsigned long long *grid = ctx->leafgrid.as<unsigned long long>();
  unsigned long long *b1 = ctx->bitmap.as<unsigned long long>();
  if (t->leaf_count)
    hipLaunchKernelGGL(k_leaves_to_grid, dim3(uint32_t(div_up(t->leaf_count, kTB))), dim3(kTB), 0,
                       s, reinterpret_cast<const unsigned long long *>(t->leaves),
                       reinterpret_cast<const unsigned long long *>(t->codes), t->leaf_count,
                       grid, b1);
  hipLaunchKernelGGL(k_edit_new_bricks, dim3(blocks), dim3(kTB), 0, s, d_edits, touched, E, grid,
                     b1);
  VXL_HIP(hipGetLastError());
  vxl_tree *nt = build_tree_from_grid(ctx, t->n, t->k, nullptr, nullptr);
  if (t->materials && nt->leaf_count) {
    try {
      void *p = nullptr;
      if (hipMalloc(&p, size_t(nt->leaf_count) * 1024) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(VXL_OOM, "edit: materials allocation failed");
      }
      nt->materials = static_cast<uint16_t *>(p);
      const uint64_t rb = div_up(nt->leaf_count * 512, kTB);
      hipLaunchKernelGGL(k_remap_materials, dim3(uint32_t(rb < (1ull << 22) ? rb : (1ull << 22))),
                         dim3(kTB), 0, s, reinterpret_cast<const unsigned long long *>(nt->codes),
                         nt->leaf_count, reinterpret_cast<const unsigned long long *>(t->codes),
                         t->leaf_count, t->materials, nt->materials);
      VXL_HIP(hipGetLastError());
    } catch (...) {
      tree_free_device(nt);
      delete nt;
      throw;
    }
  }
  // the caller's handle takes the new arrays; the old ones are released with nt
  std::swap(t->node_count, nt->node_count);
  std::swap(t->leaf_count, nt->leaf_count);
  std::swap(t->nodes, nt->nodes);
  std::swap(t->leaves, nt->leaves);
  std::swap(t->bounds, nt->bounds);
  std::swap(t->codes, nt->codes);
  std::swap(t->materials, nt->materials);
  std::swap(t->level_nodes, nt->level_nodes);
  std::swap(t->bytes, nt->bytes);
  std::swap(t->slab, nt->slab);
  std::swap(t->shared, nt->shared);
  VXL_HIP(hipStreamSynchronize(s));  // the remap read the old arrays
  tree_free_device(nt);
  delete nt;
  return res;
}

}  // namespace vxl
