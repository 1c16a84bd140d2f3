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

__device__ __forceinline__ uint32_t select_bit(unsigned long long m, uint32_t s) {
  // position of the s-th (0-based) set bit of m
  uint32_t pos = 0;
#pragma unroll
  for (int width = 32; width >= 1; width >>= 1) {
    unsigned long long low = m & ((1ull << width) - 1ull);
    uint32_t c = __popcll(low);
    if (s >= c) {
      s -= c;
      m >>= width;
      pos += width;
    } else {
      m = low;
    }
  }
  return pos;
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
__global__ __launch_bounds__(kTB) void k_leaf_gather(
    uint64_t n2, const unsigned long long *__restrict__ a_mask, const uint32_t *__restrict__ a_base,
    const uint32_t *__restrict__ a_code, unsigned long long *__restrict__ leafgrid,
    unsigned long long *__restrict__ brickbits, unsigned long long *__restrict__ leaves,
    uint32_t *__restrict__ bounds, unsigned long long *__restrict__ codes) {
  const uint64_t node = (uint64_t(blockIdx.x) * kTB + threadIdx.x) >> 6;
  if (node >= n2) return;
  const int lane = threadIdx.x & 63;
  const unsigned long long mask = a_mask[node];
  const uint32_t base = a_base[node];
  const uint32_t j = a_code[node];
  const uint32_t nch = __popcll(mask);
  const int sub = lane >> 3, w = lane & 7;
  for (uint32_t g = 0; g < nch; g += 8) {
    const uint32_t s = g + sub;
    const bool act = s < nch;
    const uint32_t bit = act ? select_bit(mask, s) : 0;
    const uint64_t code = (uint64_t(j) << 6) | bit;
    unsigned long long v = 0;
    if (act) {
      v = leafgrid[code * 8 + w];
      leafgrid[code * 8 + w] = 0ull;
      leaves[(uint64_t(base) + s) * 8 + w] = v;
    }
    // LeafBounds over the 8 words of this leaf (8-lane groups)
    uint32_t cnt = __popcll(v);
    uint32_t mn[3], mx[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      word_extent(v, a, mn[a], mx[a]);
      const uint32_t off = ((w >> a) & 1) * 4;
      if (v) {
        mn[a] += off;
        mx[a] += off;
      } else {
        mn[a] = 7;
        mx[a] = 0;
      }
    }
#pragma unroll
    for (int d = 1; d < 8; d <<= 1) {
      cnt += __shfl_xor(cnt, d, 64);
#pragma unroll
      for (int a = 0; a < 3; a++) {
        mn[a] = min(mn[a], (uint32_t)__shfl_xor((int)mn[a], d, 64));
        mx[a] = max(mx[a], (uint32_t)__shfl_xor((int)mx[a], d, 64));
      }
    }
    if (act && w == 0) {
      uint32_t b = (cnt > 64 || cnt == 0)
                       ? kFull
                       : (mn[0] | (mn[1] << 3) | (mn[2] << 6) | (mx[0] << 9) | (mx[1] << 12) |
                          (mx[2] << 15));
      bounds[base + s] = b;
      codes[base + s] = code;
    }
  }
  if (lane == 0) brickbits[j] = 0ull;  // B1 back to all-zero
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

void tree_free_device(vxl_tree *t) {
  if (t->nodes) (void)hipFree(t->nodes);
  if (t->leaves) (void)hipFree(t->leaves);
  if (t->bounds) (void)hipFree(t->bounds);
  if (t->codes) (void)hipFree(t->codes);
  t->nodes = nullptr;
  t->leaves = nullptr;
  t->bounds = nullptr;
  t->codes = nullptr;
}

vxl_tree *build_tree_from_grid(vxl_ctx *ctx, uint32_t n, uint32_t k) {
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
  VXL_HIP(hipStreamSynchronize(s));
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
                       reinterpret_cast<unsigned long long *>(t->codes));
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

}  // namespace vxl
