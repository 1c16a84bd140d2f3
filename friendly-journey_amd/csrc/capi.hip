// extern "C" entry points of libvxl.so (declared in include/vxl.h).
#include <algorithm>
#include <cmath>
#include <thread>
#include <unordered_map>
#include <vector>
#include <cstdio>
#include <cstring>

#include "vxl_internal.hpp"

namespace vxl {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

void DevBuf::reserve(size_t want) {
  if (want <= bytes && ptr) return;
  // kernels enqueued earlier may still read the old buffer: let them finish first
  if (ptr) (void)hipDeviceSynchronize();
  release();
  size_t b = want < 256 ? 256 : want;
  hipError_t e = hipMalloc(&ptr, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    bytes = 0;
    throw Error(VXL_OOM, std::string("device allocation of ") + std::to_string(b) +
                             " bytes failed: " + hipGetErrorString(e));
  }
  bytes = b;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

void wait_stream(vxl_ctx *ctx, hipStream_t s) {
  if (!ctx->ev_wait) VXL_HIP(hipEventCreateWithFlags(&ctx->ev_wait, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ctx->ev_wait, s));
  for (;;) {
    const hipError_t e = hipEventQuery(ctx->ev_wait);
    if (e == hipSuccess) return;
    if (e != hipErrorNotReady) {
      (void)hipGetLastError();
      throw Error(VXL_DEVICE_ERROR, std::string("stream wait failed: ") + hipGetErrorString(e));
    }
  }
}

bool resolution_k(uint32_t n, uint32_t *k) {
  // voxel-core/src/resolution.rs:76-99: n = 8 * 4^k
  if (n < 8 || n % 8) return false;
  uint32_t m = n / 8;
  if (m & (m - 1)) return false;
  uint32_t tz = __builtin_ctz(m);
  if (tz % 2) return false;
  *k = tz / 2;
  return true;
}

}  // namespace vxl

using vxl::Error;

namespace {

constexpr uint32_t kMaxK = 5;  // the traversal's frame stack holds 6 levels (mirror.rs:22)
// Rays one allocation-free traversal takes (its limit; the work counter is 64-bit).
constexpr uint64_t kNoAllocMaxRays = 1ull << 40;

template <class F>
vxl_status guarded(vxl_ctx *ctx, F &&f) {
  if (!ctx) {
    vxl::set_error("null context");
    return VXL_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    (void)hipGetLastError();
    vxl::set_error("hipSetDevice failed");
    return VXL_DEVICE_ERROR;
  }
  try {
    f();
    return VXL_OK;
  } catch (const Error &e) {
    vxl::set_error(e.msg);
    ctx->leafgrid_dirty = true;  // a failed call may leave scratch non-zero
    ctx->pool_dirty = true;
    return e.code;
  } catch (const std::exception &e) {
    vxl::set_error(e.what());
    ctx->leafgrid_dirty = true;
    ctx->pool_dirty = true;
    return VXL_DEVICE_ERROR;
  }
}

// VoxelGrid::validate (voxelizer/src/core.rs:116-132) + Resolution::new (resolution.rs:76).
void validate_grid(const vxl_grid *g, uint32_t *k) {
  if (!g) throw Error(VXL_INVALID_ARG, "grid is null");
  if (!vxl::resolution_k(g->n, k)) {
    throw Error(VXL_NOT_REPRESENTABLE,
                "resolution " + std::to_string(g->n) + " is not representable (must be 8*4^k)");
  }
  if (*k > kMaxK)
    throw Error(VXL_STORAGE_EXCEEDED,
                "resolution " + std::to_string(g->n) + " exceeds the supported 8*4^5 = 8192");
  if (!std::isfinite(g->voxel_size) || g->voxel_size <= 0.0f)
    throw Error(VXL_INVALID_ARG, "voxel_size must be finite and > 0");
  if (!g->has_world_to_grid && !std::isfinite(1.0f / g->voxel_size))
    throw Error(VXL_INVALID_ARG, "voxel_size is too small: its reciprocal is not finite");
  for (int a = 0; a < 3; a++)
    if (!std::isfinite(g->origin[a])) throw Error(VXL_INVALID_ARG, "origin_world must be finite");
  if (g->has_world_to_grid)
    for (int i = 0; i < 16; i++)
      if (!std::isfinite(g->world_to_grid[i]))
        throw Error(VXL_INVALID_ARG, "world_to_grid must be finite");
}

// Whether any of n binary32 values is not finite (all 8 exponent bits set): branch-free,
// so the compiler vectorises it.
bool any_non_finite(const uint32_t *src, uint64_t n) {
  uint32_t bad = 0;
  for (uint64_t i = 0; i < n; i++) bad |= uint32_t((src[i] & 0x7f800000u) == 0x7f800000u);
  return bad != 0;
}

// The same check while copying the values from src to dst.
bool copy_check_finite(const uint32_t *src, uint32_t *dst, uint64_t n) {
  uint32_t bad = 0;
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t v = src[i];
    dst[i] = v;
    bad |= uint32_t((v & 0x7f800000u) == 0x7f800000u);
  }
  return bad != 0;
}

// The two pinned chunk buffers host transfers go through (allocated on first use).
constexpr size_t kStageBytes = size_t(64) << 20;
void stage_buffers(vxl_ctx *ctx) {
  for (int b = 0; b < 2; b++) {
    if (!ctx->up_host[b]) VXL_HIP(hipHostMalloc(&ctx->up_host[b], kStageBytes, hipHostMallocDefault));
    if (!ctx->up_ev[b]) VXL_HIP(hipEventCreateWithFlags(&ctx->up_ev[b], hipEventDisableTiming));
  }
}

unsigned stage_threads() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, std::min(8u, hw ? hw : 1u));
}

// fn(t, lo, hi) over [0, len) in nth slices: slice 0 and any slice whose helper thread
// cannot be started run on the calling thread, the others on helper threads; returns once
// every slice is done. fn must not throw.
template <class F>
void run_slices(size_t len, unsigned nth, const F &fn) {
  const size_t slice = (len + nth - 1) / nth;
  std::vector<std::thread> pool;
  std::vector<unsigned> here;
  pool.reserve(nth);  // no allocation once a thread runs
  here.reserve(nth);
  here.push_back(0u);
  for (unsigned t = 1; t < nth; t++) {
    if (std::min(len, t * slice) >= len) break;
    try {
      pool.emplace_back([&fn, t, slice, len] {
        const size_t lo = t * slice;
        fn(t, lo, std::min(len, lo + slice));
      });
    } catch (...) {
      here.push_back(t);
    }
  }
  for (unsigned t : here) {
    const size_t lo = std::min(len, t * slice);
    fn(t, lo, std::min(len, lo + slice));
  }
  for (auto &th : pool) th.join();
}

// Device -> host copy of `bytes` from src (device) into dst (the caller's host memory), after
// the work already enqueued on ctx->stream; complete on return. Large copies go in 64 MiB
// chunks through the two pinned buffers: a chunk's DMA runs while up to 8 threads copy the
// previous chunk out of the other buffer into dst (pageable copies of 1 GiB ran ~8 GB/s).
void download_host(vxl_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  if (bytes < (size_t(4) << 20)) {
    VXL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VXL_HIP(hipStreamSynchronize(ctx->stream));
    return;
  }
  stage_buffers(ctx);
  const unsigned nth = stage_threads();
  char *out = static_cast<char *>(dst);
  const char *in = static_cast<const char *>(src);
  auto copy_out = [&](int b, size_t off, size_t len) {
    VXL_HIP(hipEventSynchronize(ctx->up_ev[b]));
    ctx->up_pending[b] = false;
    const char *stage = static_cast<const char *>(ctx->up_host[b]);
    run_slices(len, nth, [=](unsigned, size_t lo, size_t hi) {
      std::memcpy(out + off + lo, stage + lo, hi - lo);
    });
  };
  int prev = -1;
  size_t prev_off = 0, prev_len = 0;
  for (size_t off = 0, c = 0; off < bytes; off += kStageBytes, c++) {
    const size_t len = std::min(kStageBytes, bytes - off);
    const int b = int(c & 1);
    if (ctx->up_pending[b]) {  // an earlier transfer's DMA from this buffer
      VXL_HIP(hipEventSynchronize(ctx->up_ev[b]));
      ctx->up_pending[b] = false;
    }
    VXL_HIP(hipMemcpyAsync(ctx->up_host[b], in + off, len, hipMemcpyDeviceToHost, ctx->stream));
    VXL_HIP(hipEventRecord(ctx->up_ev[b], ctx->stream));
    ctx->up_pending[b] = true;
    if (prev >= 0) copy_out(prev, prev_off, prev_len);
    prev = b;
    prev_off = off;
    prev_len = len;
  }
  copy_out(prev, prev_off, prev_len);
}

// Host -> device copy of `bytes` from src (the caller's host memory) to dst (device),
// enqueued on ctx->stream; src may be reused on return. Large copies go in 64 MiB chunks:
// up to 8 threads copy a chunk into a pinned buffer, whose DMA then overlaps the copy of
// the next chunk into the other buffer.
void upload_host(vxl_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!bytes) return;
  if (bytes < (size_t(4) << 20)) {
    VXL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return;
  }
  stage_buffers(ctx);
  const unsigned nth = stage_threads();
  char *out = static_cast<char *>(dst);
  const char *in = static_cast<const char *>(src);
  for (size_t off = 0, c = 0; off < bytes; off += kStageBytes, c++) {
    const size_t len = std::min(kStageBytes, bytes - off);
    const int b = int(c & 1);
    if (ctx->up_pending[b]) {
      VXL_HIP(hipEventSynchronize(ctx->up_ev[b]));
      ctx->up_pending[b] = false;
    }
    char *stage = static_cast<char *>(ctx->up_host[b]);
    run_slices(len, nth, [=](unsigned, size_t lo, size_t hi) {
      std::memcpy(stage + lo, in + off + lo, hi - lo);
    });
    VXL_HIP(hipMemcpyAsync(out + off, stage, len, hipMemcpyHostToDevice, ctx->stream));
    VXL_HIP(hipEventRecord(ctx->up_ev[b], ctx->stream));
    ctx->up_pending[b] = true;
  }
}

// MeshInput::validate (core.rs:427-459: every vertex finite) fused with the upload of the
// host triangles into ctx->tris_in on ctx->stream. Large meshes go in chunks of 16M values
// through two pinned buffers: up to 8 threads copy and check a chunk, then its DMA runs
// while the next chunk is copied into the other buffer (a buffer is reused once its last
// DMA has completed). A non-finite value fails the call with VXL_INVALID_ARG before any
// kernel runs, as the reference validates the mesh before any device work.
void upload_mesh_host(vxl_ctx *ctx, const float *tris, uint64_t T) {
  if (T && !tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
  if (T == 0) return;
  const uint64_t nf = T * 9;
  ctx->tris_in.reserve(nf * 4);
  uint32_t *dst = ctx->tris_in.as<uint32_t>();
  const uint32_t *src = reinterpret_cast<const uint32_t *>(tris);
  const Error bad_vertex(VXL_INVALID_ARG, "triangle contains a non-finite vertex");
  constexpr uint64_t kSmall = 1u << 18;  // values: below this, one thread and a direct copy
  if (nf < kSmall) {
    if (any_non_finite(src, nf)) throw bad_vertex;
    VXL_HIP(hipMemcpyAsync(dst, src, nf * 4, hipMemcpyHostToDevice, ctx->stream));
    return;
  }
  constexpr uint64_t kChunk = kStageBytes / 4;  // values per chunk
  stage_buffers(ctx);
  const unsigned nth = stage_threads();
  for (uint64_t off = 0, c = 0; off < nf; off += kChunk, c++) {
    const uint64_t len = std::min(kChunk, nf - off);
    const int b = int(c & 1);
    if (ctx->up_pending[b]) {
      VXL_HIP(hipEventSynchronize(ctx->up_ev[b]));
      ctx->up_pending[b] = false;
    }
    uint32_t *stage = static_cast<uint32_t *>(ctx->up_host[b]);
    std::vector<char> bad(nth, 0);
    char *badp = bad.data();
    run_slices(len, nth, [=](unsigned t, size_t lo, size_t hi) {
      badp[t] = copy_check_finite(src + off + lo, stage + lo, hi - lo);
    });
    for (char x : bad)
      if (x) throw bad_vertex;
    VXL_HIP(hipMemcpyAsync(dst + off, stage, len * 4, hipMemcpyHostToDevice, ctx->stream));
    VXL_HIP(hipEventRecord(ctx->up_ev[b], ctx->stream));
    ctx->up_pending[b] = true;
  }
}


void validate_count(uint64_t T) {
  if (T > 0xffffffffull) throw Error(VXL_INVALID_ARG, "triangle count must be < 2^32");
}

// VoxelizeOpts::validate (core.rs:492-500)
void validate_eps(float eps) {
  if (!std::isfinite(eps) || eps < 0.0f)
    throw Error(VXL_INVALID_ARG, "voxelize epsilon must be finite and >= 0");
}

vxl::VoxelizeParams make_params(const vxl_grid *g, float eps) {
  vxl::VoxelizeParams p;
  p.n = g->n;
  // world_to_grid_matrix (core.rs:101-107): scale(1/vs) * translate(-origin);
  // per component g = (inv * p) + (inv * -o), evaluated in binary32.
  volatile float inv = 1.0f / g->voxel_size;
  p.inv = inv;
  for (int a = 0; a < 3; a++) {
    volatile float t = p.inv * (-g->origin[a]);
    p.tx[a] = t;
  }
  p.eps = eps;
  p.has_mat = g->has_world_to_grid ? 1u : 0u;
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 3; r++) p.m[3 * c + r] = p.has_mat ? g->world_to_grid[4 * c + r] : 0.0f;
  return p;
}


// The pair count of the last SAT launch (one host round trip).
uint64_t read_pairs(vxl_ctx *ctx) {
  VXL_HIP(hipMemcpyAsync(ctx->host_pinned, vxl::sat_pairs_dev(ctx), 8, hipMemcpyDeviceToHost,
                         ctx->stream));
  vxl::wait_stream(ctx, ctx->stream);
  return *static_cast<uint64_t *>(ctx->host_pinned);
}

vxl_tree *empty_tree(vxl_ctx *ctx, uint32_t n, uint32_t k) {
  vxl_tree *t = new vxl_tree();
  t->ctx = ctx;
  t->n = n;
  t->k = k;
  // one padded entry each so every pointer is valid (buffers.rs:26-44)
  VXL_HIP(hipMalloc(&t->nodes, 16));
  VXL_HIP(hipMalloc(&t->leaves, 64));
  VXL_HIP(hipMalloc(&t->bounds, 16));
  VXL_HIP(hipMalloc(&t->codes, 16));
  return t;
}

// The device part of voxelize -> tree, on ctx->stream.
vxl_tree *voxelize_tree_impl(vxl_ctx *ctx, const float *d_tris, uint64_t T, const vxl_grid *g,
                             uint32_t k, float eps, uint64_t *pairs_out) {
  if (T == 0) {
    if (pairs_out) *pairs_out = 0;
    return empty_tree(ctx, g->n, k);
  }
  vxl::mark(ctx, 0);
  vxl::prepare_grid(ctx, g->n);
  vxl::run_sat(ctx, d_tris, T, make_params(g, eps));
  uint64_t P = 0;
  vxl_tree *t = vxl::build_tree_from_grid(ctx, g->n, k, vxl::sat_pairs_dev(ctx), &P);
  if (pairs_out) *pairs_out = P;
  vxl::mark(ctx, 3);
  if (ctx->timing && P) {
    VXL_HIP(hipEventSynchronize(ctx->ev[3]));
    VXL_HIP(hipEventElapsedTime(&ctx->last_ms[0], ctx->ev[0], ctx->ev[1]));
    VXL_HIP(hipEventElapsedTime(&ctx->last_ms[1], ctx->ev[1], ctx->ev[2]));
    VXL_HIP(hipEventElapsedTime(&ctx->last_ms[2], ctx->ev[2], ctx->ev[3]));
  }
  return t;
}

// Validated batch voxelize on ctx->stream: empty meshes get empty trees, meshes the
// batched tree build covers (k = 2..4, n = 128..2048) go through grouped launches, the
// rest through the single-mesh path. Every grid and the epsilon are checked before any
// device work.
void voxelize_batch_impl(vxl_ctx *ctx, const float *d_tris, const uint64_t *off, uint32_t count,
                         const vxl_grid *grids, float eps, vxl_tree **out, uint64_t *pairs_out) {
  std::vector<uint32_t> ks(count);
  for (uint32_t g = 0; g < count; g++) {
    validate_grid(&grids[g], &ks[g]);
    if (off[g + 1] < off[g]) throw Error(VXL_INVALID_ARG, "tri_offsets must be non-decreasing");
    validate_count(off[g + 1] - off[g]);
  }
  if (count && off[count] && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
  validate_eps(eps);
  for (uint32_t g = 0; g < count; g++) out[g] = nullptr;
  try {
    std::vector<vxl::BatchInput> batch;
    for (uint32_t g = 0; g < count; g++) {
      const uint64_t T = off[g + 1] - off[g];
      if (T == 0) {
        out[g] = empty_tree(ctx, grids[g].n, ks[g]);
        ctx->trees.insert(out[g]);
        if (pairs_out) pairs_out[g] = 0;
      } else if (ks[g] >= 2 && ks[g] <= 4) {
        batch.push_back(vxl::BatchInput{g, d_tris + 9 * off[g], T, grids[g].n, ks[g],
                                        make_params(&grids[g], eps)});
      } else {
        out[g] = voxelize_tree_impl(ctx, d_tris + 9 * off[g], T, &grids[g], ks[g], eps,
                                    pairs_out ? pairs_out + g : nullptr);
        ctx->trees.insert(out[g]);
      }
    }
    vxl::voxelize_batch(ctx, batch, out, pairs_out);
  } catch (...) {
    for (uint32_t g = 0; g < count; g++)
      if (out[g]) {
        ctx->trees.erase(out[g]);
        vxl::tree_free_device(out[g]);
        delete out[g];
        out[g] = nullptr;
      }
    throw;
  }
}

// Segments of a batched traversal over one ray array delimited by offsets.
std::vector<vxl::RaySegment> segments(const vxl_tree *const *trees, uint32_t count,
                                      const uint64_t *off, const float *rays, uint32_t *hits,
                                      double *t) {
  std::vector<vxl::RaySegment> segs(count);
  for (uint32_t g = 0; g < count; g++) {
    const uint64_t a = off[g];
    segs[g] = vxl::RaySegment{trees[g], rays + 6 * a, off[g + 1] - off[g], hits + 4 * a,
                              t ? t + a : nullptr};
  }
  return segs;
}

void traverse_batch_checks(const vxl_tree *const *trees, uint32_t count, const uint64_t *off) {
  if (count && (!trees || !off)) throw Error(VXL_INVALID_ARG, "null trees or ray offsets");
  for (uint32_t g = 0; g < count; g++) {
    vxl::check_tree(trees[g]);
    if (off[g + 1] < off[g]) throw Error(VXL_INVALID_ARG, "ray_offsets must be non-decreasing");
  }
}

// A caller-supplied School-B structure, checked once on the host before upload: the walk
// indexes nodes and leaves by child_base + popcount rank without bounds checks (the
// reference's WGSL relies on clamped storage reads instead), so every node reachable from
// the root must address children inside the arrays -- nodes for levels >= 3, leaves for
// level-2 nodes (school_b.rs:68-98, node.rs:64-72) -- and no node may be reached at two
// levels. k = 0 trees have no nodes: the root is leaf 0.
void validate_school_b(uint32_t k, const vxl_node *nodes, uint64_t node_count,
                       uint64_t leaf_count) {
  if (k == 0) return;
  std::vector<uint8_t> level(node_count, 0);  // level each node was reached at (0: not yet)
  std::vector<uint64_t> stack{0};
  level[0] = uint8_t(k + 1);
  while (!stack.empty()) {
    const uint64_t i = stack.back();
    stack.pop_back();
    const vxl_node &nd = nodes[i];
    const uint64_t kids = uint64_t(__builtin_popcount(nd.mask_lo)) + __builtin_popcount(nd.mask_hi);
    const uint32_t L = level[i];
    const uint64_t end = uint64_t(nd.child_base) + kids;
    if (L == 2) {
      if (end > leaf_count)
        throw Error(VXL_INVALID_ARG, "node " + std::to_string(i) +
                                         " addresses leaves past leaf_count");
      continue;
    }
    if (end > node_count)
      throw Error(VXL_INVALID_ARG, "node " + std::to_string(i) +
                                       " addresses nodes past node_count");
    for (uint64_t c = nd.child_base; c < end; c++) {
      if (level[c] == 0) {
        level[c] = uint8_t(L - 1);
        stack.push_back(c);
      } else if (level[c] != L - 1) {
        throw Error(VXL_INVALID_ARG, "node " + std::to_string(c) + " is reached at two levels");
      }
    }
  }
}

// Makes `to` wait for the work enqueued on `from` so far (no-op for null or equal streams).
void join_streams(hipStream_t from, hipStream_t to) {
  if (!from || !to || from == to) return;
  hipEvent_t ev;
  VXL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ev, from));
  VXL_HIP(hipStreamWaitEvent(to, ev, 0));
  VXL_HIP(hipEventDestroy(ev));
}

}  // namespace

namespace vxl {
uint32_t check_voxelize_args(const vxl_grid *grid, uint64_t T, float eps) {
  uint32_t k = 0;
  validate_grid(grid, &k);
  validate_count(T);
  validate_eps(eps);
  return k;
}
VoxelizeParams voxelize_params(const vxl_grid *grid, float eps) { return make_params(grid, eps); }
}  // namespace vxl

extern "C" {

const char *vxl_last_error(void) { return vxl::g_last_error.c_str(); }

const char *vxl_status_str(vxl_status s) {
  switch (s) {
    case VXL_OK: return "ok";
    case VXL_NO_DEVICE: return "no device";
    case VXL_INVALID_ARG: return "invalid argument";
    case VXL_STORAGE_EXCEEDED: return "storage exceeded";
    case VXL_WORKGROUPS_EXCEEDED: return "workgroups exceeded";
    case VXL_OOM: return "out of device memory";
    case VXL_BUSY: return "busy";
    case VXL_DEVICE_ERROR: return "device error";
    case VXL_NOT_REPRESENTABLE: return "resolution not representable";
  }
  return "unknown";
}

vxl_status vxl_ctx_create(int device, vxl_ctx **out) {
  if (!out) {
    vxl::set_error("out is null");
    return VXL_INVALID_ARG;
  }
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    (void)hipGetLastError();
    vxl::set_error("no HIP device present");
    return VXL_NO_DEVICE;
  }
  if (device < 0 || device >= count) {
    vxl::set_error("device index out of range");
    return VXL_NO_DEVICE;
  }
  vxl_ctx *ctx = new vxl_ctx();
  ctx->device = device;
  try {
    VXL_HIP(hipSetDevice(device));
    VXL_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // fine-grained (coherent) and mapped: kernels write sizes and flags here that the host
    // polls while they run
    VXL_HIP(hipHostMalloc(&ctx->host_pinned, 4096, hipHostMallocMapped | hipHostMallocCoherent));
    // pinned memory can come back from an earlier owner (another context) with its contents
    std::memset(ctx->host_pinned, 0, 4096);
    VXL_HIP(hipHostGetDevicePointer(&ctx->host_pinned_dev, ctx->host_pinned, 0));
    hipDeviceProp_t prop;
    VXL_HIP(hipGetDeviceProperties(&prop, device));
    ctx->cu_count = prop.multiProcessorCount;
    for (auto &e : ctx->ev) VXL_HIP(hipEventCreate(&e));
    vxl::noalloc_setup(device);  // once per device: the allocation-free path's scratch
  } catch (const Error &e) {
    vxl::set_error(e.msg);
    delete ctx;
    return e.code;
  }
  *out = ctx;
  return VXL_OK;
}

void vxl_ctx_destroy(vxl_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  vxl::DevBuf *bufs[] = {&ctx->tris_in, &ctx->tri_count, &ctx->tri_flag, &ctx->tri_cidx,
                         &ctx->rec,     &ctx->ccount,    &ctx->coff,     &ctx->bstart,
                         &ctx->leafgrid, &ctx->bitmap,   &ctx->levelscratch, &ctx->scan_tmp,
                         &ctx->words_io, &ctx->rays_io,  &ctx->hits_io,  &ctx->t_io,
                         &ctx->pool_grid, &ctx->pool_b1, &ctx->batch_scratch, &ctx->batch_l2,
                         &ctx->batch_desc, &ctx->trav_desc,  &ctx->trav_ctr,
                         &ctx->owner_io, &ctx->color_io, &ctx->owner_scratch,
                         &ctx->leaf_map, &ctx->compact_io, &ctx->edit_io,
                         &ctx->edit_scratch, &ctx->setup_sums};
  for (auto *b : bufs) b->release();
  // Trees outlive nothing they depend on: release their device memory now and detach
  // them; vxl_tree_free on a detached tree only frees the handle.
  for (vxl_tree *t : ctx->trees) {
    vxl::tree_free_device(t);
    t->ctx = nullptr;
    t->leaf_count = t->node_count = 0;
    t->bytes = 0;
  }
  ctx->trees.clear();
  vxl::arena_destroy(ctx);
  for (auto &e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
  if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
  for (int b = 0; b < 2; b++) {
    if (ctx->up_host[b]) (void)hipHostFree(ctx->up_host[b]);
    if (ctx->up_ev[b]) (void)hipEventDestroy(ctx->up_ev[b]);
  }
  if (ctx->staging_free) (void)hipEventDestroy(ctx->staging_free);
  if (ctx->trav_done) (void)hipEventDestroy(ctx->trav_done);
  if (ctx->ev_wait) (void)hipEventDestroy(ctx->ev_wait);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void *vxl_ctx_stream(vxl_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

void vxl_set_timing(vxl_ctx *ctx, int enable) {
  if (ctx) ctx->timing = enable != 0;
}

void vxl_last_timing(vxl_ctx *ctx, float out_ms[8]) {
  if (!ctx || !out_ms) return;
  for (int i = 0; i < 8; i++) out_ms[i] = ctx->last_ms[i];
}

vxl_status vxl_set_tuning(vxl_ctx *ctx, int sat_waves_per_cu, int sat_units_per_wave,
                          int dda_block) {
  return guarded(ctx, [&] {
    if (sat_waves_per_cu) {
      if (sat_waves_per_cu < 1 || sat_waves_per_cu > 64)
        throw Error(VXL_INVALID_ARG, "sat_waves_per_cu must be in [1, 64]");
      ctx->sat_waves_per_cu = sat_waves_per_cu;
    }
    if (sat_units_per_wave) {
      if (sat_units_per_wave < 1 || sat_units_per_wave > 1024)
        throw Error(VXL_INVALID_ARG, "sat_units_per_wave must be in [1, 1024]");
      ctx->sat_units_per_wave = sat_units_per_wave;
    }
    if (dda_block) {
      if (dda_block % 64 || dda_block < 64 || dda_block > 256)
        throw Error(VXL_INVALID_ARG, "dda_block must be a multiple of 64 in [64, 256]");
      ctx->dda_block = dda_block;
    }
  });
}

vxl_status vxl_set_merge_words(vxl_ctx *ctx, int words) {
  return guarded(ctx, [&] {
    if (words < 1 || words > 64) throw Error(VXL_INVALID_ARG, "merge words must be in [1, 64]");
    ctx->sat_merge_words = words;
  });
}

vxl_status vxl_set_batch_refill(vxl_ctx *ctx, int refill_min) {
  return guarded(ctx, [&] {
    if (refill_min < 0 || refill_min > 64)
      throw Error(VXL_INVALID_ARG, "batch refill_min must be 0 (one lane per ray) or in [1, 64]");
    ctx->dda_batch_refill = refill_min;
  });
}

vxl_status vxl_set_dda_refill(vxl_ctx *ctx, int refill_min) {
  return guarded(ctx, [&] {
    if (refill_min < -1 || refill_min > 64)
      throw Error(VXL_INVALID_ARG, "refill_min must be in [-1, 64]");
    ctx->dda_refill = refill_min;
  });
}

vxl_status vxl_voxelize_tree_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                 const vxl_grid *grid, float epsilon, vxl_tree **out,
                                 uint64_t *pairs_out, void *stream) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    uint32_t k = 0;
    validate_grid(grid, &k);
    if (T && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    validate_count(T);
    validate_eps(epsilon);
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user && user != ctx->stream) {
      hipEvent_t ev;
      VXL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      VXL_HIP(hipEventRecord(ev, user));
      VXL_HIP(hipStreamWaitEvent(ctx->stream, ev, 0));
      VXL_HIP(hipEventDestroy(ev));
    }
    *out = voxelize_tree_impl(ctx, d_tris, T, grid, k, epsilon, pairs_out);
    ctx->trees.insert(*out);
    if (user && user != ctx->stream) {
      hipEvent_t ev;
      VXL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      VXL_HIP(hipEventRecord(ev, ctx->stream));
      VXL_HIP(hipStreamWaitEvent(user, ev, 0));
      VXL_HIP(hipEventDestroy(ev));
    }
  });
}

vxl_status vxl_voxelize_tree(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                             float epsilon, vxl_tree **out, uint64_t *pairs_out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    uint32_t k = 0;
    validate_grid(grid, &k);
    validate_count(T);
    upload_mesh_host(ctx, tris, T);  // validates, then copies
    validate_eps(epsilon);
    const float *d = nullptr;
    if (T) d = ctx->tris_in.as<float>();
    *out = voxelize_tree_impl(ctx, d, T, grid, k, epsilon, pairs_out);
    ctx->trees.insert(*out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_voxelize_dense_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                  const vxl_grid *grid, float epsilon, uint32_t *d_occ_words,
                                  uint64_t *pairs_out, void *stream) {
  return guarded(ctx, [&] {
    uint32_t k = 0;
    validate_grid(grid, &k);
    if (T && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    validate_count(T);
    validate_eps(epsilon);
    if (!d_occ_words) throw Error(VXL_INVALID_ARG, "occupancy output is null");
    const uint64_t nn = grid->n;
    const uint64_t nwords = (nn * nn * nn + 31) / 32;
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(user));
    if (T == 0) {
      VXL_HIP(hipMemsetAsync(d_occ_words, 0, nwords * 4, ctx->stream));
      if (pairs_out) *pairs_out = 0;
    } else {
      vxl::prepare_grid(ctx, grid->n);
      vxl::run_sat(ctx, d_tris, T, make_params(grid, epsilon));
      if (pairs_out) *pairs_out = read_pairs(ctx);
      vxl::grid_to_words(ctx, grid->n, d_occ_words);
    }
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_voxelize_dense(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                              float epsilon, uint32_t *occ_words, uint64_t *pairs_out) {
  return guarded(ctx, [&] {
    uint32_t k = 0;
    validate_grid(grid, &k);
    validate_count(T);
    upload_mesh_host(ctx, tris, T);  // validates, then copies
    validate_eps(epsilon);
    if (!occ_words) throw Error(VXL_INVALID_ARG, "occupancy output is null");
    const uint64_t nn = grid->n;
    const uint64_t nwords = (nn * nn * nn + 31) / 32;
    if (T == 0) {
      std::memset(occ_words, 0, nwords * 4);
      if (pairs_out) *pairs_out = 0;
      return;
    }
    ctx->words_io.reserve(nwords * 4);
    vxl::prepare_grid(ctx, grid->n);
    vxl::run_sat(ctx, ctx->tris_in.as<float>(), T, make_params(grid, epsilon));
    if (pairs_out) *pairs_out = read_pairs(ctx);
    vxl::grid_to_words(ctx, grid->n, ctx->words_io.as<uint32_t>());
    download_host(ctx, occ_words, ctx->words_io.ptr, nwords * 4);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

// The dense voxelize with per-voxel owners and colours into device arrays: occupancy words,
// owner (nullable) and color (nullable), n^3 u32 each; a colour without an owner array
// takes its owners from ctx scratch.
static void dense_owner_device(vxl_ctx *ctx, const float *d_tris, uint64_t T, const vxl_grid *grid,
                               float epsilon, uint32_t *d_words, uint32_t *d_owner,
                               uint32_t *d_color, uint64_t *pairs_out) {
  const uint64_t nn = grid->n;
  const uint64_t nv = nn * nn * nn;
  const uint64_t nwords = (nv + 31) / 32;
  uint32_t *own = d_owner;
  if (!own && d_color) {
    ctx->owner_scratch.reserve(nv * 4);
    own = ctx->owner_scratch.as<uint32_t>();
  }
  if (own) VXL_HIP(hipMemsetAsync(own, 0xFF, nv * 4, ctx->stream));  // UINT32_MAX: no owner
  if (T == 0) {
    VXL_HIP(hipMemsetAsync(d_words, 0, nwords * 4, ctx->stream));
    if (pairs_out) *pairs_out = 0;
  } else {
    vxl::prepare_grid(ctx, grid->n);
    vxl::run_sat(ctx, d_tris, T, make_params(grid, epsilon), own);
    if (pairs_out) *pairs_out = read_pairs(ctx);
    vxl::grid_to_words(ctx, grid->n, d_words);
  }
  if (d_color) vxl::owner_colors(ctx, own, d_color, nv);
}

vxl_status vxl_voxelize_dense_owner_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                        const vxl_grid *grid, float epsilon,
                                        uint32_t *d_occ_words, uint32_t *d_owner,
                                        uint32_t *d_color, uint64_t *pairs_out, void *stream) {
  return guarded(ctx, [&] {
    uint32_t k = 0;
    validate_grid(grid, &k);
    if (T && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    validate_count(T);
    validate_eps(epsilon);
    if (!d_occ_words) throw Error(VXL_INVALID_ARG, "occupancy output is null");
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(user));
    dense_owner_device(ctx, d_tris, T, grid, epsilon, d_occ_words, d_owner, d_color, pairs_out);
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_voxelize_dense_owner(vxl_ctx *ctx, const float *tris, uint64_t T,
                                    const vxl_grid *grid, float epsilon, uint32_t *occ_words,
                                    uint32_t *owner, uint32_t *color, uint64_t *pairs_out) {
  return guarded(ctx, [&] {
    uint32_t k = 0;
    validate_grid(grid, &k);
    validate_count(T);
    upload_mesh_host(ctx, tris, T);  // validates, then copies
    validate_eps(epsilon);
    if (!occ_words) throw Error(VXL_INVALID_ARG, "occupancy output is null");
    const uint64_t nn = grid->n;
    const uint64_t nv = nn * nn * nn;
    const uint64_t nwords = (nv + 31) / 32;
    const float *d = nullptr;
    if (T) d = ctx->tris_in.as<float>();
    ctx->words_io.reserve(nwords * 4);
    uint32_t *d_owner = nullptr, *d_color = nullptr;
    if (owner) {
      ctx->owner_io.reserve(nv * 4);
      d_owner = ctx->owner_io.as<uint32_t>();
    }
    if (color) {
      ctx->color_io.reserve(nv * 4);
      d_color = ctx->color_io.as<uint32_t>();
    }
    dense_owner_device(ctx, d, T, grid, epsilon, ctx->words_io.as<uint32_t>(), d_owner, d_color,
                       pairs_out);
    download_host(ctx, occ_words, ctx->words_io.ptr, nwords * 4);
    if (owner)
      download_host(ctx, owner, d_owner, nv * 4);
    if (color)
      download_host(ctx, color, d_color, nv * 4);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_voxelize_tree_materials(vxl_ctx *ctx, const float *tris, uint64_t T,
                                       const vxl_grid *grid, float epsilon,
                                       const uint32_t *material_table, uint32_t table_len,
                                       vxl_tree **out, uint64_t *pairs_out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    uint32_t k = 0;
    validate_grid(grid, &k);
    validate_count(T);
    upload_mesh_host(ctx, tris, T);  // validates, then copies
    validate_eps(epsilon);
    if (table_len && !material_table) throw Error(VXL_INVALID_ARG, "material table is null");

    const uint32_t *d_table = nullptr;
    if (table_len) {
      ctx->words_io.reserve(uint64_t(table_len) * 4);
      VXL_HIP(hipMemcpyAsync(ctx->words_io.ptr, material_table, uint64_t(table_len) * 4,
                             hipMemcpyHostToDevice, ctx->stream));
      d_table = ctx->words_io.as<uint32_t>();
    }
    *out = vxl::voxelize_tree_materials(ctx, ctx->tris_in.as<float>(), T, grid->n, k,
                                        make_params(grid, epsilon), d_table, table_len,
                                        pairs_out);
    ctx->trees.insert(*out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_voxelize_tree_materials_dev(vxl_ctx *ctx, const float *d_tris, uint64_t T,
                                           const vxl_grid *grid, float epsilon,
                                           const uint32_t *d_table, uint32_t table_len,
                                           vxl_tree **out, uint64_t *pairs_out, void *stream) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    uint32_t k = 0;
    validate_grid(grid, &k);
    if (T && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    validate_count(T);
    validate_eps(epsilon);
    if (table_len && !d_table) throw Error(VXL_INVALID_ARG, "material table is null");
    hipStream_t user = static_cast<hipStream_t>(stream);
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(user));
    if (ctx->timing) vxl::mark(ctx, 0);
    *out = vxl::voxelize_tree_materials(ctx, d_tris, T, grid->n, k, make_params(grid, epsilon),
                                        d_table, table_len, pairs_out);
    ctx->trees.insert(*out);
    if (ctx->timing) {
      vxl::mark(ctx, 3);
      VXL_HIP(hipEventSynchronize(ctx->ev[3]));
      VXL_HIP(hipEventElapsedTime(&ctx->last_ms[6], ctx->ev[0], ctx->ev[3]));
    }
    if (user && user != ctx->stream) VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_compact_surface(vxl_ctx *ctx, const float *tris, uint64_t T, const vxl_grid *grid,
                               float epsilon, const uint32_t *material_table, uint32_t table_len,
                               const int32_t g_origin[3], vxl_compact_voxel *out,
                               uint64_t capacity, uint64_t *count) {
  vxl_tree *tree = nullptr;
  vxl_status st = vxl_voxelize_tree_materials(ctx, tris, T, grid, epsilon, material_table,
                                              table_len, &tree, nullptr);
  if (st != VXL_OK) return st;
  st = guarded(ctx, [&] {
    if (!count) throw Error(VXL_INVALID_ARG, "count is null");
    const int32_t g0[3] = {g_origin ? g_origin[0] : 0, g_origin ? g_origin[1] : 0,
                           g_origin ? g_origin[2] : 0};
    // count first; stage and copy only when the caller's buffer holds them all
    const uint64_t total = vxl::compact_tree(ctx, tree, g0, nullptr, 0);
    *count = total;
    if (out && total) {
      if (capacity < total)
        throw Error(VXL_STORAGE_EXCEEDED, "compact: " + std::to_string(total) +
                                              " voxels exceed the output capacity");
      ctx->compact_io.reserve(total * 16);
      vxl::compact_tree(ctx, tree, g0, ctx->compact_io.as<int32_t>(), total);
      download_host(ctx, out, ctx->compact_io.ptr, total * 16);
      VXL_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
  vxl_tree_free(tree);
  return st;
}

vxl_status vxl_tree_from_compact(vxl_ctx *ctx, uint32_t n, const vxl_compact_voxel *voxels,
                                 uint64_t count, vxl_tree **out, uint64_t *dropped) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    vxl_grid g{n, {0, 0, 0}, 1.0f};
    uint32_t k = 0;
    validate_grid(&g, &k);
    if (count && !voxels) throw Error(VXL_INVALID_ARG, "voxels pointer is null");
    if (count) {
      ctx->compact_io.reserve(count * 16);
      upload_host(ctx, ctx->compact_io.ptr, voxels, count * 16);
    }
    *out = vxl::tree_from_compact(ctx, n, k, ctx->compact_io.as<int32_t>(), count, dropped);
    ctx->trees.insert(*out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_tree_materials_download(const vxl_tree *tree, uint16_t *materials) {
  if (!tree) {
    vxl::set_error("null tree");
    return VXL_INVALID_ARG;
  }
  return guarded(tree->ctx, [&] {
    if (!materials) throw Error(VXL_INVALID_ARG, "materials is null");
    const size_t bytes = size_t(tree->leaf_count) * 1024;
    if (!bytes) return;
    if (tree->materials) {
      download_host(tree->ctx, materials, tree->materials, bytes);
      VXL_HIP(hipStreamSynchronize(tree->ctx->stream));
    } else {
      std::memset(materials, 0, bytes);  // a tree built without materials: all MISSING
    }
  });
}

vxl_status vxl_tree_set_voxels(vxl_ctx *ctx, vxl_tree *tree, const vxl_voxel_edit *edits,
                               uint64_t count, uint64_t *changed, int *topology) {
  return guarded(ctx, [&] {
    vxl::check_tree(tree);
    if (tree->ctx != ctx) throw Error(VXL_INVALID_ARG, "tree belongs to another context");
    if (count && !edits) throw Error(VXL_INVALID_ARG, "edits pointer is null");
    if (changed) *changed = 0;
    if (topology) *topology = 0;
    // in order, last edit per voxel wins; out-of-bounds edits are no-ops (Edit::Unchanged)
    std::unordered_map<uint64_t, uint64_t> last;
    last.reserve(size_t(count));
    const uint64_t nn = tree->n;
    for (uint64_t i = 0; i < count; i++) {
      const vxl_voxel_edit &e = edits[i];
      if (e.x >= nn || e.y >= nn || e.z >= nn) continue;
      last[uint64_t(e.x) + nn * (uint64_t(e.y) + nn * uint64_t(e.z))] = i;
    }
    if (last.empty()) return;
    std::vector<uint64_t> order;
    order.reserve(last.size());
    for (const auto &kv : last) order.push_back(kv.second);
    std::sort(order.begin(), order.end());
    const uint64_t E = order.size();
    vxl::staging_begin(ctx, E * sizeof(vxl_voxel_edit));
    vxl_voxel_edit *h = static_cast<vxl_voxel_edit *>(vxl::staging_take(ctx, E * sizeof(vxl_voxel_edit)));
    for (uint64_t i = 0; i < E; i++) {
      h[i] = edits[order[i]];
      h[i].occupied = h[i].occupied ? 1u : 0u;
    }
    ctx->edit_io.reserve(E * sizeof(vxl_voxel_edit));
    VXL_HIP(hipMemcpyAsync(ctx->edit_io.ptr, h, E * sizeof(vxl_voxel_edit), hipMemcpyHostToDevice,
                           ctx->stream));
    vxl::staging_end(ctx, ctx->stream);
    const vxl::EditResult r =
        vxl::edit_tree(ctx, tree, ctx->edit_io.as<vxl_voxel_edit>(), E);
    if (changed) *changed = r.changed;
    if (topology) *topology = r.topology ? 1 : 0;
  });
}

vxl_status vxl_generate_noise_tree(vxl_ctx *ctx, uint32_t n, const vxl_noise *field,
                                   vxl_tree **out) {
  return guarded(ctx, [&] {
    if (!out || !field) throw Error(VXL_INVALID_ARG, "null field or out");
    *out = nullptr;
    vxl_grid g{n, {0, 0, 0}, 1.0f};
    uint32_t k = 0;
    validate_grid(&g, &k);
    if (field->octaves > 64) throw Error(VXL_INVALID_ARG, "more than 64 octaves");
    *out = vxl::generate_noise_tree(ctx, n, k, *field);
    ctx->trees.insert(*out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

struct vxl_graph {
  vxl_ctx *ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

vxl_status vxl_traverse_graph_create(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                                     const float *d_rays, const uint64_t *ray_offsets,
                                     uint32_t *d_hits, double *d_t, vxl_graph **out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    traverse_batch_checks(trees, count, ray_offsets);
    if (count && ray_offsets[count] && (!d_rays || !d_hits))
      throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    const auto segs = segments(trees, count, ray_offsets, d_rays, d_hits, d_t);
    // The traversals are independent: the graph forks into kBranches chains (segments
    // dealt by ray count to the least loaded chain) that join at the end, so its kernel
    // nodes can run side by side instead of one after another.
    constexpr int kBranches = 4;
    hipStream_t cs = nullptr, side[kBranches] = {};
    hipEvent_t fork = nullptr, joins[kBranches] = {};
    auto *g = new vxl_graph();
    g->ctx = ctx;
    bool capturing = false;
    auto release = [&] {
      for (int b = 0; b < kBranches; b++) {
        if (side[b]) (void)hipStreamDestroy(side[b]);
        if (joins[b]) (void)hipEventDestroy(joins[b]);
      }
      if (fork) (void)hipEventDestroy(fork);
      if (cs) (void)hipStreamDestroy(cs);
    };
    try {
      VXL_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      VXL_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
      for (int b = 0; b < kBranches; b++) {
        VXL_HIP(hipStreamCreateWithFlags(&side[b], hipStreamNonBlocking));
        VXL_HIP(hipEventCreateWithFlags(&joins[b], hipEventDisableTiming));
      }
      VXL_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      capturing = true;
      VXL_HIP(hipEventRecord(fork, cs));
      for (int b = 0; b < kBranches; b++) VXL_HIP(hipStreamWaitEvent(side[b], fork, 0));
      uint64_t load[kBranches] = {};
      for (const auto &sg : segs) {  // kernel launches only: no memset or copy nodes
        int b = 0;
        for (int i = 1; i < kBranches; i++)
          if (load[i] < load[b]) b = i;
        load[b] += sg.R;
        vxl::launch_traverse_plain(ctx, sg.tree, sg.rays, sg.R, sg.hits, sg.t, side[b]);
      }
      for (int b = 0; b < kBranches; b++) {
        VXL_HIP(hipEventRecord(joins[b], side[b]));
        VXL_HIP(hipStreamWaitEvent(cs, joins[b], 0));
      }
      capturing = false;
      VXL_HIP(hipStreamEndCapture(cs, &g->graph));
      VXL_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    } catch (...) {
      if (capturing) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(cs, &dead);
        if (dead) (void)hipGraphDestroy(dead);
      }
      if (g->graph) (void)hipGraphDestroy(g->graph);
      delete g;
      release();
      throw;
    }
    release();
    *out = g;
  });
}

vxl_status vxl_graph_launch(vxl_graph *graph, void *stream) {
  if (!graph) {
    vxl::set_error("null graph");
    return VXL_INVALID_ARG;
  }
  return guarded(graph->ctx, [&] {
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : graph->ctx->stream;
    VXL_HIP(hipGraphLaunch(graph->exec, s));
  });
}

void vxl_graph_destroy(vxl_graph *graph) {
  if (!graph) return;
  if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph) (void)hipGraphDestroy(graph->graph);
  delete graph;
}

vxl_status vxl_batch_graph_create(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                                  uint32_t count, const vxl_grid *grids, float epsilon,
                                  const float *d_rays, const uint64_t *ray_offsets,
                                  uint32_t *d_hits, double *d_t, double headroom,
                                  vxl_batch_graph **out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    if (count == 0) throw Error(VXL_INVALID_ARG, "batch graph: no meshes");
    if (!tri_offsets || !grids) throw Error(VXL_INVALID_ARG, "null offsets or grids");
    if (headroom == 0.0) headroom = 1.5;
    if (!(headroom >= 1.0 && headroom <= 64.0))
      throw Error(VXL_INVALID_ARG, "batch graph: headroom must lie in [1, 64] (0: 1.5)");
    validate_eps(epsilon);
    std::vector<vxl::BatchInput> in;
    for (uint32_t g = 0; g < count; g++) {
      uint32_t k = 0;
      validate_grid(&grids[g], &k);
      if (k < 2 || k > 4)
        throw Error(VXL_INVALID_ARG, "batch graph: every mesh needs n in 128..2048");
      if (tri_offsets[g + 1] < tri_offsets[g])
        throw Error(VXL_INVALID_ARG, "tri_offsets must be non-decreasing");
      const uint64_t T = tri_offsets[g + 1] - tri_offsets[g];
      validate_count(T);
      in.push_back(vxl::BatchInput{g, d_tris + 9 * tri_offsets[g], T, grids[g].n, k,
                                   make_params(&grids[g], epsilon)});
    }
    if (tri_offsets[count] > tri_offsets[0] && !d_tris)
      throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    std::vector<uint64_t> roff(count + 1, 0);
    if (ray_offsets) {
      for (uint32_t g = 0; g < count; g++)
        if (ray_offsets[g + 1] < ray_offsets[g])
          throw Error(VXL_INVALID_ARG, "ray_offsets must be non-decreasing");
      for (uint32_t g = 0; g <= count; g++) roff[g] = ray_offsets[g];
      if (roff[count] > roff[0] && (!d_rays || !d_hits))
        throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    }
    auto *bg = new vxl_batch_graph();
    bg->ctx = ctx;
    try {
      vxl::batch_graph_build(bg, in, d_rays, roff.data(), d_hits, d_t, headroom);
    } catch (...) {
      vxl::batch_graph_release(bg);
      delete bg;
      throw;
    }
    *out = bg;
  });
}

vxl_status vxl_batch_graph_launch(vxl_batch_graph *graph, void *stream) {
  if (!graph) {
    vxl::set_error("null batch graph");
    return VXL_INVALID_ARG;
  }
  return guarded(graph->ctx, [&] {
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : graph->ctx->stream;
    vxl::batch_graph_launch(graph, s);
  });
}

vxl_status vxl_batch_graph_trees(vxl_batch_graph *graph, vxl_tree **trees, uint64_t *pairs) {
  if (!graph) {
    vxl::set_error("null batch graph");
    return VXL_INVALID_ARG;
  }
  return guarded(graph->ctx, [&] { vxl::batch_graph_result(graph, trees, pairs); });
}

vxl_status vxl_batch_graph_info(const vxl_batch_graph *graph, uint64_t *arena_bytes,
                                uint32_t *groups, uint64_t *graph_nodes) {
  if (!graph) {
    vxl::set_error("null batch graph");
    return VXL_INVALID_ARG;
  }
  if (arena_bytes) *arena_bytes = graph->arena_bytes;
  if (groups) *groups = uint32_t(graph->groups ? graph->groups->size() : 0);
  if (graph_nodes) *graph_nodes = graph->kernel_nodes;
  return VXL_OK;
}

void vxl_batch_graph_destroy(vxl_batch_graph *graph) {
  if (!graph) return;
  {
    std::lock_guard<std::mutex> lock(graph->ctx->mu);
    (void)hipSetDevice(graph->ctx->device);
    vxl::batch_graph_release(graph);
  }
  delete graph;
}

vxl_status vxl_traverse_stats(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                              uint64_t out[3]) {
  return guarded(ctx, [&] {
    vxl::check_tree(tree);
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    if (R && !d_rays) throw Error(VXL_INVALID_ARG, "null ray buffer");
    vxl::traverse_stats(ctx, tree, d_rays, R, out);
  });
}

static void validate_camera(const vxl_tree *tree, const vxl_camera *cam) {
  vxl::check_tree(tree);
  if (!cam) throw Error(VXL_INVALID_ARG, "camera is null");
  if (cam->n != float(tree->n))
    throw Error(VXL_INVALID_ARG, "camera n must equal the tree's grid size");
  if (cam->width > 65536 || cam->height > 65536)
    throw Error(VXL_INVALID_ARG, "image larger than 65536 pixels on a side");
}

vxl_status vxl_render_dev(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera *camera,
                          const uint32_t *d_colors, uint32_t ncolors, uint32_t *d_rgba,
                          void *stream) {
  return guarded(ctx, [&] {
    validate_camera(tree, camera);
    if (ncolors && !d_colors) throw Error(VXL_INVALID_ARG, "colour table is null");
    if (uint64_t(camera->width) * camera->height && !d_rgba)
      throw Error(VXL_INVALID_ARG, "image buffer is null");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (ctx->timing) VXL_HIP(hipEventRecord(ctx->ev[4], s));
    vxl::launch_render(ctx, tree, *camera, d_colors, ncolors, d_rgba, s);
    if (ctx->timing) {
      VXL_HIP(hipEventRecord(ctx->ev[5], s));
      VXL_HIP(hipEventSynchronize(ctx->ev[5]));
      VXL_HIP(hipEventElapsedTime(&ctx->last_ms[7], ctx->ev[4], ctx->ev[5]));
    }
  });
}

vxl_status vxl_render(vxl_ctx *ctx, const vxl_tree *tree, const vxl_camera *camera,
                      const uint32_t *colors, uint32_t ncolors, uint32_t *rgba) {
  return guarded(ctx, [&] {
    validate_camera(tree, camera);
    if (ncolors && !colors) throw Error(VXL_INVALID_ARG, "colour table is null");
    const uint64_t P = uint64_t(camera->width) * camera->height;
    if (P && !rgba) throw Error(VXL_INVALID_ARG, "image buffer is null");
    if (!P) return;
    ctx->hits_io.reserve(P * 4 + uint64_t(ncolors) * 4 + 256);
    uint32_t *d_img = ctx->hits_io.as<uint32_t>();
    uint32_t *d_col = nullptr;
    if (ncolors) {
      d_col = d_img + ((P + 63) / 64) * 64;
      VXL_HIP(hipMemcpyAsync(d_col, colors, uint64_t(ncolors) * 4, hipMemcpyHostToDevice,
                             ctx->stream));
    }
    vxl::launch_render(ctx, tree, *camera, d_col, ncolors, d_img, ctx->stream);
    download_host(ctx, rgba, d_img, P * 4);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_tree_from_words(vxl_ctx *ctx, const uint32_t *words, uint32_t n, vxl_tree **out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    vxl_grid g{n, {0, 0, 0}, 1.0f};
    uint32_t k = 0;
    validate_grid(&g, &k);
    if (!words) throw Error(VXL_INVALID_ARG, "words is null");
    const uint64_t nn = n;
    const uint64_t nwords = (nn * nn * nn + 31) / 32;
    ctx->words_io.reserve(nwords * 4);
    upload_host(ctx, ctx->words_io.ptr, words, nwords * 4);
    vxl::prepare_grid(ctx, n);
    vxl::words_to_grid(ctx, n, ctx->words_io.as<uint32_t>());
    *out = vxl::build_tree_from_grid(ctx, n, k);
    ctx->trees.insert(*out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_tree_info_get(const vxl_tree *tree, vxl_tree_info *info) {
  if (!tree || !info) {
    vxl::set_error("null tree or info");
    return VXL_INVALID_ARG;
  }
  info->n = tree->n;
  info->k = tree->k;
  info->node_count = tree->node_count;
  info->leaf_count = tree->leaf_count;
  for (int i = 0; i < 16; i++) info->level_nodes[i] = tree->level_nodes[i];
  info->device_bytes = tree->bytes;
  return VXL_OK;
}

vxl_status vxl_tree_download(const vxl_tree *tree, vxl_node *nodes, uint64_t *leaf_words,
                             uint32_t *bounds, uint64_t *codes) {
  if (!tree) {
    vxl::set_error("null tree");
    return VXL_INVALID_ARG;
  }
  return guarded(tree->ctx, [&] {
    vxl_ctx *ctx = tree->ctx;
    if (nodes && tree->node_count)
      download_host(ctx, nodes, tree->nodes, tree->node_count * 12);
    if (leaf_words && tree->leaf_count)
      download_host(ctx, leaf_words, tree->leaves, tree->leaf_count * 64);
    if (bounds && tree->leaf_count)
      download_host(ctx, bounds, tree->bounds, tree->leaf_count * 4);
    if (codes && tree->leaf_count)
      download_host(ctx, codes, tree->codes, tree->leaf_count * 8);
  });
}

vxl_status vxl_tree_upload(vxl_ctx *ctx, uint32_t n, const vxl_node *nodes, uint64_t node_count,
                           const uint64_t *leaf_words, const uint32_t *bounds,
                           uint64_t leaf_count, vxl_tree **out) {
  return guarded(ctx, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    *out = nullptr;
    vxl_grid g{n, {0, 0, 0}, 1.0f};
    uint32_t k = 0;
    validate_grid(&g, &k);
    if ((node_count && !nodes) || (leaf_count && (!leaf_words || !bounds)))
      throw Error(VXL_INVALID_ARG, "null structure array");
    if (leaf_count && k >= 1 && node_count == 0)
      throw Error(VXL_INVALID_ARG, "a non-empty tree with k >= 1 needs nodes");
    if (leaf_count) validate_school_b(k, nodes, node_count, leaf_count);
    vxl_tree *t = new vxl_tree();
    t->ctx = ctx;
    t->n = n;
    t->k = k;
    t->node_count = node_count;
    t->leaf_count = leaf_count;
    try {
      VXL_HIP(hipMalloc(&t->nodes, node_count ? node_count * 12 : 16));
      VXL_HIP(hipMalloc(&t->leaves, leaf_count ? leaf_count * 64 : 64));
      VXL_HIP(hipMalloc(&t->bounds, leaf_count ? leaf_count * 4 : 16));
      VXL_HIP(hipMalloc(&t->codes, leaf_count ? leaf_count * 8 : 16));
      hipStream_t s = ctx->stream;
      if (node_count)
        upload_host(ctx, t->nodes, nodes, node_count * 12);
      if (leaf_count) {
        upload_host(ctx, t->leaves, leaf_words, leaf_count * 64);
        upload_host(ctx, t->bounds, bounds, leaf_count * 4);
        VXL_HIP(hipMemsetAsync(t->codes, 0, leaf_count * 8, s));
      }
      VXL_HIP(hipStreamSynchronize(s));
    } catch (...) {
      vxl::tree_free_device(t);
      delete t;
      throw;
    }
    t->bytes = node_count * 12 + leaf_count * 76;
    ctx->trees.insert(t);
    *out = t;
  });
}

void vxl_tree_free(vxl_tree *tree) {
  if (!tree) return;
  if (tree->ctx) {
    std::lock_guard<std::mutex> lock(tree->ctx->mu);
    (void)hipSetDevice(tree->ctx->device);
    (void)hipStreamSynchronize(tree->ctx->stream);
    vxl::tree_free_device(tree);
    tree->ctx->trees.erase(tree);
  }
  delete tree;
}

vxl_status vxl_traverse_dev(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays, uint64_t R,
                            uint32_t *d_hits, double *d_t, void *stream) {
  return guarded(ctx, [&] {
    vxl::check_tree(tree);
    if (R && (!d_rays || !d_hits)) throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (ctx->timing) VXL_HIP(hipEventRecord(ctx->ev[4], s));
    vxl::launch_traverse(ctx, tree, d_rays, R, d_hits, d_t, s);
    if (ctx->timing) {
      VXL_HIP(hipEventRecord(ctx->ev[5], s));
      VXL_HIP(hipEventSynchronize(ctx->ev[5]));
      VXL_HIP(hipEventElapsedTime(&ctx->last_ms[3], ctx->ev[4], ctx->ev[5]));
    }
  });
}

vxl_status vxl_traverse_dev_noalloc(vxl_ctx *ctx, const vxl_tree *tree, const float *d_rays,
                                    uint64_t R, uint32_t *d_hits, double *d_t, void *stream) {
  return guarded(ctx, [&] {
    vxl::check_tree(tree);
    if (R && (!d_rays || !d_hits)) throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    if (R > kNoAllocMaxRays)
      throw Error(VXL_STORAGE_EXCEEDED, "allocation-free traversal takes at most 2^40 rays");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    // kernels only: the persistent schedule's counter is a slot of the per-device scratch
    vxl::launch_traverse(ctx, tree, d_rays, R, d_hits, d_t, s,
                         R ? vxl::noalloc_counter(ctx->device) : nullptr);
  });
}

vxl_status vxl_voxelize_batch_dev(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                                  uint32_t count, const vxl_grid *grids, float epsilon,
                                  vxl_tree **out, uint64_t *pairs_out, void *stream) {
  return guarded(ctx, [&] {
    if (count && (!tri_offsets || !grids || !out))
      throw Error(VXL_INVALID_ARG, "null offsets, grids or out");
    hipStream_t user = static_cast<hipStream_t>(stream);
    join_streams(user, ctx->stream);
    if (ctx->timing) VXL_HIP(hipEventRecord(ctx->ev[6], ctx->stream));
    voxelize_batch_impl(ctx, d_tris, tri_offsets, count, grids, epsilon, out, pairs_out);
    if (ctx->timing) {
      VXL_HIP(hipEventRecord(ctx->ev[7], ctx->stream));
      VXL_HIP(hipEventSynchronize(ctx->ev[7]));
      VXL_HIP(hipEventElapsedTime(&ctx->last_ms[4], ctx->ev[6], ctx->ev[7]));
    }
    join_streams(ctx->stream, user);
  });
}

vxl_status vxl_voxelize_batch(vxl_ctx *ctx, const float *tris, const uint64_t *tri_offsets,
                              uint32_t count, const vxl_grid *grids, float epsilon, vxl_tree **out,
                              uint64_t *pairs_out) {
  return guarded(ctx, [&] {
    if (count && (!tri_offsets || !grids || !out))
      throw Error(VXL_INVALID_ARG, "null offsets, grids or out");
    const uint64_t T = count ? tri_offsets[count] - tri_offsets[0] : 0;
    if (count && tri_offsets[0] != 0) throw Error(VXL_INVALID_ARG, "tri_offsets[0] must be 0");
    upload_mesh_host(ctx, tris, T);  // validates, then copies
    const float *d = nullptr;
    if (T) d = ctx->tris_in.as<float>();
    voxelize_batch_impl(ctx, d, tri_offsets, count, grids, epsilon, out, pairs_out);
    VXL_HIP(hipStreamSynchronize(ctx->stream));
  });
}

vxl_status vxl_traverse_batch_dev(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                                  const float *d_rays, const uint64_t *ray_offsets,
                                  uint32_t *d_hits, double *d_t, void *stream) {
  return guarded(ctx, [&] {
    traverse_batch_checks(trees, count, ray_offsets);
    if (count && ray_offsets[count] > ray_offsets[0] && (!d_rays || !d_hits))
      throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    if (ctx->timing) VXL_HIP(hipEventRecord(ctx->ev[4], s));
    const auto segs = segments(trees, count, ray_offsets, d_rays, d_hits, d_t);
    vxl::launch_traverse_batch(ctx, segs.data(), count, s);
    if (ctx->timing) {
      VXL_HIP(hipEventRecord(ctx->ev[5], s));
      VXL_HIP(hipEventSynchronize(ctx->ev[5]));
      VXL_HIP(hipEventElapsedTime(&ctx->last_ms[5], ctx->ev[4], ctx->ev[5]));
    }
  });
}

vxl_status vxl_traverse_batch(vxl_ctx *ctx, const vxl_tree *const *trees, uint32_t count,
                              const float *rays, const uint64_t *ray_offsets, uint32_t *hits,
                              double *t) {
  return guarded(ctx, [&] {
    traverse_batch_checks(trees, count, ray_offsets);
    if (count && ray_offsets[0] != 0) throw Error(VXL_INVALID_ARG, "ray_offsets[0] must be 0");
    const uint64_t R = count ? ray_offsets[count] : 0;
    if (R == 0) return;
    if (!rays || !hits) throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    hipStream_t s = ctx->stream;
    ctx->rays_io.reserve(R * 24);
    ctx->hits_io.reserve(R * 16);
    if (t) ctx->t_io.reserve(R * 8);
    upload_host(ctx, ctx->rays_io.ptr, rays, R * 24);
    const auto segs = segments(trees, count, ray_offsets, ctx->rays_io.as<float>(),
                               ctx->hits_io.as<uint32_t>(), t ? ctx->t_io.as<double>() : nullptr);
    vxl::launch_traverse_batch(ctx, segs.data(), count, s);
    download_host(ctx, hits, ctx->hits_io.ptr, R * 16);
    if (t) download_host(ctx, t, ctx->t_io.ptr, R * 8);
    VXL_HIP(hipStreamSynchronize(s));
  });
}

vxl_status vxl_set_batch_budget(vxl_ctx *ctx, uint64_t bytes) {
  return guarded(ctx, [&] {
    if (bytes < (1ull << 30) || bytes > (128ull << 30))
      throw Error(VXL_INVALID_ARG, "batch budget must be in [1 GiB, 128 GiB]");
    ctx->pool_budget = bytes;
  });
}

vxl_status vxl_autotune_batch(vxl_ctx *ctx, const float *d_tris, const uint64_t *tri_offsets,
                              uint32_t count, const vxl_grid *grids, float epsilon,
                              const float *d_rays, const uint64_t *ray_offsets, uint32_t *d_hits,
                              double *d_t, vxl_tuning *result) {
  return guarded(ctx, [&] {
    if (!result) throw Error(VXL_INVALID_ARG, "result is null");
    if (count && (!tri_offsets || !grids)) throw Error(VXL_INVALID_ARG, "null offsets or grids");
    std::vector<vxl_tree *> trees(count, nullptr);
    hipEvent_t e0, e1;
    VXL_HIP(hipEventCreate(&e0));
    VXL_HIP(hipEventCreate(&e1));
    auto free_trees = [&] {
      VXL_HIP(hipStreamSynchronize(ctx->stream));
      for (auto &t : trees)
        if (t) {
          ctx->trees.erase(t);
          vxl::tree_free_device(t);
          delete t;
          t = nullptr;
        }
    };
    auto time_voxelize = [&] {
      free_trees();
      VXL_HIP(hipEventRecord(e0, ctx->stream));
      voxelize_batch_impl(ctx, d_tris, tri_offsets, count, grids, epsilon, trees.data(), nullptr);
      VXL_HIP(hipEventRecord(e1, ctx->stream));
      VXL_HIP(hipEventSynchronize(e1));
      float ms = 0;
      VXL_HIP(hipEventElapsedTime(&ms, e0, e1));
      return ms;
    };
    const int waves[] = {16, 20, 24, 28, 32};
    const int units[] = {1, 2, 4};
    const int blocks[] = {64, 128, 256};
    const int merge[] = {16, 32, 48, 64};
    const int refills[] = {8, 16, 32};
    const int keep_r = ctx->dda_batch_refill;
    const int keep_w = ctx->sat_waves_per_cu, keep_u = ctx->sat_units_per_wave;
    const int keep_m = ctx->sat_merge_words;
    try {
      time_voxelize();  // warm: scratch and pools reach their size
      float best = 1e30f;
      int bw = keep_w, bu = keep_u;
      for (int w : waves)
        for (int u : units) {
          ctx->sat_waves_per_cu = w;
          ctx->sat_units_per_wave = u;
          const float ms = time_voxelize();
          if (ms < best) {
            best = ms;
            bw = w;
            bu = u;
          }
        }
      ctx->sat_waves_per_cu = bw;
      ctx->sat_units_per_wave = bu;
      int bm = keep_m;
      for (int m : merge) {
        ctx->sat_merge_words = m;
        const float ms = time_voxelize();
        if (ms < best) {
          best = ms;
          bm = m;
        }
      }
      ctx->sat_merge_words = bm;
      result->sat_waves_per_cu = bw;
      result->sat_units_per_wave = bu;
      result->sat_merge_words = bm;
      result->voxelize_ms = time_voxelize();
      result->dda_block = ctx->dda_block;
      result->dda_batch_refill = ctx->dda_batch_refill;
      result->traverse_ms = 0;
      const bool rays = count && ray_offsets && ray_offsets[count] > ray_offsets[0] && d_rays &&
                        d_hits;
      if (rays) {
        VXL_HIP(hipStreamSynchronize(ctx->stream));
        float bt = 1e30f;
        int bb = ctx->dda_block;
        const auto segs = segments(trees.data(), count, ray_offsets, d_rays, d_hits, d_t);
        ctx->dda_batch_refill = 0;
        for (int b : blocks) {
          ctx->dda_block = b;
          vxl::launch_traverse_batch(ctx, segs.data(), count, ctx->stream);  // warm
          VXL_HIP(hipEventRecord(e0, ctx->stream));
          vxl::launch_traverse_batch(ctx, segs.data(), count, ctx->stream);
          VXL_HIP(hipEventRecord(e1, ctx->stream));
          VXL_HIP(hipEventSynchronize(e1));
          float ms = 0;
          VXL_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms < bt) {
            bt = ms;
            bb = b;
          }
        }
        ctx->dda_block = bb;
        // persistent lanes over the batch's run table, refilled at 8, 16 or 32 idle lanes
        int br = 0;
        for (int refill : refills) {
          ctx->dda_batch_refill = refill;
          vxl::launch_traverse_batch(ctx, segs.data(), count, ctx->stream);  // warm
          VXL_HIP(hipEventRecord(e0, ctx->stream));
          vxl::launch_traverse_batch(ctx, segs.data(), count, ctx->stream);
          VXL_HIP(hipEventRecord(e1, ctx->stream));
          VXL_HIP(hipEventSynchronize(e1));
          float ms = 0;
          VXL_HIP(hipEventElapsedTime(&ms, e0, e1));
          if (ms < bt) {
            bt = ms;
            br = refill;
          }
        }
        ctx->dda_batch_refill = br;
        result->dda_block = bb;
        result->dda_batch_refill = br;
        result->traverse_ms = bt;
      }
      free_trees();
    } catch (...) {
      ctx->sat_waves_per_cu = keep_w;
      ctx->sat_units_per_wave = keep_u;
      ctx->sat_merge_words = keep_m;
      ctx->dda_batch_refill = keep_r;
      (void)hipStreamSynchronize(ctx->stream);
      for (auto &t : trees)
        if (t) {
          ctx->trees.erase(t);
          vxl::tree_free_device(t);
          delete t;
        }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

vxl_status vxl_traverse(vxl_ctx *ctx, const vxl_tree *tree, const float *rays, uint64_t R,
                        uint32_t *hits, double *t) {
  return guarded(ctx, [&] {
    vxl::check_tree(tree);
    if (R == 0) return;
    if (!rays || !hits) throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    hipStream_t s = ctx->stream;
    ctx->rays_io.reserve(R * 24);
    ctx->hits_io.reserve(R * 16);
    if (t) ctx->t_io.reserve(R * 8);
    upload_host(ctx, ctx->rays_io.ptr, rays, R * 24);
    vxl::launch_traverse(ctx, tree, ctx->rays_io.as<float>(), R, ctx->hits_io.as<uint32_t>(),
                         t ? ctx->t_io.as<double>() : nullptr, s);
    download_host(ctx, hits, ctx->hits_io.ptr, R * 16);
    if (t) download_host(ctx, t, ctx->t_io.ptr, R * 8);
    VXL_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
