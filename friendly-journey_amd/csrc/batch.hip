// Many meshes per launch: the batch group driver and the pinned staging it uploads from.
//
// A group is a run of meshes whose dense leaf grids fit the context's batch budget. Each
// mesh gets a region of the pooled leaf grid (its (n/8)^3 bricks, rounded up to a whole
// upper-bitmap word of 64 x 64 bricks) and the matching words of the pooled brick bitmap.
// Per group: one setup launch over every triangle, one scan, one persistent SAT launch
// over every (triangle, voxel) pair, one batched tree build -- one host round trip (the
// per-mesh tree sizes), whatever the number of meshes. The pooled grid and bitmap are all-zero between groups
// (the leaf gather clears what it reads), like the single-mesh scratch.
#include "vxl_internal.hpp"

#include <algorithm>

namespace vxl {

void *staging_begin(vxl_ctx *ctx, size_t bytes) {
  if (ctx->staging_free) VXL_HIP(hipEventSynchronize(ctx->staging_free));
  if (ctx->batch_host_bytes < bytes) {
    if (ctx->batch_host) (void)hipHostFree(ctx->batch_host);
    ctx->batch_host = nullptr;
    ctx->batch_host_bytes = 0;
    const size_t b = bytes < (64u << 10) ? (64u << 10) : (bytes + 4095) / 4096 * 4096;
    VXL_HIP(hipHostMalloc(&ctx->batch_host, b, hipHostMallocDefault));
    ctx->batch_host_bytes = b;
    VXL_HIP(hipHostGetDevicePointer(&ctx->batch_host_dev, ctx->batch_host, 0));
  }
  ctx->batch_host_used = 0;
  return ctx->batch_host;
}

void *staging_take(vxl_ctx *ctx, size_t bytes) {
  const size_t b = (bytes + 255) / 256 * 256;
  if (ctx->batch_host_used + b > ctx->batch_host_bytes)
    throw Error(VXL_DEVICE_ERROR, "internal: batch staging overflow");
  void *p = static_cast<char *>(ctx->batch_host) + ctx->batch_host_used;
  ctx->batch_host_used += b;
  return p;
}

void *staging_dev(vxl_ctx *ctx, void *host) {
  return static_cast<char *>(ctx->batch_host_dev) +
         (static_cast<char *>(host) - static_cast<char *>(ctx->batch_host));
}

void staging_end(vxl_ctx *ctx, hipStream_t s) {
  if (!ctx->staging_free)
    VXL_HIP(hipEventCreateWithFlags(&ctx->staging_free, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ctx->staging_free, s));
}

namespace {

uint64_t region_words(uint32_t n) {  // brick-bitmap words of a mesh region
  return (bitmap_level1_words(n) + 63) / 64 * 64;
}

void ensure_pool(vxl_ctx *ctx, uint64_t W1) {
  const size_t grid_bytes = size_t(W1) * 512 * 8;
  const size_t b1_bytes = size_t(W1) * 8 + 256;
  bool fresh = false;
  if (ctx->pool_grid.bytes < grid_bytes) {
    ctx->pool_grid.reserve(grid_bytes);
    fresh = true;
  }
  if (ctx->pool_b1.bytes < b1_bytes) {
    ctx->pool_b1.reserve(b1_bytes);
    fresh = true;
  }
  if (fresh || ctx->pool_dirty) {
    VXL_HIP(hipMemsetAsync(ctx->pool_grid.ptr, 0, ctx->pool_grid.bytes, ctx->stream));
    VXL_HIP(hipMemsetAsync(ctx->pool_b1.ptr, 0, ctx->pool_b1.bytes, ctx->stream));
    ctx->pool_dirty = false;
  }
}

void run_group(vxl_ctx *ctx, const BatchInput *in, uint32_t G, vxl_tree **out,
               uint64_t *pairs) {
  hipStream_t s = ctx->stream;
  staging_begin(ctx, batch_staging_bytes(G));
  BatchMesh *h_mesh = static_cast<BatchMesh *>(staging_take(ctx, G * sizeof(BatchMesh)));
  uint32_t *h_sb = static_cast<uint32_t *>(staging_take(ctx, (G + 1) * 4));
  uint64_t *h_ref = static_cast<uint64_t *>(staging_take(ctx, G * 8));  // zeroed counters
  std::memset(h_ref, 0, G * 8);
  std::vector<BatchMesh> meshes;
  std::vector<uint32_t> sb;
  uint64_t W1 = 0, blocks = 0;
  batch_group_meshes(in, G, meshes, sb, &blocks, &W1);
  std::memcpy(h_sb, sb.data(), (G + 1) * 4);
  std::memcpy(h_mesh, meshes.data(), G * sizeof(BatchMesh));
  ensure_pool(ctx, W1);
  const size_t mesh_bytes = (G * sizeof(BatchMesh) + 255) / 256 * 256;
  const size_t sb_bytes = ((G + 1) * 4 + 255) / 256 * 256;
  const size_t desc_bytes = mesh_bytes + sb_bytes + G * 8;
  ctx->batch_desc.reserve(desc_bytes + 256);
  char *d = static_cast<char *>(ctx->batch_desc.ptr);
  VXL_HIP(hipMemcpyAsync(d, h_mesh, desc_bytes, hipMemcpyHostToDevice, s));
  const BatchMesh *d_mesh = reinterpret_cast<const BatchMesh *>(d);
  unsigned long long *d_ref = reinterpret_cast<unsigned long long *>(d + mesh_bytes + sb_bytes);
  run_sat_batch(ctx, d_mesh, reinterpret_cast<const uint32_t *>(d + mesh_bytes), G,
                uint32_t(blocks), ctx->pool_grid.as<unsigned long long>(),
                ctx->pool_b1.as<unsigned long long>(), d_ref);  // no host round trip
  std::vector<vxl_tree *> trees(G, nullptr);
  std::vector<uint64_t> pr(G, 0);
  build_trees_batch(ctx, meshes, d_mesh, W1, trees.data(), pr.data(), d_ref);
  staging_end(ctx, s);
  for (uint32_t g = 0; g < G; g++) {
    out[in[g].index] = trees[g];
    ctx->trees.insert(trees[g]);
    if (pairs) pairs[in[g].index] = pr[g];
  }
}

uint64_t up256(uint64_t b) { return (b + 255) / 256 * 256; }

}  // namespace

std::vector<std::pair<size_t, size_t>> batch_groups(const vxl_ctx *ctx,
                                                    const std::vector<BatchInput> &in) {
  std::vector<std::pair<size_t, size_t>> out;
  size_t i = 0;
  while (i < in.size()) {
    uint64_t words = 0;
    size_t j = i;
    while (j < in.size()) {
      const uint64_t w = region_words(in[j].n);
      if (j > i && (words + w) * 4096 > ctx->pool_budget) break;
      words += w;
      j++;
    }
    out.emplace_back(i, j);
    i = j;
  }
  return out;
}

void batch_group_meshes(const BatchInput *in, uint32_t G, std::vector<BatchMesh> &meshes,
                        std::vector<uint32_t> &sb_begin, uint64_t *blocks_out, uint64_t *W1_out) {
  const uint64_t per_block = setup_slots_per_block();
  meshes.assign(G, BatchMesh{});
  sb_begin.assign(G + 1, 0);
  uint64_t W1 = 0, blocks = 0;
  for (uint32_t g = 0; g < G; g++) {
    BatchMesh &m = meshes[g];
    m.tris = in[g].tris;
    m.T = in[g].T;
    m.b1_begin = W1;
    m.w1 = bitmap_level1_words(in[g].n);
    m.n = in[g].n;
    m.k = in[g].k;
    m.p = in[g].p;
    sb_begin[g] = uint32_t(blocks);
    m.slot_begin = blocks * per_block;
    blocks += (m.T + per_block - 1) / per_block;
    m.slot_end = blocks * per_block;
    W1 += region_words(m.n);
    if (blocks > 0xffffffffull || W1 > 0xffffffffull)
      throw Error(VXL_WORKGROUPS_EXCEEDED, "voxelize batch: group exceeds the 32-bit limits");
  }
  sb_begin[G] = uint32_t(blocks);
  *blocks_out = blocks;
  *W1_out = W1;
}

void voxelize_batch(vxl_ctx *ctx, const std::vector<BatchInput> &in, vxl_tree **out,
                    uint64_t *pairs) {
  for (const auto &r : batch_groups(ctx, in))
    run_group(ctx, in.data() + r.first, uint32_t(r.second - r.first), out, pairs);
}

// ---- the batch as one HIP graph --------------------------------------------------------

void batch_graph_build(vxl_batch_graph *bg, const std::vector<BatchInput> &in,
                       const float *d_rays, const uint64_t *roff, uint32_t *d_hits,
                       double *d_t, double headroom) {
  vxl_ctx *ctx = bg->ctx;
  const uint32_t count = uint32_t(in.size());
  bg->count = count;
  bg->groups = new std::vector<GraphGroup>();
  std::vector<GraphGroup> &groups = *bg->groups;
  {
    vxl_ctx *own = nullptr;
    const vxl_status st = vxl_ctx_create(ctx->device, &own);
    if (st != VXL_OK) throw Error(st, std::string("batch graph: ") + vxl_last_error());
    bg->own = own;
  }
  vxl_ctx *own = bg->own;
  own->sat_waves_per_cu = ctx->sat_waves_per_cu;
  own->sat_units_per_wave = ctx->sat_units_per_wave;
  own->sat_merge_words = ctx->sat_merge_words;
  own->dda_block = ctx->dda_block;
  own->dda_batch_refill = ctx->dda_batch_refill;
  own->pool_budget = ctx->pool_budget;
  for (const BatchInput &b : in) {
    bg->n.push_back(b.n);
    bg->k.push_back(b.k);
  }

  // 1. one eager run on the private context: grows its scratch to what every group needs
  //    (nothing is allocated while capturing) and sizes each group's arena region
  std::vector<vxl_tree *> eager(count, nullptr);
  std::vector<uint64_t> eager_pairs(count, 0);
  voxelize_batch(own, in, eager.data(), eager_pairs.data());
  wait_stream(own, own->stream);
  const auto ranges = batch_groups(own, in);
  std::vector<uint64_t> need_nodes(ranges.size(), 0), need_leaves(ranges.size(), 0);
  for (size_t r = 0; r < ranges.size(); r++)
    for (size_t g = ranges[r].first; g < ranges[r].second; g++) {
      const vxl_tree *t = eager[g];
      if (t && t->leaf_count) {
        need_nodes[r] += up256(t->node_count * 12);
        need_leaves[r] += t->leaf_count;
      }
    }
  for (vxl_tree *t : eager)
    if (t) {
      own->trees.erase(t);
      tree_free_device(t);
      delete t;
    }

  // 2. descriptors (uploaded once) and the arena
  size_t desc_bytes = 0, arena_bytes = 0, scratch = 0;
  std::vector<std::vector<uint32_t>> sbs(ranges.size());
  groups.resize(ranges.size());
  for (size_t r = 0; r < ranges.size(); r++) {
    GraphGroup &gg = groups[r];
    gg.first = uint32_t(ranges[r].first);
    gg.G = uint32_t(ranges[r].second - ranges[r].first);
    uint64_t blocks = 0;
    batch_group_meshes(in.data() + gg.first, gg.G, gg.meshes, sbs[r], &blocks, &gg.W1);
    gg.setup_blocks = uint32_t(blocks);
    const uint32_t G = gg.G;
    desc_bytes += up256(G * sizeof(BatchMesh)) + up256((G + 1) * 4) + up256(G * 8) +
                  up256(G * sizeof(LevelPtrs)) + up256(G * sizeof(BatchTotals)) +
                  up256((G + 1) * 8) + up256(G * 8) + up256(G * 8) + 256;
    gg.node_cap = up256(uint64_t(double(need_nodes[r]) * headroom)) + 4096;
    gg.leaf_cap = uint64_t(double(need_leaves[r]) * headroom) + 64;
    arena_bytes += gg.node_cap + up256(gg.leaf_cap * 64) + up256(gg.leaf_cap * 4) +
                   up256(gg.leaf_cap * 8);
    scratch = std::max(scratch, group_scratch_bytes(gg.meshes, gg.W1));
    if (own->pool_grid.bytes < gg.W1 * 4096 || own->pool_b1.bytes < gg.W1 * 8)
      throw Error(VXL_DEVICE_ERROR, "internal: batch graph pool smaller than a group");
  }
  const size_t seg_bytes = up256(count * sizeof(TraverseSeg));
  const size_t blk_bytes = up256((count + 1) * 4);
  const size_t run_bytes = up256((count + 1) * 8);
  desc_bytes += seg_bytes + blk_bytes + run_bytes + 256;
  own->batch_scratch.reserve(scratch);
  VXL_HIP(hipMalloc(&bg->desc, desc_bytes));
  bg->desc_bytes = desc_bytes;
  VXL_HIP(hipMalloc(&bg->arena, arena_bytes));
  bg->arena_bytes = arena_bytes;

  std::vector<char> image(desc_bytes, 0);
  char *dbase = static_cast<char *>(bg->desc);
  char *abase = static_cast<char *>(bg->arena);
  size_t doff = 0, aoff = 0;
  auto take = [&](size_t bytes) {
    const size_t o = doff;
    doff += up256(bytes);
    return o;
  };
  std::vector<GroupScratch> gss(groups.size());
  for (size_t r = 0; r < groups.size(); r++) {
    GraphGroup &gg = groups[r];
    const uint32_t G = gg.G;
    const size_t o_mesh = take(G * sizeof(BatchMesh)), o_sb = take((G + 1) * 4),
                 o_ref = take(G * 8), o_lps = take(G * sizeof(LevelPtrs)),
                 o_tot = take(G * sizeof(BatchTotals)), o_n2 = take((G + 1) * 8),
                 o_np = take(G * 8), o_off = take(G * 8), o_st = take(sizeof(GraphState));
    gg.d_mesh = reinterpret_cast<BatchMesh *>(dbase + o_mesh);
    gg.d_sb = reinterpret_cast<uint32_t *>(dbase + o_sb);
    gg.d_ref = reinterpret_cast<unsigned long long *>(dbase + o_ref);
    gg.d_lps = reinterpret_cast<LevelPtrs *>(dbase + o_lps);
    gg.d_tot = reinterpret_cast<BatchTotals *>(dbase + o_tot);
    gg.n2_beg = reinterpret_cast<uint64_t *>(dbase + o_n2);
    gg.nodes_ptr = reinterpret_cast<uint32_t **>(dbase + o_np);
    gg.node_off = reinterpret_cast<uint64_t *>(dbase + o_off);
    gg.state = reinterpret_cast<GraphState *>(dbase + o_st);
    std::memcpy(image.data() + o_mesh, gg.meshes.data(), G * sizeof(BatchMesh));
    std::memcpy(image.data() + o_sb, sbs[r].data(), (G + 1) * 4);
    std::vector<LevelPtrs> lps(G);
    gss[r] = carve_group_scratch(gg.meshes, gg.W1, static_cast<char *>(own->batch_scratch.ptr),
                                 lps.data());
    std::memcpy(image.data() + o_lps, lps.data(), G * sizeof(LevelPtrs));
    gg.node_region = abase + aoff;
    aoff += gg.node_cap;
    gg.leaves_base = abase + aoff;
    aoff += up256(gg.leaf_cap * 64);
    gg.bounds_base = abase + aoff;
    aoff += up256(gg.leaf_cap * 4);
    gg.codes_base = abase + aoff;
    aoff += up256(gg.leaf_cap * 8);
  }
  const size_t o_seg = take(count * sizeof(TraverseSeg)), o_blk = take((count + 1) * 4),
               o_run = take((count + 1) * 8), o_ctr = take(8);
  bg->d_segs = reinterpret_cast<TraverseSeg *>(dbase + o_seg);
  const uint32_t *d_blk = reinterpret_cast<const uint32_t *>(dbase + o_blk);
  const uint64_t *d_run = reinterpret_cast<const uint64_t *>(dbase + o_run);
  unsigned long long *d_ctr = reinterpret_cast<unsigned long long *>(dbase + o_ctr);
  TraverseSeg *hs = reinterpret_cast<TraverseSeg *>(image.data() + o_seg);
  uint32_t *hb = reinterpret_cast<uint32_t *>(image.data() + o_blk);
  uint64_t *hr = reinterpret_cast<uint64_t *>(image.data() + o_run);
  const uint32_t tb = uint32_t(own->dda_block);
  uint64_t blocks = 0, runs = 0;
  uint32_t kmax = 0;
  bool want_t = false;
  for (uint32_t g = 0; g < count; g++) {
    const uint64_t a = roff[g], R = roff[g + 1] - roff[g];
    hs[g] = TraverseSeg{nullptr, nullptr, nullptr, d_rays + 6 * a,
                        reinterpret_cast<uint4 *>(d_hits + 4 * a), d_t ? d_t + a : nullptr, 0, R,
                        in[g].n, in[g].k};
    hb[g] = uint32_t(blocks);
    hr[g] = runs;
    blocks += (R + tb - 1) / tb;
    runs += (R + kBatchRayRun - 1) / kBatchRayRun;
    kmax = std::max(kmax, in[g].k);
    want_t = want_t || (d_t && R);
  }
  hb[count] = uint32_t(blocks);
  hr[count] = runs;
  if (blocks > 0xffffffffull || runs > 0xffffffffull)
    throw Error(VXL_WORKGROUPS_EXCEEDED, "batch graph: rays exceed the grid limit");
  VXL_HIP(hipMemcpy(bg->desc, image.data(), desc_bytes, hipMemcpyHostToDevice));

  // 3. capture: kernels only (the counters are cleared by a kernel, no memset or copy nodes)
  hipStream_t cs = own->stream;
  VXL_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  try {
    for (size_t r = 0; r < groups.size(); r++) {
      const GraphGroup &gg = groups[r];
      fill_u64(gg.d_ref, gg.G, 0ull, cs);
      run_sat_batch(own, gg.d_mesh, gg.d_sb, gg.G, gg.setup_blocks,
                    own->pool_grid.as<unsigned long long>(), own->pool_b1.as<unsigned long long>(),
                    gg.d_ref);
      enqueue_trees_graph(own, gg, gss[r], bg->d_segs + gg.first, cs);
    }
    launch_traverse_segs(own, bg->d_segs, d_blk, d_run, count, blocks, runs, kmax, want_t,
                         d_ctr, cs);
  } catch (...) {
    hipGraph_t dead = nullptr;
    (void)hipStreamEndCapture(cs, &dead);
    if (dead) (void)hipGraphDestroy(dead);
    throw;
  }
  VXL_HIP(hipStreamEndCapture(cs, &bg->graph));
  VXL_HIP(hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0));
  VXL_HIP(hipGraphGetNodes(bg->graph, nullptr, &bg->kernel_nodes));
  VXL_HIP(hipEventCreateWithFlags(&bg->done, hipEventDisableTiming));
}

void batch_graph_launch(vxl_batch_graph *bg, hipStream_t s) {
  // replays share the graph's scratch: one launched on another stream waits for the last
  if (bg->launched) VXL_HIP(hipStreamWaitEvent(s, bg->done, 0));
  VXL_HIP(hipGraphLaunch(bg->exec, s));
  VXL_HIP(hipEventRecord(bg->done, s));
  bg->launched = true;
}

void batch_graph_result(vxl_batch_graph *bg, vxl_tree **trees, uint64_t *pairs) {
  if (!bg->launched) throw Error(VXL_INVALID_ARG, "batch graph: not launched yet");
  VXL_HIP(hipEventSynchronize(bg->done));
  vxl_ctx *ctx = bg->ctx;
  if (trees)
    for (uint32_t g = 0; g < bg->count; g++) trees[g] = nullptr;
  std::string overflow;
  struct Host {
    GraphState st;
    std::vector<BatchTotals> tot;
    std::vector<uint64_t> off;
  };
  std::vector<Host> hs(bg->groups->size());
  for (size_t r = 0; r < bg->groups->size(); r++) {
    const GraphGroup &gg = (*bg->groups)[r];
    Host &h = hs[r];
    h.tot.resize(gg.G);
    h.off.resize(gg.G);
    VXL_HIP(hipMemcpy(&h.st, gg.state, sizeof(GraphState), hipMemcpyDeviceToHost));
    VXL_HIP(hipMemcpy(h.tot.data(), gg.d_tot, gg.G * sizeof(BatchTotals), hipMemcpyDeviceToHost));
    VXL_HIP(hipMemcpy(h.off.data(), gg.node_off, gg.G * 8, hipMemcpyDeviceToHost));
    for (uint32_t g = 0; g < gg.G; g++)
      if (pairs) pairs[gg.first + g] = h.tot[g].pairs;
    if (!h.st.ok && overflow.empty())
      overflow = "batch graph: group " + std::to_string(r) + " needs " +
                 std::to_string(h.st.node_bytes) + " node bytes and " + std::to_string(h.st.leaves) +
                 " leaves; its arena region holds " + std::to_string(gg.node_cap) + " and " +
                 std::to_string(gg.leaf_cap) + " (the hits of its meshes are misses)";
  }
  if (!overflow.empty()) throw Error(VXL_STORAGE_EXCEEDED, overflow);
  if (!trees) return;
  try {
    for (size_t r = 0; r < bg->groups->size(); r++) {
      const GraphGroup &gg = (*bg->groups)[r];
      for (uint32_t g = 0; g < gg.G; g++) {
        const BatchTotals &t = hs[r].tot[g];
        const uint32_t idx = gg.first + g;
        uint64_t lv[16] = {0};
        uint64_t nodes = 0;
        if (t.leaves) {
          lv[2] = t.n2;
          nodes = t.n2;
          for (int L = 3; L <= int(gg.meshes[g].k) + 1; L++) {
            lv[L] = t.nodes[L - 3];
            nodes += t.nodes[L - 3];
          }
        }
        trees[idx] = tree_copy_arrays(ctx, bg->n[idx], bg->k[idx], t.leaves, nodes, lv,
                                      gg.node_region + hs[r].off[g],
                                      gg.leaves_base + t.leaf_begin * 64,
                                      gg.bounds_base + t.leaf_begin * 4,
                                      gg.codes_base + t.leaf_begin * 8, ctx->stream);
        ctx->trees.insert(trees[idx]);
      }
    }
    wait_stream(ctx, ctx->stream);
  } catch (...) {
    for (uint32_t g = 0; g < bg->count; g++)
      if (trees[g]) {
        ctx->trees.erase(trees[g]);
        tree_free_device(trees[g]);
        delete trees[g];
        trees[g] = nullptr;
      }
    throw;
  }
}

void batch_graph_release(vxl_batch_graph *bg) {
  if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
  if (bg->graph) (void)hipGraphDestroy(bg->graph);
  if (bg->done) {
    (void)hipEventSynchronize(bg->done);
    (void)hipEventDestroy(bg->done);
  }
  if (bg->own) vxl_ctx_destroy(bg->own);  // waits for its stream's work
  if (bg->desc) (void)hipFree(bg->desc);
  if (bg->arena) (void)hipFree(bg->arena);
  delete bg->groups;
  bg->exec = nullptr;
  bg->graph = nullptr;
  bg->done = nullptr;
  bg->own = nullptr;
  bg->desc = nullptr;
  bg->arena = nullptr;
  bg->groups = nullptr;
}

}  // namespace vxl
