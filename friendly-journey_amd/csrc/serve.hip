// Request loop that keeps trees resident and serves ray queries and voxelize requests
// with continuous batching.
//
// Two lanes, each with its own context (stream, scratch) and worker thread: the ray lane
// packs every admitted ray query into one multi-tree traversal launch and keeps up to two
// such launches queued on its stream, so the next batch forms while the current one runs;
// the voxelize lane packs admitted meshes into one grouped voxelize batch. The lanes'
// streams run concurrently, so ray and voxelize work overlap on the device. A request is
// admitted when its device footprint fits the byte budget next to everything in flight
// and the bytes the caller holds resident; it leaves the budget when its batch completes.
// Requests that arrive while batches are on the GPU wait in FIFO order and are admitted
// as slots and memory free up.
//
// The plain mode is the baseline: one worker, one stream, one request per launch in
// arrival order, each synchronised before the next starts.
//
// Batching changes only grouping and order of independent requests: every kernel is
// deterministic and a request's result depends on its own inputs alone, so each result
// equals the one the request gets alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "vxl_internal.hpp"

namespace vxl {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

enum Kind { kRays = 0, kVoxelize = 1 };

struct Request {
  uint64_t id = 0;
  Kind kind = kRays;
  RaySegment seg{};
  const float *tris = nullptr;
  uint64_t T = 0;
  uint32_t n = 0, k = 0;
  VoxelizeParams p{};
  uint64_t footprint = 0;
  double t_submit = 0, t_done = 0;
  vxl_tree *result = nullptr;
  uint64_t pairs = 0;
  vxl_status status = VXL_OK;
  std::string error;
  bool done = false;
  bool claimed = false;  // a vxl_server_wait call owns the request (it deletes it)
};

struct InFlight {
  std::vector<Request *> reqs;
  hipEvent_t ev = nullptr;
};

}  // namespace
}  // namespace vxl

struct vxl_server {
  int device = 0;
  vxl_server_config cfg{};
  vxl_ctx *lane[2] = {nullptr, nullptr};  // rays, voxelize; the plain mode uses lane[0]
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<vxl::Request *> pending[2];  // plain mode: pending[0] holds both kinds
  std::unordered_map<uint64_t, vxl::Request *> live;
  uint64_t next_id = 1;
  uint64_t held = 0, inflight = 0, peak = 0;
  bool stopping = false;
  std::vector<std::thread> workers;
  std::vector<double> lat[2];
  uint64_t submitted = 0, completed = 0, failed = 0, batches = 0, batched_requests = 0;
  double first_submit = 0, last_done = 0;
};

namespace vxl {
namespace {

uint64_t ray_footprint(uint64_t R) { return R * (24 + 16 + 8); }

uint64_t voxelize_footprint(uint64_t T, uint32_t n) {
  const uint64_t nb = uint64_t(n / 8);
  const uint64_t bricks = nb * nb * nb;
  const uint64_t region = (bitmap_level1_words(n) + 63) / 64 * 64 * 4096;
  const uint64_t leaves = std::min<uint64_t>(bricks, 27 * T + 64);
  return T * (36 + 64 + 16) + region + leaves * 76 + leaves / 8 * 12;
}

// Marks a request finished; caller holds S->mu.
void finish(vxl_server *S, Request *r, double t, vxl_status st, const std::string &msg) {
  r->t_done = t;
  r->status = st;
  r->error = msg;
  r->done = true;
  S->inflight -= r->footprint;
  if (st == VXL_OK) {
    S->completed++;
    S->lat[r->kind].push_back((t - r->t_submit) * 1e3);
    S->last_done = std::max(S->last_done, t);
  } else {
    S->failed++;
  }
}

// Pops the FIFO head requests of queue q that fit the budget and the batch caps. With
// nothing in flight anywhere, a head that can never fit fails instead of blocking.
std::vector<Request *> admit(vxl_server *S, int q, bool single) {
  std::vector<Request *> out;
  uint64_t rays = 0, tris = 0;
  auto &Q = S->pending[q];
  while (!Q.empty()) {
    Request *r = Q.front();
    if (S->held + S->inflight + r->footprint > S->cfg.budget_bytes) {
      if (S->inflight == 0 && out.empty()) {
        Q.pop_front();
        S->inflight += r->footprint;
        finish(S, r, now_s(), VXL_STORAGE_EXCEEDED,
               "request footprint exceeds the admission budget");
        S->cv_done.notify_all();
        continue;
      }
      break;
    }
    if (!out.empty()) {
      if (single || out.size() >= S->cfg.max_batch_requests) break;
      if (r->kind == kRays && rays + r->seg.R > S->cfg.max_batch_rays) break;
      if (r->kind == kVoxelize && tris + r->T > S->cfg.max_batch_tris) break;
    }
    Q.pop_front();
    S->inflight += r->footprint;
    S->peak = std::max(S->peak, S->held + S->inflight);
    (r->kind == kRays ? rays : tris) += r->kind == kRays ? r->seg.R : r->T;
    out.push_back(r);
    if (single) break;
  }
  return out;
}

vxl_tree *voxelize_one(vxl_ctx *ctx, const Request *r, uint64_t *pairs) {
  if (r->T == 0) {
    *pairs = 0;
    return empty_tree_device(ctx, r->n, r->k);
  }
  prepare_grid(ctx, r->n);
  run_sat(ctx, r->tris, r->T, r->p);
  return build_tree_from_grid(ctx, r->n, r->k, sat_pairs_dev(ctx), pairs);
}

// Enqueues one batch on the lane's stream; returns its completion event. A batch of one
// request (every plain-mode batch, and a batching lane's when requests arrive one at a
// time) runs through the per-request entry points' own kernels, with no segment table to
// upload. A batch of rays takes the configured persistent-lane schedule only when it fills
// the device's resident waves several times over: a smaller one runs one lane per ray,
// where every ray starts at once (persistent waves would walk their runs' rays in turn).
hipEvent_t launch(vxl_ctx *ctx, Kind kind, std::vector<Request *> &batch, bool plain,
                  uint32_t refill) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (plain || batch.size() == 1) {
    Request *r = batch[0];
    if (kind == kRays) {
      launch_traverse(ctx, r->seg.tree, r->seg.rays, r->seg.R, r->seg.hits, r->seg.t,
                      ctx->stream);
    } else {
      r->result = voxelize_one(ctx, r, &r->pairs);
      ctx->trees.insert(r->result);
    }
  } else if (kind == kRays) {
    std::vector<RaySegment> segs(batch.size());
    uint64_t rays = 0;
    for (size_t i = 0; i < batch.size(); i++) {
      segs[i] = batch[i]->seg;
      rays += segs[i].R;
    }
    ctx->dda_batch_refill = refill && rays >= persistent_batch_min_rays(ctx) ? int(refill) : 0;
    launch_traverse_batch(ctx, segs.data(), uint32_t(segs.size()), ctx->stream);
  } else {
    std::vector<BatchInput> in;
    std::vector<vxl_tree *> out(batch.size(), nullptr);
    std::vector<uint64_t> pairs(batch.size(), 0);
    try {
      for (size_t i = 0; i < batch.size(); i++) {
        Request *r = batch[i];
        if (r->T && r->k >= 2 && r->k <= 4) {
          in.push_back(BatchInput{uint32_t(i), r->tris, r->T, r->n, r->k, r->p});
        } else {
          out[i] = voxelize_one(ctx, r, &pairs[i]);
          ctx->trees.insert(out[i]);
        }
      }
      voxelize_batch(ctx, in, out.data(), pairs.data());
    } catch (...) {
      for (vxl_tree *t : out)
        if (t) {
          ctx->trees.erase(t);
          tree_free_device(t);
          delete t;
        }
      throw;
    }
    for (size_t i = 0; i < batch.size(); i++) {
      batch[i]->result = out[i];
      batch[i]->pairs = pairs[i];
    }
  }
  hipEvent_t ev;
  VXL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  VXL_HIP(hipEventRecord(ev, ctx->stream));
  return ev;
}

// Worker of one lane (batching) or of the whole server (plain: lane 0, both kinds).
void worker(vxl_server *S, int lane_id) {
  (void)hipSetDevice(S->device);
  const bool plain = !S->cfg.batching;
  vxl_ctx *ctx = S->lane[lane_id];
  const size_t max_fly = plain ? 1 : (lane_id == kRays ? 2 : 1);
  std::deque<InFlight> flying;
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    if (flying.size() < max_fly) {
      std::vector<Request *> batch = admit(S, lane_id, plain);
      if (!batch.empty()) {
        S->batches++;
        S->batched_requests += batch.size();
        lk.unlock();
        InFlight f;
        f.reqs = batch;
        vxl_status st = VXL_OK;
        std::string msg;
        try {
          f.ev = launch(ctx, batch[0]->kind, f.reqs, plain, S->cfg.batch_refill);
        } catch (const Error &e) {
          st = e.code;
          msg = e.msg;
          ctx->leafgrid_dirty = ctx->pool_dirty = true;
        } catch (const std::exception &e) {
          st = VXL_DEVICE_ERROR;
          msg = e.what();
          ctx->leafgrid_dirty = ctx->pool_dirty = true;
        }
        lk.lock();
        if (st != VXL_OK) {
          const double t = now_s();
          for (Request *r : batch) finish(S, r, t, st, msg);
          S->cv_done.notify_all();
          S->cv_work.notify_all();
        } else {
          flying.push_back(f);
        }
        continue;
      }
    }
    if (!flying.empty()) {
      InFlight f = flying.front();
      flying.pop_front();
      lk.unlock();
      const hipError_t e = hipEventSynchronize(f.ev);
      (void)hipEventDestroy(f.ev);
      const double t = now_s();
      lk.lock();
      for (Request *r : f.reqs)
        finish(S, r, t, e == hipSuccess ? VXL_OK : VXL_DEVICE_ERROR,
               e == hipSuccess ? "" : hipGetErrorString(e));
      S->cv_done.notify_all();
      S->cv_work.notify_all();  // budget freed: the other lane may admit
      continue;
    }
    if (S->stopping && S->pending[lane_id].empty()) break;
    S->cv_work.wait(lk);
  }
}

Request *find(vxl_server *S, uint64_t id) {
  auto it = S->live.find(id);
  if (it == S->live.end()) throw Error(VXL_INVALID_ARG, "unknown request id");
  return it->second;
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0;
  const size_t i = std::min(v.size() - 1, size_t(q * double(v.size() - 1) + 0.5));
  std::nth_element(v.begin(), v.begin() + i, v.end());
  return v[i];
}

template <class F>
vxl_status server_guarded(vxl_server *S, F &&f) {
  if (!S) {
    set_error("null server");
    return VXL_INVALID_ARG;
  }
  try {
    f();
    return VXL_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception &e) {
    set_error(e.what());
    return VXL_DEVICE_ERROR;
  }
}

uint64_t submit(vxl_server *S, Request *r) {
  std::lock_guard<std::mutex> lock(S->mu);
  if (S->stopping) {
    delete r;
    throw Error(VXL_BUSY, "server is shutting down");
  }
  if (r->footprint + S->held > S->cfg.budget_bytes) {
    delete r;
    throw Error(VXL_STORAGE_EXCEEDED, "request footprint exceeds the admission budget");
  }
  r->id = S->next_id++;
  r->t_submit = now_s();
  if (S->submitted == 0) S->first_submit = r->t_submit;
  S->submitted++;
  S->live[r->id] = r;
  S->pending[S->cfg.batching ? r->kind : 0].push_back(r);
  S->cv_work.notify_all();
  return r->id;
}

}  // namespace
}  // namespace vxl

using vxl::Error;

extern "C" {

vxl_status vxl_server_create(int device, const vxl_server_config *cfg, vxl_server **out) {
  if (!out || !cfg) {
    vxl::set_error("null config or out");
    return VXL_INVALID_ARG;
  }
  *out = nullptr;
  if (cfg->batch_refill > 64) {
    vxl::set_error("batch_refill must be 0 (one lane per ray) or in [1, 64]");
    return VXL_INVALID_ARG;
  }
  vxl_server *S = new vxl_server();
  S->device = device;
  S->cfg = *cfg;
  if (!S->cfg.budget_bytes) S->cfg.budget_bytes = 200ull * 1000 * 1000 * 1000;
  if (!S->cfg.max_batch_requests) S->cfg.max_batch_requests = 1024;
  if (!S->cfg.max_batch_rays) S->cfg.max_batch_rays = 64ull << 20;
  if (!S->cfg.max_batch_tris) S->cfg.max_batch_tris = 16ull << 20;
  const int lanes = S->cfg.batching ? 2 : 1;
  for (int i = 0; i < lanes; i++) {
    const vxl_status st = vxl_ctx_create(device, &S->lane[i]);
    if (st != VXL_OK) {
      for (int j = 0; j < i; j++) vxl_ctx_destroy(S->lane[j]);
      delete S;
      return st;
    }
  }
  for (int i = 0; i < lanes; i++) S->workers.emplace_back(vxl::worker, S, i);
  *out = S;
  return VXL_OK;
}

void vxl_server_destroy(vxl_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lock(S->mu);
    S->stopping = true;
    S->cv_work.notify_all();
  }
  for (auto &w : S->workers) w.join();
  for (auto &kv : S->live) delete kv.second;  // results not collected stay with the lanes
  for (vxl_ctx *c : S->lane)
    if (c) vxl_ctx_destroy(c);
  delete S;
}

vxl_status vxl_server_hold(vxl_server *S, int64_t delta_bytes) {
  return vxl::server_guarded(S, [&] {
    std::lock_guard<std::mutex> lock(S->mu);
    if (delta_bytes < 0 && uint64_t(-delta_bytes) > S->held)
      throw Error(VXL_INVALID_ARG, "releasing more bytes than held");
    S->held = uint64_t(int64_t(S->held) + delta_bytes);
    S->peak = std::max(S->peak, S->held + S->inflight);
    S->cv_work.notify_all();
  });
}

vxl_status vxl_server_submit_rays(vxl_server *S, const vxl_tree *tree, const float *d_rays,
                                  uint64_t R, uint32_t *d_hits, double *d_t, uint64_t *id) {
  return vxl::server_guarded(S, [&] {
    vxl::check_tree(tree);
    if (!id) throw Error(VXL_INVALID_ARG, "id is null");
    if (R && (!d_rays || !d_hits)) throw Error(VXL_INVALID_ARG, "null ray or hit buffer");
    auto *r = new vxl::Request();
    r->kind = vxl::kRays;
    r->seg = vxl::RaySegment{tree, d_rays, R, d_hits, d_t};
    r->footprint = vxl::ray_footprint(R);
    *id = vxl::submit(S, r);
  });
}

vxl_status vxl_server_submit_voxelize(vxl_server *S, const float *d_tris, uint64_t T,
                                      const vxl_grid *grid, float epsilon, uint64_t *id) {
  return vxl::server_guarded(S, [&] {
    if (!grid || !id) throw Error(VXL_INVALID_ARG, "null grid or id");
    if (T && !d_tris) throw Error(VXL_INVALID_ARG, "triangles pointer is null");
    auto *r = new vxl::Request();
    try {
      r->k = vxl::check_voxelize_args(grid, T, epsilon);
    } catch (...) {
      delete r;
      throw;
    }
    r->kind = vxl::kVoxelize;
    r->tris = d_tris;
    r->T = T;
    r->n = grid->n;
    r->p = vxl::voxelize_params(grid, epsilon);
    r->footprint = vxl::voxelize_footprint(T, grid->n);
    *id = vxl::submit(S, r);
  });
}

vxl_status vxl_server_poll(vxl_server *S, uint64_t id, int *done) {
  return vxl::server_guarded(S, [&] {
    if (!done) throw Error(VXL_INVALID_ARG, "done is null");
    std::lock_guard<std::mutex> lock(S->mu);
    *done = vxl::find(S, id)->done ? 1 : 0;
  });
}

vxl_status vxl_server_wait(vxl_server *S, uint64_t id, vxl_tree **tree_out, uint64_t *pairs_out,
                           double *latency_ms) {
  if (tree_out) *tree_out = nullptr;
  vxl::Request *r = nullptr;
  const vxl_status st = vxl::server_guarded(S, [&] {
    std::unique_lock<std::mutex> lk(S->mu);
    r = vxl::find(S, id);
    if (r->claimed) throw Error(VXL_INVALID_ARG, "request is already being waited on");
    r->claimed = true;
    S->cv_done.wait(lk, [&] { return r->done; });
    S->live.erase(id);
  });
  if (st != VXL_OK) return st;
  const vxl_status rs = r->status;
  if (rs != VXL_OK) vxl::set_error(r->error);
  if (pairs_out) *pairs_out = r->pairs;
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  if (r->result) {
    if (tree_out) *tree_out = r->result;
    else vxl_tree_free(r->result);
  }
  delete r;
  return rs;
}

vxl_status vxl_server_stats_get(vxl_server *S, vxl_server_stats *out) {
  return vxl::server_guarded(S, [&] {
    if (!out) throw Error(VXL_INVALID_ARG, "out is null");
    std::lock_guard<std::mutex> lock(S->mu);
    std::vector<double> all = S->lat[0];
    all.insert(all.end(), S->lat[1].begin(), S->lat[1].end());
    *out = vxl_server_stats{};
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->failed = S->failed;
    out->batches = S->batches;
    out->peak_bytes = S->peak;
    out->mean_batch_requests = S->batches ? double(S->batched_requests) / double(S->batches) : 0;
    out->p50_ms = vxl::percentile(all, 0.50);
    out->p99_ms = vxl::percentile(all, 0.99);
    out->max_ms = all.empty() ? 0 : *std::max_element(all.begin(), all.end());
    double sum = 0;
    for (double v : all) sum += v;
    out->mean_ms = all.empty() ? 0 : sum / double(all.size());
    out->ray_p50_ms = vxl::percentile(S->lat[0], 0.50);
    out->ray_p99_ms = vxl::percentile(S->lat[0], 0.99);
    out->vox_p50_ms = vxl::percentile(S->lat[1], 0.50);
    out->vox_p99_ms = vxl::percentile(S->lat[1], 0.99);
    const double span = S->last_done - S->first_submit;
    out->throughput_rps = span > 0 ? double(S->completed) / span : 0;
  });
}

vxl_status vxl_server_reset_stats(vxl_server *S) {
  return vxl::server_guarded(S, [&] {
    std::lock_guard<std::mutex> lock(S->mu);
    S->lat[0].clear();
    S->lat[1].clear();
    S->submitted = S->completed = S->failed = S->batches = S->batched_requests = 0;
    S->first_submit = S->last_done = 0;
    S->peak = S->held + S->inflight;
  });
}

}  // extern "C"
