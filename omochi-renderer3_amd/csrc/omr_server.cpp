// omr_server.cpp -- served mode: continuous batching on the shared queues.
//
// A server thread owns the device stream.  Between bounce-step chunks it
// (1) admits queued requests while the job ring and the HBM budget allow,
//     appending their samples to the global sample range and waking idle
//     path slots (refill), so new work overlaps with work already in flight;
// (2) runs one chunk of bounce steps (a hipGraph), or, once every admitted
//     sample is handed out and few paths remain, one drain launch that ends
//     them all;
// (3) detects finished jobs from their device completion counters, resolves
//     them and copies them to the caller's buffer on a second stream while
//     the next chunk runs, then frees their memory.
// Results are identical to running each job alone: per-path seeds depend only
// on (job seed, pixel, sub-pixel, sample, branch) and accumulation is integer.
#include <algorithm>
#include <chrono>
#include <cstddef>
#include <cstring>

#include "omr_engine.h"

namespace omr {

namespace {
double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}
uint64_t samples_of(const omr_camera &c, const omr_render_params &p) {
  return (uint64_t)c.width * c.height * p.supersamples * p.supersamples * (uint64_t)(p.spp_end - p.spp_begin);
}
}  // namespace

int Engine::server_start(const omr_server_config &cfg) {
  if (server_on_.load()) return fail(OMR_ERR_STATE, "server already running");
  scfg_ = cfg;
  if (scfg_.hbm_budget_bytes == 0) scfg_.hbm_budget_bytes = 64ull << 30;
  if (scfg_.max_live_jobs == 0 || scfg_.max_live_jobs > kJobRing) scfg_.max_live_jobs = 512;
  {
    std::lock_guard<std::mutex> g(mu);
    srv_record_bytes_ = std::min<uint64_t>(scfg_.hbm_budget_bytes / 4, (uint64_t)tuning.record_mb << 20);
  }
  if (scfg_.pool_paths) {
    if (scfg_.pool_paths < 64 || scfg_.pool_paths > (1u << 26)) return fail(OMR_ERR_INVALID, "pool_paths out of range");
    std::lock_guard<std::mutex> g(mu);
    drop_graphs();
    tuning.pool_paths = (uint32_t)scfg_.pool_paths;
  }
  {
    std::lock_guard<std::mutex> g(smu);
    queue_.clear();
    all_.clear();
    used_bytes_ = 0;
    n_submitted_ = n_done_ = n_rejected_ = n_cancelled_ = 0;
    srv_segments_ = srv_shadow_ = 0;
    busy_ms_ = 0;
    latencies_.clear();
  }
  // Request buffers come from the device's default pool (stream-ordered);
  // keep freed blocks in the pool while serving instead of handing them back
  // to the driver at every synchronization, so admissions reuse mapped memory.
  {
    std::lock_guard<std::mutex> g(mu);
    hipSetDevice(device_);
    hipMemPool_t pool = nullptr;
    uint64_t keep = UINT64_MAX;
    if (hipDeviceGetDefaultMemPool(&pool, device_) == hipSuccess && pool)
      hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  server_stop_ = false;
  server_on_ = true;
  server_ = std::thread([this] { server_loop(); });
  return OMR_OK;
}

int Engine::server_stop() {
  if (!server_on_.load()) return OMR_OK;
  server_stop_ = true;
  scv.notify_all();
  if (server_.joinable()) server_.join();
  server_on_ = false;
  {
    std::lock_guard<std::mutex> g(mu);
    hipSetDevice(device_);
    hipMemPool_t pool = nullptr;
    uint64_t keep = 0;
    if (hipDeviceGetDefaultMemPool(&pool, device_) == hipSuccess && pool) {
      hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      hipMemPoolTrimTo(pool, 0);
    }
  }
  return OMR_OK;
}

int Engine::submit(omr_scene_id scene_id, const omr_camera &cam, const omr_render_params &p, float *out,
                   omr_job_id *id) {
  if (!server_on_.load()) return OMR_ERR_STATE;
  SceneHost *s = scene(scene_id);
  if (!s) return OMR_ERR_NOTFOUND;
  auto r = std::make_shared<Req>();
  r->scene = scene_id;
  r->cam = cam;
  r->params = p;
  r->out = out;
  r->t_submit = std::chrono::steady_clock::now();
  size_t px = (size_t)cam.width * cam.height;
  r->bytes = px * 3 * 8 + px * 3 * 4;
  // per-sample records (no atomics) for requests whose records take at most a
  // quarter of the budget and fit in record_mb; larger ones use pixel atomics
  const size_t rec = (size_t)samples_of(cam, p) * 24;
  r->records = rec <= srv_record_bytes_;
  if (r->records) r->bytes += rec;
  {
    std::lock_guard<std::mutex> g(smu);
    all_.push_back(r);
    r->id = (omr_job_id)all_.size();
    n_submitted_++;
    if (r->bytes > scfg_.hbm_budget_bytes) {
      r->state = 4;
      n_rejected_++;
      *id = r->id;
      scv.notify_all();
      return OMR_ERR_BUDGET;
    }
    s->users++;
    queue_.push_back(r);
    *id = r->id;
  }
  scv.notify_all();
  return OMR_OK;
}

int Engine::poll(omr_job_id id, int *state) {
  std::lock_guard<std::mutex> g(smu);
  if (id < 1 || (size_t)id > all_.size()) return OMR_ERR_NOTFOUND;
  *state = all_[(size_t)id - 1]->state;
  return OMR_OK;
}

int Engine::wait(omr_job_id id, int64_t timeout_ms) {
  std::unique_lock<std::mutex> g(smu);
  if (id < 1 || (size_t)id > all_.size()) return OMR_ERR_NOTFOUND;
  auto r = all_[(size_t)id - 1];
  auto done = [&] { return r->state >= 2; };
  if (timeout_ms < 0) scv.wait(g, done);
  else if (!scv.wait_for(g, std::chrono::milliseconds(timeout_ms), done)) return OMR_ERR_TIMEOUT;
  if (r->state == 3) return OMR_ERR_CANCELLED;
  if (r->state == 4) return OMR_ERR_BUDGET;
  return OMR_OK;
}

int Engine::cancel(omr_job_id id) {
  std::lock_guard<std::mutex> g(smu);
  if (id < 1 || (size_t)id > all_.size()) return OMR_ERR_NOTFOUND;
  all_[(size_t)id - 1]->cancel_req = true;
  scv.notify_all();
  return OMR_OK;
}

int Engine::server_stats(omr_server_stats *s) {
  std::lock_guard<std::mutex> g(smu);
  std::memset(s, 0, sizeof(*s));
  s->jobs_submitted = n_submitted_;
  s->jobs_done = n_done_;
  s->jobs_rejected = n_rejected_;
  s->jobs_cancelled = n_cancelled_;
  s->segments = srv_segments_;
  s->shadow_rays = srv_shadow_;
  s->busy_ms = busy_ms_;
  if (!latencies_.empty()) {
    std::vector<double> l = latencies_;
    std::sort(l.begin(), l.end());
    auto pct = [&](double q) {
      size_t i = (size_t)std::min<double>((double)l.size() - 1, q * (double)(l.size() - 1) + 0.5);
      return l[i];
    };
    s->lat_p50_ms = pct(0.50);
    s->lat_p99_ms = pct(0.99);
    s->lat_max_ms = l.back();
  }
  return OMR_OK;
}

void Engine::server_loop() {
  hipSetDevice(device_);
  {
    std::lock_guard<std::mutex> g(mu);
    solo_ = kNoSolo;  // requests share the run
  }
  std::deque<std::shared_ptr<Req>> live;
  uint64_t lo = 0, hi = 0;          // logical ring window [lo, hi)
  unsigned long long total = 0;     // samples admitted
  unsigned long long cursor[kShards] = {};  // device sample cursors as of the last readback
  JobDev *h_jobs = nullptr;         // pinned staging for job records
  struct Stage {  // same layout as Ctrl::next_sample, total_samples (one copy updates both)
    Sharded<unsigned long long> cursor;
    alignas(kLine) unsigned long long total;
    uint32_t lo, hi;
  } *h_stage = nullptr;
  static_assert(offsetof(Stage, total) - offsetof(Stage, cursor) ==
                    offsetof(Ctrl, total_samples) - offsetof(Ctrl, next_sample),
                "staged cursors must mirror the control block");
  static_assert(offsetof(Ctrl, job_hi) == offsetof(Ctrl, job_lo) + 4 && offsetof(Stage, hi) == offsetof(Stage, lo) + 4,
                "the job window is copied as one 8-byte pair");
  bool ok = true;
  hipStream_t cs = nullptr;  // result copies, beside the compute stream
  {
    std::lock_guard<std::mutex> g(mu);
    if (ensure_pool() != OMR_OK) ok = false;
    if (ok && hipStreamCreate(&cs) != hipSuccess) ok = false;
    if (ok && hipHostMalloc(&h_jobs, kJobRing * sizeof(JobDev), hipHostMallocDefault) != hipSuccess) ok = false;
    if (ok && hipHostMalloc(&h_stage, sizeof(Stage), hipHostMallocDefault) != hipSuccess) ok = false;
    if (ok) {
      Ctrl c;
      std::memset(&c, 0, sizeof(c));
      ok = hipMemcpy(d_ctrl_, &c, sizeof(Ctrl), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemset(d_done_, 0, kJobRing * kShards * sizeof(unsigned long long)) == hipSuccess &&
           hipMemset(pool_.misc, 0, (size_t)pool_cap_ * 16) == hipSuccess;
    }
  }
  auto fail_all = [&](const std::string &why) {
    std::lock_guard<std::mutex> g(smu);
    err = why;
    for (auto &r : live) r->state = 4;
    for (auto &r : queue_) r->state = 4;
    queue_.clear();
    scv.notify_all();
  };
  if (!ok) {
    fail_all("server start: device allocation failed");
    return;
  }
  // Requests whose image is resolved on the compute stream but not yet copied
  // out.  Their copies run on `cs` while the compute stream works on the next
  // chunk of steps (the copy to a caller's pageable buffer holds this thread,
  // not the GPU); then their device buffers go back and they complete.
  std::vector<std::shared_ptr<Req>> copying;
  auto flush_copies = [&]() {
    if (copying.empty()) return;
    bool copied = true;
    for (auto &r : copying) {
      copied = copied && hipStreamWaitEvent(cs, r->ev_copy, 0) == hipSuccess &&
               hipMemcpyAsync(r->out, r->run.d_out, (size_t)r->cam.width * r->cam.height * 3 * 4,
                              hipMemcpyDeviceToHost, cs) == hipSuccess;
    }
    copied = hipStreamSynchronize(cs) == hipSuccess && copied;
    for (auto &r : copying) {
      hipEventDestroy(r->ev_copy);
      r->ev_copy = nullptr;
      hipFreeAsync(r->run.d_accum, stream_);
      hipFreeAsync(r->run.d_out, stream_);
      if (r->run.d_recs) hipFreeAsync(r->run.d_recs, stream_);
    }
    std::lock_guard<std::mutex> g(smu);
    for (auto &r : copying) {
      used_bytes_ -= r->bytes;
      if (SceneHost *s = scene(r->scene)) s->users--;
      if (!copied) {
        r->state = 4;  // its image did not reach the caller
        err = "result copy failed";
      } else if (r->cancel_req) {
        r->state = 3;
        n_cancelled_++;
      } else {
        r->state = 2;
        n_done_++;
        latencies_.push_back(ms_since(r->t_submit));
      }
    }
    copying.clear();
    scv.notify_all();
  };
  auto busy_t0 = std::chrono::steady_clock::now();
  while (!server_stop_.load()) {
    // ---- 1. admission
    std::vector<std::shared_ptr<Req>> admit;
    {
      std::unique_lock<std::mutex> g(smu);
      if (live.empty() && queue_.empty()) {
        scv.wait_for(g, std::chrono::milliseconds(20));
        if (queue_.empty()) continue;
      }
      while (!queue_.empty() && (hi - lo) < scfg_.max_live_jobs) {
        auto r = queue_.front();
        if (r->cancel_req) {
          queue_.pop_front();
          r->state = 3;
          n_cancelled_++;
          if (SceneHost *s = scene(r->scene)) s->users--;
          scv.notify_all();
          continue;
        }
        if (used_bytes_ + r->bytes > scfg_.hbm_budget_bytes) break;  // wait for memory
        queue_.pop_front();
        used_bytes_ += r->bytes;
        r->state = 1;
        admit.push_back(r);
        hi++;
      }
    }
    std::unique_lock<std::mutex> dev(mu);
    if (!admit.empty()) {
      const unsigned long long total_before = total;
      if (live.empty()) busy_t0 = std::chrono::steady_clock::now();
      uint32_t need_stack = 1;
      for (auto &r : admit) {
        SceneHost *s = scene(r->scene);
        need_stack = std::max(need_stack, s ? s->bvh.stack_bound : 1u);
        JobRun &j = r->run;
        j.scene = (uint32_t)r->scene;
        j.cam = r->cam;
        j.params = r->params;
        j.n_samples = (uint32_t)samples_of(r->cam, r->params);
        j.logical = hi - admit.size() + (&r - &admit[0]);
        j.ring = (uint32_t)(j.logical % kJobRing);
        size_t px = (size_t)r->cam.width * r->cam.height;
        j.d_recs = nullptr;
        if (hipMallocAsync((void **)&j.d_accum, px * 3 * 8, stream_) != hipSuccess ||
            hipMallocAsync((void **)&j.d_out, px * 3 * 4, stream_) != hipSuccess ||
            (r->records && hipMallocAsync((void **)&j.d_recs, (size_t)j.n_samples * 24, stream_) != hipSuccess)) {
          dev.unlock();
          fail_all("device allocation for a job failed");
          return;
        }
        hipMemsetAsync(j.d_accum, 0, px * 3 * 8, stream_);
        if (j.d_recs) hipMemsetAsync(j.d_recs, 0, (size_t)j.n_samples * 24, stream_);
        k::clear_done(d_done_, j.ring, stream_);
        JobDev &D = h_jobs[j.ring];
        std::memset(&D, 0, sizeof(D));
        D.sample_base = total;
        D.n_samples = j.n_samples;
        D.width = j.cam.width;
        D.height = j.cam.height;
        D.ss = j.params.supersamples;
        D.nspp = j.params.spp_end - j.params.spp_begin;
        D.spp_begin = j.params.spp_begin;
        D.max_depth = j.params.max_depth;
        D.rr_depth = j.params.rr_depth;
        D.flags = j.params.env_is ? 1u : 0u;
        D.scene = j.scene;
        D.seed = j.params.seed;
        for (int c = 0; c < 3; c++) {
          D.cam_pos[c] = j.cam.position[c];
          D.cam_c[c] = j.cam.screen_center[c];
          D.cam_x[c] = j.cam.screen_x[c];
          D.cam_y[c] = j.cam.screen_y[c];
        }
        D.aperture = j.cam.aperture_radius;
        D.focal = j.cam.focal_distance;
        D.accum = j.d_accum;
        D.recs = j.d_recs;
        hipMemcpyAsync(d_jobs_ + j.ring, &D, sizeof(JobDev), hipMemcpyHostToDevice, stream_);
        total += j.n_samples;
        live.push_back(r);
      }
      LaunchShape sh = shape();
      if (ensure_gstack(need_stack, (size_t)sh.grid * sh.wg) != OMR_OK) {
        dev.unlock();
        fail_all("traversal stack allocation failed");
        return;
      }
      // the cursor may have run past the old total once work ran out: new
      // samples start exactly at the old total
      for (uint32_t k = 0; k < kShards; k++) h_stage->cursor[k] = std::min(cursor[k], shard_limit(total_before, k));
      h_stage->total = total;
      h_stage->lo = (uint32_t)lo;
      h_stage->hi = (uint32_t)hi;
      hipMemcpyAsync(&d_ctrl_->next_sample, &h_stage->cursor, offsetof(Stage, total) + 8, hipMemcpyHostToDevice,
                     stream_);
      hipMemcpyAsync(&d_ctrl_->job_lo, &h_stage->lo, 8, hipMemcpyHostToDevice, stream_);
      k::refill(args(), sh, 0, stream_);
    }
    if (live.empty()) continue;
    // ---- 2. one chunk of bounce steps
    bool env_is = false;
    for (auto &r : live) {
      SceneHost *s = scene(r->scene);
      if (r->params.env_is && s && s->has_env) env_is = true;
    }
    if (launch_steps(tuning.steps_per_sync, env_is) != OMR_OK) {
      dev.unlock();
      fail_all(err);
      return;
    }
    flush_copies();
    hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_);
    hipMemcpyAsync(h_done_, d_done_, kJobRing * kShards * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_);
    if (hipStreamSynchronize(stream_) != hipSuccess) {
      dev.unlock();
      fail_all("device error in the bounce loop");
      return;
    }
    for (uint32_t k = 0; k < kShards; k++) cursor[k] = h_ctrl_->next_sample[k];
    // ---- 2b. drain: every admitted sample is handed out and few paths are
    // left (no request could be admitted above): finish them in one launch
    // instead of many short steps, then start the next run on empty queues.
    const uint64_t queued = queue_len(*h_ctrl_, 0);
    if (tuning.drain_below && queued > 0 && queued <= tuning.drain_below && samples_exhausted(*h_ctrl_)) {
      bool ok2 = k::drain(args(), shape(), 0, false, stream_) == hipSuccess &&
                 hipMemsetAsync(&d_ctrl_->qcount[0], 0, sizeof(d_ctrl_->qcount[0]), stream_) == hipSuccess &&
                 hipMemsetAsync(&d_ctrl_->qhead[0], 0, sizeof(d_ctrl_->qhead[0]), stream_) == hipSuccess;
      hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_);
      hipMemcpyAsync(h_done_, d_done_, kJobRing * kShards * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                     stream_);
      if (!ok2 || hipStreamSynchronize(stream_) != hipSuccess) {
        dev.unlock();
        fail_all("device error in the drain");
        return;
      }
    }
    // ---- 3. completions
    std::vector<std::shared_ptr<Req>> finished;
    for (auto &r : live) {
      JobRun &j = r->run;
      if (j.done || job_done(h_done_, j.ring) != j.n_samples) continue;
      j.done = true;
      double nspp = (double)j.params.supersamples * j.params.supersamples *
                    (double)(j.params.spp_end - j.params.spp_begin);
      if (j.d_recs) k::gather_records(j.d_recs, (uint32_t)nspp, j.cam.width * j.cam.height, j.d_accum, stream_);
      k::resolve(j.d_accum, j.d_out, j.cam.width * j.cam.height, nspp, stream_);
      if (hipEventCreateWithFlags(&r->ev_copy, hipEventDisableTiming) != hipSuccess ||
          hipEventRecord(r->ev_copy, stream_) != hipSuccess) {
        dev.unlock();
        fail_all("result event");
        return;
      }
      copying.push_back(r);
      finished.push_back(r);
    }
    if (!finished.empty()) {
      while (!live.empty() && live.front()->run.done) {
        live.pop_front();
        lo++;
      }
      h_stage->lo = (uint32_t)lo;
      h_stage->hi = (uint32_t)hi;
      hipMemcpyAsync(&d_ctrl_->job_lo, &h_stage->lo, 8, hipMemcpyHostToDevice, stream_);
    }
    // nothing left to overlap with: copy out now
    if (live.empty()) flush_copies();
    dev.unlock();
    {
      std::lock_guard<std::mutex> g(smu);
      srv_segments_ = sum_shards(h_ctrl_->segments);
      srv_shadow_ = sum_shards(h_ctrl_->shadow_rays);
      if (live.empty()) busy_ms_ += ms_since(busy_t0);
    }
  }
  {
    std::lock_guard<std::mutex> g(mu);
    hipStreamSynchronize(stream_);
    flush_copies();
    hipStreamSynchronize(stream_);
    hipStreamDestroy(cs);
    for (auto &r : live) {
      hipFree(r->run.d_accum);
      if (r->run.d_recs) hipFree(r->run.d_recs);
      hipFree(r->run.d_out);
    }
    hipHostFree(h_jobs);
    hipHostFree(h_stage);
  }
  std::lock_guard<std::mutex> g(smu);
  for (auto &r : live) r->state = 3;
  for (auto &r : queue_) r->state = 3;
  queue_.clear();
  scv.notify_all();
}

}  // namespace omr
