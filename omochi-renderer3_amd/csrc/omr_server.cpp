// omr_server.cpp -- served mode: continuous batching on the shared queues.
//
// A server thread owns the device stream.  It keeps two chunks of bounce
// steps in flight: after it enqueues chunk N+1 it reads the status of chunk N
// (one small kernel writes the queue length, cursor state and the live jobs'
// completion counts, one copy brings them back), so the device never waits
// for the host.  From each status it
// (1) admits queued requests while the job ring and the HBM budget allow:
//     their samples are appended to the global sample range on the device
//     (k::admit, stream-ordered after the chunks in flight) and idle path
//     slots take them (refill), so new work overlaps with work in flight;
// (2) once every admitted sample is handed out and few paths remain, ends
//     them in one drain launch instead of many short steps;
// (3) resolves finished jobs and copies them to the caller's buffer on a
//     second stream while the next chunks run, then frees their memory.
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
}  // namespace

bool Engine::arena_take(uint64_t bytes, uint64_t *off) {
  std::lock_guard<std::mutex> g(arena_mu_);
  return arena_map_.take(bytes, off);
}

void Engine::arena_give(uint64_t off, uint64_t bytes) {
  std::lock_guard<std::mutex> g(arena_mu_);
  arena_map_.give(off, bytes);
}

int Engine::server_start(const omr_server_config &cfg) {
  if (server_on_.load()) return fail(OMR_ERR_STATE, "server already running");
  omr_server_config c = cfg;
  if (c.hbm_budget_bytes == 0) c.hbm_budget_bytes = 64ull << 30;
  if (c.max_live_jobs == 0 || c.max_live_jobs > kJobRing) c.max_live_jobs = 512;
  if (c.pool_paths && (c.pool_paths < 64 || c.pool_paths > (1u << 26)))
    return fail(OMR_ERR_INVALID, "pool_paths out of range");
  {
    std::lock_guard<std::mutex> g(mu);
    hipSetDevice(device_);
    saved_pool_paths_ = tuning.pool_paths;
    if (c.pool_paths && c.pool_paths != tuning.pool_paths) {
      drop_graphs();
      tuning.pool_paths = (uint32_t)c.pool_paths;
    }
    // the served loop runs one lane: the second pool of two-lane runs is
    // not kept while serving
    hipStreamSynchronize(stream_);
    swap_lane();
    hipStreamSynchronize(stream_);
    drop_graphs();
    free_lane_pool();
    if (d_gstack_) hipFree(d_gstack_);
    d_gstack_ = nullptr;
    gcap_ = 0;
    gstack_threads_ = 0;
    swap_lane();
    // the pool is part of the budget: allocate it now so its size is known
    int rc = ensure_pool();
    if (rc) {
      tuning.pool_paths = saved_pool_paths_;
      return rc;
    }
    srv_record_bytes_ = std::min<uint64_t>(c.hbm_budget_bytes / 4, (uint64_t)tuning.record_mb << 20);
    fixed_bytes_ = pool_bytes();
  }
  const uint64_t fixed = fixed_bytes_.load() + scene_bytes();
  if (fixed >= c.hbm_budget_bytes) {
    std::lock_guard<std::mutex> g(mu);
    if (tuning.pool_paths != saved_pool_paths_) {
      drop_graphs();
      tuning.pool_paths = saved_pool_paths_;
    }
    return fail(OMR_ERR_BUDGET, "the path pool and resident scenes (" + std::to_string(fixed) +
                                    " bytes) leave nothing of the HBM budget (" +
                                    std::to_string(c.hbm_budget_bytes) + " bytes)");
  }
  scfg_ = c;
  {
    std::lock_guard<std::mutex> g(smu);
    queue_.clear();
    all_.clear();
    used_bytes_ = 0;
    n_submitted_ = n_done_ = n_rejected_ = n_cancelled_ = 0;
    srv_segments_ = srv_shadow_ = 0;
    busy_ms_ = 0;
    latencies_.clear();
    server_failed_ = false;
  }
  // Request buffers: one arena of what the pool and scenes leave of the
  // budget, allocated now; serving itself allocates nothing
  {
    std::lock_guard<std::mutex> g(mu);
    const uint64_t room = (c.hbm_budget_bytes - fixed) & ~uint64_t(255);
    if (room == 0 || hipMalloc((void **)&arena_, room) != hipSuccess) {
      (void)hipGetLastError();
      arena_ = nullptr;
      if (tuning.pool_paths != saved_pool_paths_) {
        drop_graphs();
        tuning.pool_paths = saved_pool_paths_;
      }
      return fail(OMR_ERR_NOMEM, "request arena of " + std::to_string(room) + " bytes: device allocation failed");
    }
    arena_bytes_ = room;
    std::lock_guard<std::mutex> ag(arena_mu_);
    arena_map_.reset(room);
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
    // the pool size a server_start override chose does not outlive the server
    if (tuning.pool_paths != saved_pool_paths_) {
      drop_graphs();
      tuning.pool_paths = saved_pool_paths_;
    }
    if (arena_) hipFree(arena_);
    arena_ = nullptr;
    arena_bytes_ = 0;
  }
  return OMR_OK;
}

int Engine::submit(omr_scene_id scene_id, const omr_camera &cam, const omr_render_params &p, float *out,
                   omr_job_id *id) {
  if (!server_on_.load()) return OMR_ERR_STATE;
  auto r = std::make_shared<Req>();
  {
    // lookup and hold under the lock release_scene checks the holders under
    std::lock_guard<std::mutex> g(scene_mu);
    SceneHost *s = scene(scene_id);
    if (!s) return OMR_ERR_NOTFOUND;
    if (p.nee && s->emissive_specular) return OMR_ERR_INVALID;  // as run_jobs refuses it
    s->users++;
    r->sh = s;
  }
  r->scene = scene_id;
  r->cam = cam;
  r->params = p;
  r->out = out;
  r->t_submit = std::chrono::steady_clock::now();
  // The budget covers the pool, the resident scenes and the admitted
  // requests.  A request needs its sums and image; per-sample records (no
  // atomics, the same results) are added when they fit in the record limit
  // and, with the image, in what the pool and scenes leave of the budget.
  const uint64_t room = arena_bytes_;
  const uint64_t need = image_bytes(cam.width, cam.height);
  const bool fits = plan_request(need, samples_of(cam, p), srv_record_bytes_, room, &r->place);
  {
    std::lock_guard<std::mutex> g(smu);
    all_.push_back(r);
    r->id = (omr_job_id)all_.size();
    *id = r->id;
    n_submitted_++;
    int code = OMR_OK;
    if (server_failed_) {
      code = OMR_ERR_STATE;
      r->msg = "the server stopped after an error: " + srv_err_;
      r->state = 4;
    } else if (!fits) {
      code = OMR_ERR_BUDGET;
      r->msg = "the request needs " + std::to_string(need) + " bytes; the path pool and resident scenes leave " +
               std::to_string(room) + " bytes of the HBM budget for requests";
      r->state = 5;
      n_rejected_++;
    }
    if (code != OMR_OK) {
      r->code = code;
      r->sh->users--;
      r->sh = nullptr;
      scv.notify_all();
      return code;
    }
    queue_.push_back(r);
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
  if (r->state >= 4) return r->code;
  return OMR_OK;
}

int Engine::job_error(omr_job_id id, std::string *msg) {
  std::lock_guard<std::mutex> g(smu);
  if (id < 1 || (size_t)id > all_.size()) return OMR_ERR_NOTFOUND;
  *msg = all_[(size_t)id - 1]->msg;
  return OMR_OK;
}

int Engine::job_latency(omr_job_id id, double *ms) {
  std::lock_guard<std::mutex> g(smu);
  if (id < 1 || (size_t)id > all_.size()) return OMR_ERR_NOTFOUND;
  const auto &r = all_[(size_t)id - 1];
  if (r->state != 2) return OMR_ERR_STATE;
  *ms = r->lat_ms;
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
  const uint64_t fixed = fixed_bytes_.load() + scene_bytes();
  std::lock_guard<std::mutex> g(smu);
  std::memset(s, 0, sizeof(*s));
  s->fixed_bytes = fixed;
  s->used_bytes = used_bytes_;
  s->jobs_submitted = n_submitted_;
  s->jobs_done = n_done_;
  s->jobs_rejected = n_rejected_;
  s->jobs_cancelled = n_cancelled_;
  s->segments = srv_segments_;
  s->shadow_rays = srv_shadow_;
  s->busy_ms = busy_ms_;
  s->lat_p50_ms = latencies_.percentile(0.50);
  s->lat_p99_ms = latencies_.percentile(0.99);
  s->lat_max_ms = latencies_.max();
  return OMR_OK;
}

void Engine::server_loop() {
  hipSetDevice(device_);
  {
    std::lock_guard<std::mutex> g(mu);
    solo_ = kNoSolo;  // requests share the run
    any_env_ = true;
  }
  std::deque<std::shared_ptr<Req>> live;  // admitted, by logical ring index
  uint64_t lo = 0, hi = 0;                // logical ring window [lo, hi)
  unsigned long long total = 0;           // samples admitted
  unsigned long long placed = 0;          // samples placed in rounds (drain rounds)
  // rounds are placed by the host when a full pool's fill fits the sub-queues
  // (the pool is fixed while serving); otherwise the refill hands out samples
  bool place_ok = true;
  for (uint32_t k = 0; k < kShards; k++)
    if (shard_limit(pool_cap_, k) > pool_.qcap) place_ok = false;
  JobDev *h_jobs = nullptr;               // pinned staging for job records
  // Status slots: two chunks' statuses and a drain's can be in flight at
  // once; a slot is reused only kStatusSlots enqueues later.
  constexpr int kStatusSlots = 4;
  ServeStatus *d_st = nullptr;            // status slots on the device ...
  ServeStatus *h_st = nullptr;            // ... and their pinned copies
  hipEvent_t ev_st[kStatusSlots] = {};
  hipStream_t cs = nullptr;  // result copies, beside the compute stream
  bool ok = true;
  {
    std::lock_guard<std::mutex> g(mu);
    ok = hipStreamCreate(&cs) == hipSuccess &&
         hipHostMalloc(&h_jobs, kJobRing * sizeof(JobDev), hipHostMallocDefault) == hipSuccess &&
         hipMalloc(&d_st, kStatusSlots * sizeof(ServeStatus)) == hipSuccess &&
         hipHostMalloc(&h_st, kStatusSlots * sizeof(ServeStatus), hipHostMallocDefault) == hipSuccess;
    for (int i = 0; ok && i < kStatusSlots; i++)
      ok = hipEventCreateWithFlags(&ev_st[i], hipEventDisableTiming) == hipSuccess;
    if (ok) {
      Ctrl c;
      std::memset(&c, 0, sizeof(c));
      ok = hipMemcpy(d_ctrl_, &c, sizeof(Ctrl), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemset(d_done_, 0, kJobRing * kShards * sizeof(unsigned long long)) == hipSuccess &&
           hipMemset(pool_.misc, 0, (size_t)pool_cap_ * 16) == hipSuccess;
    }
  }
  // Marks every unfinished live request and every queued one failed (the
  // server stops serving).  A live request whose samples are all done is the
  // copier's (queued for its copy, or already completed and waiting in the
  // window behind an unfinished one): its outcome is the copy's.
  auto fail_all = [&](int code, const std::string &why) {
    set_err(why);
    std::lock_guard<std::mutex> g(smu);
    srv_err_ = why;
    server_failed_ = true;
    auto fail_one = [&](const std::shared_ptr<Req> &r) {
      if (r->run.done) return;
      r->state = 4;
      r->code = code;
      r->msg = why;
      if (r->sh) r->sh->users--;
      r->sh = nullptr;
    };
    for (auto &r : live) fail_one(r);
    for (auto &r : queue_) fail_one(r);
    queue_.clear();
    scv.notify_all();
  };
  auto release_all = [&]() {
    std::lock_guard<std::mutex> g(mu);
    hipStreamSynchronize(stream_);
    if (cs) {
      hipStreamSynchronize(cs);
      hipStreamDestroy(cs);
    }
    for (auto &r : live) {
      if (r->run.d_accum) arena_give(r->place.off, r->place.bytes());  // not yet returned by the copier
      r->run.d_accum = r->run.d_recs = nullptr;
      r->run.d_out = nullptr;
      if (r->ev_copy) hipEventDestroy(r->ev_copy);
    }
    for (hipEvent_t e : ev_st)
      if (e) hipEventDestroy(e);
    if (d_st) hipFree(d_st);
    if (h_st) hipHostFree(h_st);
    if (h_jobs) hipHostFree(h_jobs);
  };
  if (!ok) {
    fail_all(OMR_ERR_NOMEM, "server start: device allocation failed");
    release_all();
    return;
  }
  // Result copies run on their own host thread and stream: for each finished
  // request the copier waits for its resolve (an event on the compute
  // stream), copies the image into the caller's (pageable) buffer, returns
  // its device buffers and completes it.  The copy to pageable memory holds
  // the copier, never the server thread, so the chunks of steps keep coming.
  std::deque<std::shared_ptr<Req>> to_copy;
  std::mutex cmu;
  std::condition_variable ccv;
  bool copier_stop = false;
  std::thread copier([&]() {
    hipSetDevice(device_);
    for (;;) {
      std::shared_ptr<Req> r;
      {
        std::unique_lock<std::mutex> g(cmu);
        ccv.wait(g, [&] { return copier_stop || !to_copy.empty(); });
        if (to_copy.empty()) return;  // stopped and drained
        r = to_copy.front();
        to_copy.pop_front();
      }
      bool copied = hipStreamWaitEvent(cs, r->ev_copy, 0) == hipSuccess &&
                    hipMemcpyAsync(r->out, r->run.d_out, (size_t)r->cam.width * r->cam.height * 3 * 4,
                                   hipMemcpyDeviceToHost, cs) == hipSuccess;
      copied = hipStreamSynchronize(cs) == hipSuccess && copied;
      hipEventDestroy(r->ev_copy);
      r->ev_copy = nullptr;
      // the copy is complete and so is all device work of the request (the
      // copy waited for its resolve): its block goes back to the arena
      arena_give(r->place.off, r->place.bytes());
      r->run.d_accum = r->run.d_recs = nullptr;
      r->run.d_out = nullptr;
      std::lock_guard<std::mutex> g(smu);
      used_bytes_ -= r->place.bytes();
      if (r->sh) r->sh->users--;
      r->sh = nullptr;
      if (!copied) {
        r->state = 4;  // its image did not reach the caller
        r->code = OMR_ERR_HIP;
        r->msg = "result copy to the caller's buffer failed";
      } else if (r->cancel_req) {
        r->state = 3;
        n_cancelled_++;
      } else {
        r->state = 2;
        n_done_++;
        r->lat_ms = ms_since(r->t_submit);
        latencies_.add(r->lat_ms);
      }
      scv.notify_all();
    }
  });
  auto stop_copier = [&]() {
    {
      std::lock_guard<std::mutex> g(cmu);
      copier_stop = true;
    }
    ccv.notify_all();
    if (copier.joinable()) copier.join();
  };
  // A device or HIP failure: every live and queued request fails with the
  // code, the copies already queued finish, then everything is released.
  auto abort_with = [&](int code, const std::string &why) {
    fail_all(code, why);
    stop_copier();
    release_all();
  };
  // Status reads in flight, oldest first: the status slot, the job window
  // and the admitted total when it was enqueued, and its sequence number.
  struct Pending {
    int slot;
    uint64_t lo;
    uint32_t n;
    unsigned long long total;
    uint64_t seq;
  };
  std::deque<Pending> pend;
  uint64_t seq = 0, drain_seq = 0;  // statuses enqueued; the first status after the last drain
  int slot_next = 0;
  auto enqueue_status = [&]() -> bool {
    const int s = slot_next;
    slot_next = (slot_next + 1) % kStatusSlots;
    const uint32_t n = (uint32_t)(hi - lo);
    const bool okq = k::serve_status(d_ctrl_, d_done_, (uint32_t)(lo % kJobRing), n, d_st + s, stream_) == hipSuccess &&
                     hipMemcpyAsync(h_st + s, d_st + s, offsetof(ServeStatus, done) + 8ull * n,
                                    hipMemcpyDeviceToHost, stream_) == hipSuccess &&
                     hipEventRecord(ev_st[s], stream_) == hipSuccess;
    pend.push_back(Pending{s, lo, n, total, seq++});
    return okq;
  };
  auto busy_t0 = std::chrono::steady_clock::now();
  while (!server_stop_.load()) {
    // ---- 1. admission
    std::vector<std::shared_ptr<Req>> admit;
    {
      std::unique_lock<std::mutex> g(smu);
      if (live.empty() && queue_.empty() && pend.empty()) {
        scv.wait_for(g, std::chrono::milliseconds(20));
        if (queue_.empty()) continue;
      }
      // in submission order while the job ring and the arena allow (a block
      // without the per-sample records when only that fits: pixel atomics,
      // the same results); cancelled requests leave without one
      std::lock_guard<std::mutex> ag(arena_mu_);
      admit_in_order(
          queue_, arena_map_, scfg_.max_live_jobs - (size_t)(hi - lo),
          [](const std::shared_ptr<Req> &r) { return r->cancel_req; },
          [&](const std::shared_ptr<Req> &r) {
            r->state = 3;
            n_cancelled_++;
            if (r->sh) r->sh->users--;
            r->sh = nullptr;
            scv.notify_all();
          },
          [](const std::shared_ptr<Req> &r) { return &r->place; },
          [&](const std::shared_ptr<Req> &r) {
            used_bytes_ += r->place.bytes();
            r->state = 1;
            admit.push_back(r);
          });
    }
    std::unique_lock<std::mutex> dev(mu);
    if (!admit.empty()) {
      const unsigned long long total_before = total;
      if (live.empty()) busy_t0 = std::chrono::steady_clock::now();
      uint32_t need_stack = 1;
      bool okd = true;
      for (auto &r : admit) {
        need_stack = std::max(need_stack, stack_need(r->sh));
        JobRun &j = r->run;
        j.scene = (uint32_t)r->scene;
        j.cam = r->cam;
        j.params = r->params;
        j.n_samples = (uint32_t)samples_of(r->cam, r->params);
        j.logical = hi++;
        j.ring = (uint32_t)(j.logical % kJobRing);
        ring_env_[j.ring] = r->sh->has_env ? 1u : 0u;
        const size_t px = (size_t)r->cam.width * r->cam.height;
        j.d_recs = nullptr;
        live.push_back(r);
        char *blk = arena_ + r->place.off;
        j.d_accum = reinterpret_cast<unsigned long long *>(blk);
        j.d_out = reinterpret_cast<float *>(blk + align256(px * 3 * 8));
        if (r->place.records) j.d_recs = reinterpret_cast<unsigned long long *>(blk + r->place.image);
        // with per-sample records the gather writes every pixel's sums
        okd = okd &&
              (j.d_recs ? hipMemsetAsync(j.d_recs, 0, (size_t)j.n_samples * 24, stream_)
                        : hipMemsetAsync(j.d_accum, 0, px * 3 * 8, stream_)) == hipSuccess &&
              k::clear_done(d_done_, j.ring, stream_) == hipSuccess;
        JobDev &D = h_jobs[j.ring];
        std::memset(&D, 0, sizeof(D));
        D.sample_base = total;
        j.first = total;
        D.n_samples = j.n_samples;
        D.width = j.cam.width;
        D.height = j.cam.height;
        D.ss = j.params.supersamples;
        D.nspp = j.params.spp_end - j.params.spp_begin;
        D.spp_begin = j.params.spp_begin;
        D.max_depth = j.params.max_depth;
        D.rr_depth = j.params.rr_depth;
        D.flags = (j.params.env_is ? 1u : 0u) | (j.params.nee ? 2u : 0u);
        D.scene = scene_slot((omr_scene_id)j.scene);  // the device scene table is indexed by slot
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
        okd = okd && hipMemcpyAsync(d_jobs_ + j.ring, &D, sizeof(JobDev), hipMemcpyHostToDevice, stream_) == hipSuccess;
        total += j.n_samples;
        if (!okd) break;
      }
      if (!okd) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "request set-up failed");
        return;
      }
      LaunchShape sh = shape();
      if (ensure_gstack(need_stack, (size_t)sh.grid * sh.wg) != OMR_OK) {
        dev.unlock();
        abort_with(OMR_ERR_NOMEM, "traversal stack allocation failed");
        return;
      }
      fixed_bytes_ = pool_bytes();
      // (in drain rounds the next round's refill hands the new samples out)
      if (k::admit(d_ctrl_, total_before, total, (uint32_t)lo, (uint32_t)hi, stream_) != hipSuccess ||
          (!tuning.drain_rounds && k::refill(args(), sh, 0, stream_) != hipSuccess)) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "admission launch failed");
        return;
      }
    }
    // ---- 2. keep two chunks of work (and their statuses) in flight: a drain
    // round (every idle slot takes a sample, all run to their end) or a chunk
    // of bounce steps
    if (!live.empty()) {
      bool env_is = false;
      // (requests that finished behind an unfinished one stay in `live`
      // until the window passes them; their scene hold is already released)
      for (auto &r : live)
        if (!r->run.done && r->params.env_is && r->sh && r->sh->has_env) env_is = true;
      bool okc = true, launched = true;
      if (tuning.drain_rounds) {
        // the next round: a range of the admitted samples in admission order;
        // a range inside one request runs the single-job kernels
        if (!place_ok) {
          okc = launch_round(false, env_is) == OMR_OK;  // the refill hands out samples
        } else if (placed < total) {
          std::vector<uint64_t> jfirst, jcount;
          std::vector<uint32_t> jring;
          for (auto &r : live) {
            if (r->run.first + r->run.n_samples <= placed) continue;
            jfirst.push_back(r->run.first);
            jcount.push_back(r->run.n_samples);
            jring.push_back(r->run.ring);
          }
          uint32_t n0 = 0, solo_job = kNoSolo;
          bool env = true;
          plan_round(jfirst, jcount, jring, placed, total, &n0, &solo_job, &env);
          okc = launch_placed_round(env_is, placed, n0, solo_job, env) == OMR_OK;
          placed += n0;
        } else {
          launched = false;  // every admitted sample is placed: the statuses show the rest
        }
      } else {
        okc = launch_steps(tuning.steps_per_sync, env_is) == OMR_OK;
      }
      if (!okc || ((launched || pend.empty()) && !enqueue_status())) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "bounce step launch failed: " + get_err());
        return;
      }
      if (launched && pend.size() < 2) continue;  // fill the pipeline before waiting
    }
    if (pend.empty()) continue;
    // ---- 3. the oldest status
    const Pending P = pend.front();
    pend.pop_front();
    dev.unlock();
    const bool st_ok = hipEventSynchronize(ev_st[P.slot]) == hipSuccess;
    dev.lock();
    if (!st_ok) {
      dev.unlock();
      abort_with(OMR_ERR_HIP, "device error in the bounce loop");
      return;
    }
    const ServeStatus &S = h_st[P.slot];
    // drain: every admitted sample is handed out and few paths are left; only
    // from a status taken after the last drain with nothing admitted since
    // (the drain ends the queued paths; newly admitted samples would wait)
    if (!tuning.drain_rounds && tuning.drain_below && S.exhausted && S.queued > 0 && S.queued <= tuning.drain_below && P.total == total &&
        P.seq >= drain_seq) {
      const bool okd = k::drain(args(), shape(), 0, false, stream_) == hipSuccess &&
                       hipMemsetAsync(&d_ctrl_->qcount[0], 0, sizeof(d_ctrl_->qcount[0]), stream_) == hipSuccess &&
                       hipMemsetAsync(&d_ctrl_->qhead[0], 0, sizeof(d_ctrl_->qhead[0]), stream_) == hipSuccess;
      drain_seq = seq;
      if (!okd || !enqueue_status()) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "drain launch failed");
        return;
      }
    }
    // completions
    bool retired = false;
    for (auto &r : live) {
      JobRun &j = r->run;
      if (j.done || j.logical < P.lo || j.logical >= P.lo + P.n) continue;
      if (S.done[j.logical - P.lo] != j.n_samples) continue;
      j.done = true;
      const double nspp =
          (double)j.params.supersamples * j.params.supersamples * (double)(j.params.spp_end - j.params.spp_begin);
      bool okr = (!j.d_recs ||
                  k::gather_records(j.d_recs, (uint32_t)nspp, j.cam.width * j.cam.height, j.d_accum, stream_) ==
                      hipSuccess) &&
                 k::resolve(j.d_accum, j.d_out, j.cam.width * j.cam.height, nspp, stream_) == hipSuccess &&
                 hipEventCreateWithFlags(&r->ev_copy, hipEventDisableTiming) == hipSuccess &&
                 hipEventRecord(r->ev_copy, stream_) == hipSuccess;
      if (!okr) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "result resolve failed");
        return;
      }
      {
        std::lock_guard<std::mutex> g(cmu);
        to_copy.push_back(r);
      }
      ccv.notify_one();
      retired = true;
    }
    if (retired) {
      while (!live.empty() && live.front()->run.done) {
        live.pop_front();
        lo++;
      }
      if (k::admit(d_ctrl_, total, total, (uint32_t)lo, (uint32_t)hi, stream_) != hipSuccess) {
        dev.unlock();
        abort_with(OMR_ERR_HIP, "job window update failed");
        return;
      }
    }
    dev.unlock();
    {
      std::lock_guard<std::mutex> g(smu);
      srv_segments_ = S.segments;
      srv_shadow_ = S.shadow_rays;
      if (retired && live.empty()) busy_ms_ += ms_since(busy_t0);
    }
  }
  {
    std::lock_guard<std::mutex> g(mu);
    hipStreamSynchronize(stream_);
  }
  stop_copier();
  {
    std::lock_guard<std::mutex> g(smu);
    for (auto &r : live) {
      if (r->run.done) continue;  // finished: completed by the copier
      r->state = 3;
      if (r->sh) r->sh->users--;
      r->sh = nullptr;
    }
    for (auto &r : queue_) {
      r->state = 3;
      if (r->sh) r->sh->users--;
      r->sh = nullptr;
    }
    queue_.clear();
    scv.notify_all();
  }
  release_all();
}

}  // namespace omr
