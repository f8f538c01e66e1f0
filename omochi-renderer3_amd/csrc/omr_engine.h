// omr_engine.h -- device runtime behind the C ABI: resident scenes, the
// wavefront path pool, the job ring and the bounce-step loop.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/omr_capi.h"
#include "omr_bvh.h"
#include "omr_kernels.h"
#include "omr_admission.h"
#include "omr_layout.h"

namespace omr {

// Resident scenes per context.  A scene id is use * kMaxScenes + slot: a
// released slot takes the next upload, under a new id, and the device scene
// table is indexed by the slot.
constexpr uint32_t kMaxScenes = kMaxSceneSlots;
inline uint32_t scene_slot(omr_scene_id id) { return (uint32_t)id % kMaxScenes; }

// Summaries of the sharded control block as read back to the host.
inline unsigned long long sum_shards(const Sharded<unsigned long long> &a) {
  unsigned long long s = 0;
  for (uint32_t k = 0; k < kShards; k++) s += a[k];
  return s;
}
// Completed samples of the job in ring slot `ring` (host copy of the
// kShards x kJobRing counter array).
inline unsigned long long job_done(const unsigned long long *done, uint32_t ring) {
  unsigned long long s = 0;
  for (uint32_t k = 0; k < kShards; k++) s += done[done_index(k, ring)];
  return s;
}
inline uint64_t queue_len(const Ctrl &c, int parity) {
  uint64_t s = 0;
  for (uint32_t k = 0; k < kShards; k++) s += c.qcount[parity][k];
  return s;
}
inline bool samples_exhausted(const Ctrl &c) {
  for (uint32_t k = 0; k < kShards; k++)
    if (c.next_sample[k] < shard_limit(c.total_samples, k)) return false;
  return true;
}

struct SceneHost {
  omr_scene_id id = -1;
  BuildOutput bvh;
  SceneDev dev{};
  std::vector<void *> allocs;
  size_t bytes = 0;
  bool has_env = false;
  bool emissive_specular = false;  // a mirror or glass material emits (refused with next-event estimation)
  uint32_t env_w = 0, env_h = 0;
  std::atomic<int> users{0};  // queued or live served jobs holding the scene (guarded by Engine::scene_mu)
};

// One job as the runtime sees it (plain, batched or served).
struct JobRun {
  uint32_t scene = 0;
  omr_camera cam{};
  omr_render_params params{};
  uint32_t n_samples = 0;
  unsigned long long *d_accum = nullptr;  // W*H*3 fixed-point sums
  unsigned long long *d_recs = nullptr;   // n_samples*3 per-sample sums, or null (atomics on d_accum)
  float *d_out = nullptr;                 // W*H*3 resolved means (optional)
  uint32_t ring = 0;                      // ring slot while live
  uint64_t logical = 0;                   // logical ring index
  unsigned long long first = 0;           // its first sample in the run's global sample order
  bool done = false;
};

class Engine {
 public:
  explicit Engine(int device);
  ~Engine();
  bool ok() const { return init_ok_; }

  // Last error text.  The server thread and the caller's thread both report
  // errors, so the text sits behind its own lock.
  void set_err(const std::string &m) {
    std::lock_guard<std::mutex> g(err_mu_);
    err_ = m;
  }
  std::string get_err() const {
    std::lock_guard<std::mutex> g(err_mu_);
    return err_;
  }
  std::mutex mu;  // guards the stream and all device state
  // Guards the scene table entries and SceneHost::users: a served request
  // takes its scene (lookup + users++) under this lock, and release checks
  // users and removes the entry under it, so neither can interleave.
  std::mutex scene_mu;

  int device_id() const { return device_; }
  int n_cu() const { return n_cu_; }
  size_t hbm_bytes() const { return hbm_; }
  const std::string &device_name() const { return name_; }

  int upload_scene(const omr_scene_desc *desc, omr_scene_id *id);
  int release_scene(omr_scene_id id);
  // The scene of `id` or null.  Entries never move (fixed table); an entry
  // is removed only by release_scene, under `mu` and `scene_mu`.
  SceneHost *scene(omr_scene_id id);
  // Device bytes of every resident scene.
  uint64_t scene_bytes();
  // traversal stack entries a scene needs at the current node width
  uint32_t stack_need(const SceneHost *s) const {
    return tuning.node_width == 8 ? s->bvh.stack_bound8 : s->bvh.stack_bound;
  }
  // Device bytes of the path pool, its queues and the traversal overflow
  // stacks as allocated now.
  uint64_t pool_bytes() const;

  int trace(omr_scene_id id, const float *d_rays, size_t n, int32_t *d_prim, float *d_t, uint64_t *stats2);
  // trace() without any device allocation: its fetch heads and overflow
  // stacks live in the fixed scratch array of this GPU (k::noalloc_scratch),
  // whose address is looked up once, at construction.
  int trace_noalloc(omr_scene_id id, const float *d_rays, size_t n, int32_t *d_prim, float *d_t);

  // Runs jobs to completion on the shared queues (plain = one job).
  int run_jobs(std::vector<JobRun> &jobs, omr_render_stats *st);

  omr_tuning tuning;
  int set_tuning(const omr_tuning &t);
  bool profile = false;

  // ---- served mode ----
  int server_start(const omr_server_config &cfg);
  int server_stop();
  int submit(omr_scene_id scene, const omr_camera &cam, const omr_render_params &p, float *out, omr_job_id *id);
  int poll(omr_job_id id, int *state);
  int wait(omr_job_id id, int64_t timeout_ms);
  int cancel(omr_job_id id);
  int server_stats(omr_server_stats *s);
  bool server_running() const { return server_on_.load(); }

  int env_tables_host(omr_scene_id id, uint64_t *row, uint64_t *marg);
  int job_error(omr_job_id id, std::string *msg);
  int job_latency(omr_job_id id, double *ms);

 private:
  mutable std::mutex err_mu_;
  std::string err_;
  bool init_ok_ = false;
  int device_ = 0;
  int n_cu_ = 0;
  size_t hbm_ = 0;
  std::string name_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr, ev_c_ = nullptr, ev_d_ = nullptr;

  void *noalloc_scratch_ = nullptr;  // device address of this GPU's allocation-free scratch
  std::string noalloc_err_;
  std::unique_ptr<SceneHost> scenes_[kMaxScenes];
  uint32_t slot_uses_[kMaxScenes] = {};  // uploads each slot has held (the id's high part)
  SceneDev *d_scenes_ = nullptr;

  // The second path pool ("lane") of a run with tuning.lanes = 2: its own
  // pool, control block, overflow stacks, graphs and stream.  swap_lane()
  // exchanges it with the current one, so every launcher works on whichever
  // lane is current; placed drain rounds alternate between the two, and one
  // round's drain tail overlaps the next round's fill and steps on the other
  // stream.  Lane 0 is current outside run_jobs.
  struct Lane {
    PoolPtrs pool{};
    uint32_t pool_cap = 0;
    Ctrl *d_ctrl = nullptr;
    Ctrl *h_ctrl = nullptr;
    uint2 *d_gstack = nullptr;
    uint32_t gcap = 0;
    size_t gstack_threads = 0;
    hipGraphExec_t graph[8] = {};
    uint32_t graph_steps = 0;
    bool graph_round = false;
    hipStream_t stream = nullptr;
  };
  Lane alt_;
  bool on_alt_ = false;  // the members below hold lane 1
  void swap_lane();
  void free_lane_pool();  // frees the current lane's pool buffers

  // pool
  PoolPtrs pool_{};
  uint32_t pool_cap_ = 0;
  uint32_t grid_threads_ = 0;
  Ctrl *d_ctrl_ = nullptr;
  Ctrl *h_ctrl_ = nullptr;               // pinned readback
  JobDev *d_jobs_ = nullptr;
  unsigned long long *d_done_ = nullptr; // kJobRing counters
  unsigned long long *h_done_ = nullptr; // pinned readback
  uint2 *d_gstack_ = nullptr;
  uint32_t gcap_ = 0;
  size_t gstack_threads_ = 0;

  // graphs: [with shadow kernel + 2 x single-job kernels]
  // bounce-step graphs by (shadow rays, single-job, single-job scene without environment map)
  hipGraphExec_t graph_[8] = {};
  uint32_t graph_steps_ = 0;
  bool graph_round_ = false;       // the graphs hold drain-round steps (no regeneration, no refill)
  uint32_t solo_ = kNoSolo;        // the current launches are single-job (Ctrl::solo_job names it), or kNoSolo
  bool any_env_ = true;            // a job of the current launches has a scene with an environment map
  uint8_t ring_env_[kJobRing] = {};  // per job-ring slot: the job's scene has an environment map

  int ensure_pool();
  int ensure_gstack(uint32_t need_entries, size_t threads);
  void drop_graphs();
  // `steps` bounce steps; round: the steps of a drain round (finished slots
  // go idle, no regeneration or bulk refill)
  int launch_steps(uint32_t steps, bool with_shadow, bool round = false);
  // One drain round on queue 0: (unless first) every idle slot takes a new
  // sample, tuning.round_steps bounce steps, then one drain launch.
  int launch_round(bool first, bool with_shadow);
  // A round the host places: samples [base, base + n0) fill the idle pool
  // without atomics, round_steps steps, then the drain; solo_job = the ring
  // slot of the round's only job (single-job kernels), or kNoSolo.
  int launch_placed_round(bool with_shadow, uint64_t base, uint32_t n0, uint32_t solo_job, bool env);
  // The next placed round at global sample `base` of `total`, over jobs whose
  // sample ranges [first[i], first[i] + count[i]) follow one another: its
  // sample count, its only job's ring slot (or kNoSolo) and whether a job in
  // it has a scene with an environment map.
  void plan_round(const std::vector<uint64_t> &first, const std::vector<uint64_t> &count,
                  const std::vector<uint32_t> &ring, uint64_t base, uint64_t total, uint32_t *n0,
                  uint32_t *solo_job, bool *env) const;
  // profiling: event pairs around each launch of the last chunk
  std::vector<hipEvent_t> prof_ev_;
  std::vector<int> prof_kind_;  // 0 extend, 1 shade, 2 shadow
  double prof_ms_[3] = {0, 0, 0};
  uint64_t prof_extend_launches_ = 0, prof_shadow_launches_ = 0;
  int collect_profile();
  LaunchShape shape() const;
  KernelArgs args() const;
  int fail(int code, const std::string &msg);
  int hipfail(hipError_t e, const char *what);
  void free_scene_mem(SceneHost &s);

  // ---- served mode state ----
  struct Req {
    omr_job_id id;
    omr_scene_id scene;
    omr_camera cam;
    omr_render_params params;
    float *out;
    std::chrono::steady_clock::time_point t_submit;
    SceneHost *sh = nullptr;  // its scene, held (users) from submission to completion
    int state = 0;  // 0 queued 1 running 2 done 3 cancelled 4 failed 5 rejected (HBM budget)
    double lat_ms = 0.0;      // submission to result in the caller's buffer (done requests)
    int code = OMR_OK;        // error code of a failed or rejected request
    std::string msg;          // its error text
    Placement place;          // its block in the request arena: sums, image and per-sample records
    JobRun run;
    bool cancel_req = false;
    hipEvent_t ev_copy = nullptr;  // its image is resolved (the copy out waits on this)
  };
  std::thread server_;
  std::atomic<bool> server_on_{false};
  std::atomic<bool> server_stop_{false};
  std::mutex smu;
  std::condition_variable scv;
  std::deque<std::shared_ptr<Req>> queue_;
  std::vector<std::shared_ptr<Req>> all_;  // by id-1
  omr_server_config scfg_{};
  uint64_t srv_record_bytes_ = 0;  // per-request record limit, fixed at server start
  uint32_t saved_pool_paths_ = 0;  // tuning.pool_paths before server_start changed it
  size_t used_bytes_ = 0;          // device bytes of the admitted requests
  // Request arena: one device allocation made at server start (what the pool
  // and scenes leave of the budget), sub-allocated first-fit per request, so
  // serving allocates and frees no device memory.  Blocks are returned by the
  // copier thread, taken by the server thread: own lock.
  char *arena_ = nullptr;
  uint64_t arena_bytes_ = 0;
  std::mutex arena_mu_;
  Arena arena_map_;  // free ranges of arena_
  bool arena_take(uint64_t bytes, uint64_t *off);
  void arena_give(uint64_t off, uint64_t bytes);
  std::atomic<uint64_t> fixed_bytes_{0};  // device bytes of the pool, queues and stacks while serving
  bool server_failed_ = false;     // the server thread stopped after an error (srv_err_)
  std::string srv_err_;
  uint64_t n_submitted_ = 0, n_done_ = 0, n_rejected_ = 0, n_cancelled_ = 0;
  uint64_t srv_segments_ = 0, srv_shadow_ = 0;
  double busy_ms_ = 0.0;
  LatencyBook latencies_;
  void server_loop();
};

}  // namespace omr
