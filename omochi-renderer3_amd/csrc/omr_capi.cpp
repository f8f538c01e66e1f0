// omr_capi.cpp -- extern "C" entry points (declared in include/omr_capi.h).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/omr_capi.h"
#include "omr_engine.h"

struct omr_ctx {
  std::unique_ptr<omr::Engine> eng;
  std::string err;
  omr_render_stats last{};
  // grow-only device scratch for per-call accumulators and outputs
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
};

namespace {

int set_err(omr_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}
int from_engine(omr_ctx *c, int code) {
  if (code != OMR_OK && c) c->err = c->eng->get_err();
  return code;
}

int check_job(omr_ctx *c, const omr_camera *cam, const omr_render_params *p, uint64_t *n_out) {
  if (!cam || !p) return set_err(c, OMR_ERR_INVALID, "null camera or params");
  if (cam->width == 0 || cam->height == 0) return set_err(c, OMR_ERR_INVALID, "image size is zero");
  if (p->supersamples == 0 || p->supersamples > 64) return set_err(c, OMR_ERR_INVALID, "supersamples must be 1..64");
  if (p->spp_end <= p->spp_begin) return set_err(c, OMR_ERR_INVALID, "spp_end must exceed spp_begin");
  if (p->max_depth > 255) return set_err(c, OMR_ERR_INVALID, "max_depth must be <= 255 (0 = unlimited)");
  if (p->nee > 1) return set_err(c, OMR_ERR_INVALID, "nee must be 0 or 1");
  uint64_t n = (uint64_t)cam->width * cam->height * p->supersamples * p->supersamples *
               (uint64_t)(p->spp_end - p->spp_begin);
  if (n > 0xFFFFFFFFull) return set_err(c, OMR_ERR_INVALID, "job exceeds 2^32-1 samples; split spp into batches");
  *n_out = n;
  return OMR_OK;
}

int ensure_scratch(omr_ctx *c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return OMR_OK;
  if (c->scratch) hipFree(c->scratch);
  c->scratch = nullptr;
  c->scratch_bytes = 0;
  if (hipMalloc(&c->scratch, bytes) != hipSuccess) return set_err(c, OMR_ERR_NOMEM, "device scratch allocation failed");
  c->scratch_bytes = bytes;
  return OMR_OK;
}

// Plain and batched rendering share this: one persistent wavefront run.
int render_impl(omr_ctx *c, size_t n, const omr_scene_id *scenes, const omr_camera *cams,
                const omr_render_params *params, float *const *outs, int64_t *sums_out) {
  if (!c) return OMR_ERR_INVALID;
  if (c->eng->server_running()) return set_err(c, OMR_ERR_STATE, "server mode is running");
  if (n == 0) return OMR_OK;
  std::vector<omr::JobRun> jobs(n);
  size_t bytes = 0, rec_bytes = 0;
  const size_t rec_cap = (size_t)c->eng->tuning.record_mb << 20;
  std::vector<size_t> off_acc(n), off_out(n), off_rec(n, SIZE_MAX);
  for (size_t i = 0; i < n; i++) {
    uint64_t ns;
    int rc = check_job(c, &cams[i], &params[i], &ns);
    if (rc) return rc;
    if (!c->eng->scene(scenes[i])) return set_err(c, OMR_ERR_NOTFOUND, "unknown scene id");
    jobs[i].scene = (uint32_t)scenes[i];
    jobs[i].cam = cams[i];
    jobs[i].params = params[i];
    jobs[i].n_samples = (uint32_t)ns;
    size_t px = (size_t)cams[i].width * cams[i].height;
    off_acc[i] = bytes;
    bytes += (px * 3 * 8 + 255) & ~(size_t)255;
    off_out[i] = bytes;
    bytes += (px * 3 * 4 + 255) & ~(size_t)255;
  }
  // per-sample records, after every job's sums and image, while the
  // render's records fit in record_mb
  for (size_t i = 0; i < n; i++) {
    const size_t rb = ((size_t)jobs[i].n_samples * 24 + 255) & ~(size_t)255;
    if (rec_bytes + rb <= rec_cap) {
      off_rec[i] = bytes + rec_bytes;
      rec_bytes += rb;
    }
  }
  std::lock_guard<std::mutex> g(c->eng->mu);
  hipSetDevice(c->eng->device_id());
  int rc = ensure_scratch(c, bytes + rec_bytes);
  if (rc && rec_bytes) {  // no room for the records: pixel atomics instead (the same results)
    (void)hipGetLastError();  // the failed allocation must not surface at the next launch check
    std::fill(off_rec.begin(), off_rec.end(), SIZE_MAX);
    rc = ensure_scratch(c, bytes);
  }
  if (rc) return rc;
  for (size_t i = 0; i < n; i++) {
    jobs[i].d_accum = (unsigned long long *)((char *)c->scratch + off_acc[i]);
    jobs[i].d_out = (float *)((char *)c->scratch + off_out[i]);
    if (off_rec[i] != SIZE_MAX) jobs[i].d_recs = (unsigned long long *)((char *)c->scratch + off_rec[i]);
  }
  omr_render_stats st{};
  rc = c->eng->run_jobs(jobs, &st);
  if (rc) return from_engine(c, rc);
  for (size_t i = 0; i < n; i++) {
    size_t px = (size_t)cams[i].width * cams[i].height;
    hipError_t e;
    if (sums_out) e = hipMemcpy(sums_out, jobs[i].d_accum, px * 3 * 8, hipMemcpyDeviceToHost);
    else e = hipMemcpy(outs[i], jobs[i].d_out, px * 3 * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_err(c, OMR_ERR_HIP, std::string("result copy: ") + hipGetErrorString(e));
  }
  c->last = st;
  return OMR_OK;
}

}  // namespace

extern "C" {

int omr_create(omr_ctx **ctx, int device) {
  if (!ctx) return OMR_ERR_INVALID;
  *ctx = nullptr;
  auto c = std::make_unique<omr_ctx>();
  c->eng = std::make_unique<omr::Engine>(device);
  if (!c->eng->ok()) return OMR_ERR_NODEVICE;
  *ctx = c.release();
  return OMR_OK;
}

void omr_destroy(omr_ctx *ctx) {
  if (!ctx) return;
  if (ctx->scratch) {
    hipSetDevice(ctx->eng->device_id());
    hipFree(ctx->scratch);
  }
  delete ctx;
}

const char *omr_last_error(omr_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int omr_device_info(omr_ctx *ctx, int *n_cu, uint64_t *hbm_bytes, char *name, size_t name_len) {
  if (!ctx) return OMR_ERR_INVALID;
  if (n_cu) *n_cu = ctx->eng->n_cu();
  if (hbm_bytes) *hbm_bytes = ctx->eng->hbm_bytes();
  if (name && name_len) {
    std::strncpy(name, ctx->eng->device_name().c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  return OMR_OK;
}

int omr_scene_upload(omr_ctx *ctx, const omr_scene_desc *desc, omr_scene_id *id) {
  if (!ctx) return OMR_ERR_INVALID;
  return from_engine(ctx, ctx->eng->upload_scene(desc, id));
}

int omr_scene_release(omr_ctx *ctx, omr_scene_id id) {
  if (!ctx) return OMR_ERR_INVALID;
  return from_engine(ctx, ctx->eng->release_scene(id));
}

int omr_scene_info_get(omr_ctx *ctx, omr_scene_id id, omr_scene_info *info) {
  if (!ctx || !info) return OMR_ERR_INVALID;
  omr::SceneHost *s = ctx->eng->scene(id);
  if (!s) return set_err(ctx, OMR_ERR_NOTFOUND, "unknown scene id");
  info->n_nodes = (uint32_t)s->bvh.nodes.size();
  info->depth = s->bvh.depth;
  info->stack_bound = s->bvh.stack_bound;
  info->n_prims = (uint32_t)s->bvh.prims.size();
  info->device_bytes = s->bytes;
  info->build_ms = s->bvh.build_ms;
  info->n_nodes8 = (uint32_t)s->bvh.nodes8.size();
  info->depth8 = s->bvh.depth8;
  info->stack_bound8 = s->bvh.stack_bound8;
  info->pad = 0;
  return OMR_OK;
}

int omr_scene_bvh(omr_ctx *ctx, omr_scene_id id, const void **nodes, size_t *n_nodes, const uint32_t **prim_order,
                  size_t *n_order) {
  if (!ctx || !nodes || !n_nodes || !prim_order || !n_order) return OMR_ERR_INVALID;
  omr::SceneHost *s = ctx->eng->scene(id);
  if (!s) return set_err(ctx, OMR_ERR_NOTFOUND, "unknown scene id");
  *nodes = s->bvh.nodes.data();
  *n_nodes = s->bvh.nodes.size();
  *prim_order = s->bvh.prim_order.data();
  *n_order = s->bvh.prim_order.size();
  return OMR_OK;
}

struct omr_bvh {
  omr::BuildOutput out;
};

int omr_bvh_build(const omr_scene_desc *d, omr_bvh **out) {
  if (!d || !out) return OMR_ERR_INVALID;
  *out = nullptr;
  omr::BuildInput bi;
  bi.n_tri = d->n_triangles;
  bi.tri = d->triangles;
  bi.n_sph = d->n_spheres;
  bi.sph = d->spheres;
  if ((bi.n_tri && !bi.tri) || (bi.n_sph && !bi.sph)) return OMR_ERR_INVALID;
  auto b = std::make_unique<omr_bvh>();
  std::string err;
  if (!omr::build_wide_bvh(bi, b->out, err)) return OMR_ERR_INVALID;
  *out = b.release();
  return OMR_OK;
}

int omr_bvh_data(const omr_bvh *b, const void **nodes, size_t *n_nodes, const uint32_t **prim_order, size_t *n_order,
                 uint32_t *depth, uint32_t *stack_bound) {
  if (!b) return OMR_ERR_INVALID;
  if (nodes) *nodes = b->out.nodes.data();
  if (n_nodes) *n_nodes = b->out.nodes.size();
  if (prim_order) *prim_order = b->out.prim_order.data();
  if (n_order) *n_order = b->out.prim_order.size();
  if (depth) *depth = b->out.depth;
  if (stack_bound) *stack_bound = b->out.stack_bound;
  return OMR_OK;
}

int omr_bvh_data8(const omr_bvh *b, const void **nodes, size_t *n_nodes, uint32_t *depth, uint32_t *stack_bound) {
  if (!b) return OMR_ERR_INVALID;
  if (nodes) *nodes = b->out.nodes8.data();
  if (n_nodes) *n_nodes = b->out.nodes8.size();
  if (depth) *depth = b->out.depth8;
  if (stack_bound) *stack_bound = b->out.stack_bound8;
  return OMR_OK;
}

void omr_bvh_free(omr_bvh *b) { delete b; }

int omr_scene_env_tables(omr_ctx *ctx, omr_scene_id id, uint64_t *row, uint64_t *marg) {
  if (!ctx || !row || !marg) return OMR_ERR_INVALID;
  return from_engine(ctx, ctx->eng->env_tables_host(id, row, marg));
}

int omr_trace_closest_device(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n, int32_t *d_prim_id,
                             float *d_t) {
  if (!ctx || (n && (!d_rays || !d_prim_id || !d_t))) return OMR_ERR_INVALID;
  std::lock_guard<std::mutex> g(ctx->eng->mu);
  return from_engine(ctx, ctx->eng->trace(scene, d_rays, n, d_prim_id, d_t, nullptr));
}

int omr_trace_closest_device_noalloc(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n,
                                     int32_t *d_prim_id, float *d_t) {
  if (!ctx || (n && (!d_rays || !d_prim_id || !d_t))) return OMR_ERR_INVALID;
  // every context on a GPU shares that GPU's scratch array: calls from this
  // process take turns per device
  static std::mutex per_device[64];
  const int dev = ctx->eng->device_id();
  std::unique_lock<std::mutex> d;
  if (dev >= 0 && dev < 64) d = std::unique_lock<std::mutex>(per_device[dev]);
  std::lock_guard<std::mutex> g(ctx->eng->mu);
  return from_engine(ctx, ctx->eng->trace_noalloc(scene, d_rays, n, d_prim_id, d_t));
}

int omr_trace_stats_device(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n, int32_t *d_prim_id,
                           float *d_t, uint64_t *node_visits, uint64_t *prim_tests) {
  if (!ctx || (n && (!d_rays || !d_prim_id || !d_t))) return OMR_ERR_INVALID;
  std::lock_guard<std::mutex> g(ctx->eng->mu);
  uint64_t s2[2] = {0, 0};
  int rc = ctx->eng->trace(scene, d_rays, n, d_prim_id, d_t, s2);
  if (rc) return from_engine(ctx, rc);
  if (node_visits) *node_visits = s2[0];
  if (prim_tests) *prim_tests = s2[1];
  return OMR_OK;
}

int omr_trace_closest(omr_ctx *ctx, omr_scene_id scene, const float *rays, size_t n, int32_t *prim_id, float *t) {
  if (!ctx || (n && (!rays || !prim_id || !t))) return OMR_ERR_INVALID;
  if (n == 0) return OMR_OK;
  std::lock_guard<std::mutex> g(ctx->eng->mu);
  hipSetDevice(ctx->eng->device_id());
  float *d_r = nullptr, *d_t = nullptr;
  int32_t *d_p = nullptr;
  auto cleanup = [&]() {
    if (d_r) hipFree(d_r);
    if (d_t) hipFree(d_t);
    if (d_p) hipFree(d_p);
  };
  if (hipMalloc(&d_r, n * 32) != hipSuccess || hipMalloc(&d_t, n * 4) != hipSuccess ||
      hipMalloc(&d_p, n * 4) != hipSuccess) {
    cleanup();
    return set_err(ctx, OMR_ERR_NOMEM, "ray buffers");
  }
  int rc = OMR_OK;
  if (hipMemcpy(d_r, rays, n * 32, hipMemcpyHostToDevice) != hipSuccess) rc = set_err(ctx, OMR_ERR_HIP, "ray upload");
  if (!rc) rc = from_engine(ctx, ctx->eng->trace(scene, d_r, n, d_p, d_t, nullptr));
  if (!rc && (hipMemcpy(prim_id, d_p, n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(t, d_t, n * 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = set_err(ctx, OMR_ERR_HIP, "hit download");
  cleanup();
  return rc;
}

int omr_render(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
               float *out_rgb) {
  if (!out_rgb) return set_err(ctx, OMR_ERR_INVALID, "null output");
  float *outs[1] = {out_rgb};
  return render_impl(ctx, 1, &scene, cam, params, outs, nullptr);
}

int omr_render_sums(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
                    int64_t *out_sums) {
  if (!out_sums) return set_err(ctx, OMR_ERR_INVALID, "null output");
  float *outs[1] = {nullptr};
  return render_impl(ctx, 1, &scene, cam, params, outs, out_sums);
}

int omr_render_batch(omr_ctx *ctx, size_t n, const omr_scene_id *scenes, const omr_camera *cams,
                     const omr_render_params *params, float *const *outs) {
  if (n && (!scenes || !cams || !params || !outs)) return set_err(ctx, OMR_ERR_INVALID, "null argument");
  for (size_t i = 0; i < n; i++)
    if (!outs[i]) return set_err(ctx, OMR_ERR_INVALID, "null output");
  return render_impl(ctx, n, scenes, cams, params, outs, nullptr);
}

int omr_last_render_stats(omr_ctx *ctx, omr_render_stats *stats) {
  if (!ctx || !stats) return OMR_ERR_INVALID;
  *stats = ctx->last;
  return OMR_OK;
}

int omr_set_profiling(omr_ctx *ctx, int enable) {
  if (!ctx) return OMR_ERR_INVALID;
  std::lock_guard<std::mutex> g(ctx->eng->mu);
  ctx->eng->profile = enable != 0;
  return OMR_OK;
}

int omr_get_tuning(omr_ctx *ctx, omr_tuning *t) {
  if (!ctx || !t) return OMR_ERR_INVALID;
  *t = ctx->eng->tuning;
  return OMR_OK;
}

int omr_set_tuning(omr_ctx *ctx, const omr_tuning *t) {
  if (!ctx || !t) return OMR_ERR_INVALID;
  if (ctx->eng->server_running()) return set_err(ctx, OMR_ERR_STATE, "server mode is running");
  return from_engine(ctx, ctx->eng->set_tuning(*t));
}

int omr_autotune(omr_ctx *ctx, size_t n, const omr_scene_id *scenes, const omr_camera *cams,
                 const omr_render_params *params, uint32_t budget_ms, omr_tuning *best_out) {
  if (!ctx || n == 0 || !scenes || !cams || !params) return OMR_ERR_INVALID;
  if (ctx->eng->server_running()) return set_err(ctx, OMR_ERR_STATE, "server mode is running");
  std::vector<std::vector<float>> bufs(n);
  std::vector<float *> outs(n);
  for (size_t i = 0; i < n; i++) {
    bufs[i].resize((size_t)cams[i].width * cams[i].height * 3);
    outs[i] = bufs[i].data();
  }
  auto t0 = std::chrono::steady_clock::now();
  auto elapsed = [&] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
  // score = device Mrays/s of the workload, the lower of two runs (run-to-run
  // noise of a few percent would otherwise let a worse setting win)
  auto score = [&](const omr_tuning &t, double *out) -> int {
    int rc = omr_set_tuning(ctx, &t);
    if (rc) return rc;
    double lo = 0.0;
    for (int r = 0; r < 2; r++) {
      rc = render_impl(ctx, n, scenes, cams, params, outs.data(), nullptr);
      if (rc) return rc;
      const omr_render_stats &s = ctx->last;
      double v = s.gpu_ms > 0 ? (double)(s.segments + s.shadow_rays) / (s.gpu_ms * 1e3) : 0.0;
      lo = r == 0 ? v : std::min(lo, v);
    }
    *out = lo;
    return OMR_OK;
  };
  omr_tuning best = ctx->eng->tuning;
  double best_s = 0;
  int rc = score(best, &best_s);  // warm-up run (allocations, graph capture)
  if (rc) return rc;
  rc = score(best, &best_s);
  if (rc) return rc;
  struct Dim {
    uint32_t omr_tuning::*field;
    std::vector<uint32_t> values;
  };
  std::vector<Dim> dims = {
      {&omr_tuning::node_width, {4, 8}},
      {&omr_tuning::wg_size, {64, 128, 256}},
      {&omr_tuning::blocks_per_cu, {4, 8, 16, 19}},
      {&omr_tuning::pool_paths, {1u << 22, 1u << 24, 1u << 25, 1u << 26}},
      {&omr_tuning::refill_every, {0, 1, 2, 4}},
      {&omr_tuning::fetch_budget, {0, 4, 8, 16}},
      {&omr_tuning::fetch_min, {8, 16, 32}},
      {&omr_tuning::stack_short, {8, 16}},
      {&omr_tuning::drain_below, {0, 1u << 20, 1u << 22, 1u << 23, 1u << 24}},
      {&omr_tuning::steps_per_sync, {2, 4, 8, 16}},
      {&omr_tuning::drain_rounds, {0, 1}},
      {&omr_tuning::round_steps, {0, 2, 4}},
      {&omr_tuning::solo_round_min, {0, 1u << 20, 1u << 22, 1u << 24}},
      {&omr_tuning::lanes, {1, 2}},
  };
  for (int pass = 0; pass < 2; pass++) {
    for (auto &d : dims) {
      for (uint32_t v : d.values) {
        if (elapsed() > budget_ms) break;
        if (best.*(d.field) == v) continue;
        omr_tuning t = best;
        t.*(d.field) = v;
        double s = 0;
        if (score(t, &s) == OMR_OK && s > best_s * 1.03) {
          best_s = s;
          best = t;
        }
      }
    }
  }
  rc = omr_set_tuning(ctx, &best);
  if (rc) return rc;
  if (best_out) *best_out = best;
  return OMR_OK;
}

int omr_server_start(omr_ctx *ctx, const omr_server_config *cfg) {
  if (!ctx) return OMR_ERR_INVALID;
  omr_server_config c{};
  if (cfg) c = *cfg;
  return from_engine(ctx, ctx->eng->server_start(c));
}
int omr_server_stop(omr_ctx *ctx) {
  if (!ctx) return OMR_ERR_INVALID;
  return from_engine(ctx, ctx->eng->server_stop());
}
int omr_job_submit(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
                   float *out_rgb, omr_job_id *job) {
  if (!ctx || !out_rgb || !job) return OMR_ERR_INVALID;
  uint64_t ns;
  int rc = check_job(nullptr, cam, params, &ns);
  if (rc) return rc;
  return ctx->eng->submit(scene, *cam, *params, out_rgb, job);
}
int omr_job_poll(omr_ctx *ctx, omr_job_id job, int *state) {
  if (!ctx || !state) return OMR_ERR_INVALID;
  return ctx->eng->poll(job, state);
}
int omr_job_wait(omr_ctx *ctx, omr_job_id job, int64_t timeout_ms) {
  if (!ctx) return OMR_ERR_INVALID;
  return ctx->eng->wait(job, timeout_ms);
}
int omr_job_error(omr_ctx *ctx, omr_job_id job, char *msg, size_t len) {
  if (!ctx || !msg || len == 0) return OMR_ERR_INVALID;
  std::string m;
  int rc = ctx->eng->job_error(job, &m);
  if (rc) return rc;
  std::strncpy(msg, m.c_str(), len - 1);
  msg[len - 1] = 0;
  return OMR_OK;
}
int omr_job_latency(omr_ctx *ctx, omr_job_id job, double *ms) {
  if (!ctx || !ms) return OMR_ERR_INVALID;
  return ctx->eng->job_latency(job, ms);
}
int omr_job_cancel(omr_ctx *ctx, omr_job_id job) {
  if (!ctx) return OMR_ERR_INVALID;
  return ctx->eng->cancel(job);
}
int omr_server_stats_get(omr_ctx *ctx, omr_server_stats *stats) {
  if (!ctx || !stats) return OMR_ERR_INVALID;
  return ctx->eng->server_stats(stats);
}

int omr_serve_simulate(const omr_sim_request *requests, size_t n, const omr_sim_config *cfg,
                       omr_sim_outcome *outcomes, omr_sim_summary *summary) {
  if (!cfg) return OMR_ERR_INVALID;
  return omr::serve_simulate(requests, n, *cfg, outcomes, summary);
}

}  // extern "C"
