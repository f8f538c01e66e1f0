// omr_engine.cpp -- device runtime: scenes, path pool, bounce-step loop.
#include "omr_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace omr {

#define OMR_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hipfail(e_, #call);   \
  } while (0)

int Engine::fail(int code, const std::string &msg) {
  set_err(msg);
  return code;
}

int Engine::hipfail(hipError_t e, const char *what) {
  set_err(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? OMR_ERR_NOMEM : OMR_ERR_HIP;
}

Engine::Engine(int device) : device_(device) {
  // defaults: the best point of a sweep on MI355X (Cornell box + 82k-triangle
  // icosphere, 512x512, 64 spp)
  tuning.node_width = 4;
  tuning.wg_size = 64;
  tuning.blocks_per_cu = 19;  // the most 64-thread groups whose 16-entry LDS stacks fit in one CU
  tuning.pool_paths = 1u << 26;  // 15 GB of path state: more paths in flight, fewer steps (c2 +4 %)
  tuning.steps_per_sync = 2;
  tuning.use_graph = 1;
  tuning.refill_every = 0;
  tuning.fetch_budget = 8;
  tuning.fetch_min = 16;
  tuning.stack_short = kShortStackMax;
  tuning.drain_below = 1u << 24;
  tuning.record_mb = 32768;
  tuning.drain_rounds = 1;
  tuning.round_steps = 2;
  tuning.solo_round_min = 1u << 22;  // 4M samples (swept: half a pool, 16M, 4M, 1M)
  tuning.lanes = 2;  // a job's next round starts beside the last paths of the one before
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    set_err("no HIP device " + std::to_string(device));
    return;
  }
  if (hipSetDevice(device) != hipSuccess) { set_err("hipSetDevice failed"); return; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) { set_err("hipGetDeviceProperties failed"); return; }
  n_cu_ = prop.multiProcessorCount;
  hbm_ = prop.totalGlobalMem;
  name_ = prop.gcnArchName;
  if (hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking) != hipSuccess) { set_err("stream"); return; }
  if (hipEventCreate(&ev_a_) != hipSuccess || hipEventCreate(&ev_b_) != hipSuccess ||
      hipEventCreate(&ev_c_) != hipSuccess || hipEventCreate(&ev_d_) != hipSuccess) {
    set_err("event");
    return;
  }
  if (hipMalloc(&d_scenes_, kMaxScenes * sizeof(SceneDev)) != hipSuccess) { set_err("scene table"); return; }
  if (hipMemset(d_scenes_, 0, kMaxScenes * sizeof(SceneDev)) != hipSuccess) { set_err("scene table"); return; }
  if (hipMalloc(&d_ctrl_, sizeof(Ctrl)) != hipSuccess) { set_err("ctrl"); return; }
  if (hipHostMalloc(&h_ctrl_, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) { set_err("ctrl pinned"); return; }
  if (hipStreamCreateWithFlags(&alt_.stream, hipStreamNonBlocking) != hipSuccess) { set_err("stream"); return; }
  if (hipMalloc(&alt_.d_ctrl, sizeof(Ctrl)) != hipSuccess) { set_err("ctrl"); return; }
  if (hipHostMalloc(&alt_.h_ctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) { set_err("ctrl pinned"); return; }
  std::memset(alt_.h_ctrl, 0, sizeof(Ctrl));
  if (hipMalloc(&d_jobs_, kJobRing * sizeof(JobDev)) != hipSuccess) { set_err("jobs"); return; }
  if (hipMalloc(&d_done_, kJobRing * kShards * sizeof(unsigned long long)) != hipSuccess) { set_err("done"); return; }
  if (hipHostMalloc(&h_done_, kJobRing * kShards * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    set_err("done pinned");
    return;
  }
  std::memset(h_ctrl_, 0, sizeof(Ctrl));
  // the allocation-free entry point's scratch array: its address on this
  // device (each GPU holds its own copy), looked up once with the device current
  hipError_t se = k::noalloc_scratch(&noalloc_scratch_);
  if (se != hipSuccess) {
    noalloc_scratch_ = nullptr;
    noalloc_err_ = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
    (void)hipGetLastError();
  }
  init_ok_ = true;
}

Engine::~Engine() {
  if (server_on_.load()) server_stop();
  if (!init_ok_) return;
  hipSetDevice(device_);
  hipStreamSynchronize(stream_);
  if (alt_.stream) hipStreamSynchronize(alt_.stream);
  for (auto &s : scenes_)
    if (s) free_scene_mem(*s);
  for (int l = 0; l < 2; l++) {
    drop_graphs();
    free_lane_pool();
    if (d_gstack_) hipFree(d_gstack_);
    d_gstack_ = nullptr;
    swap_lane();
  }
  hipFree(alt_.d_ctrl);
  hipHostFree(alt_.h_ctrl);
  if (alt_.stream) hipStreamDestroy(alt_.stream);
  hipFree(d_scenes_);
  hipFree(d_ctrl_);
  hipFree(d_jobs_);
  hipFree(d_done_);
  hipHostFree(h_ctrl_);
  hipHostFree(h_done_);
  hipEventDestroy(ev_a_);
  hipEventDestroy(ev_b_);
  hipEventDestroy(ev_c_);
  hipEventDestroy(ev_d_);
  hipStreamDestroy(stream_);
}

void Engine::drop_graphs() {
  for (auto &g : graph_) {
    if (g) hipGraphExecDestroy(g);
    g = nullptr;
  }
  graph_steps_ = 0;
}

int Engine::set_tuning(const omr_tuning &t) {
  if (t.node_width != 4 && t.node_width != 8) return fail(OMR_ERR_INVALID, "node_width must be 4 or 8");
  if (t.wg_size != 64 && t.wg_size != 128 && t.wg_size != 256) return fail(OMR_ERR_INVALID, "wg_size must be 64/128/256");
  if (t.blocks_per_cu < 1 || t.blocks_per_cu > 64) return fail(OMR_ERR_INVALID, "blocks_per_cu out of range");
  if (t.pool_paths < 64 || t.pool_paths > (1u << 26)) return fail(OMR_ERR_INVALID, "pool_paths out of range");
  if (t.steps_per_sync < 2 || t.steps_per_sync > 256) return fail(OMR_ERR_INVALID, "steps_per_sync out of range");
  if (t.refill_every > 64) return fail(OMR_ERR_INVALID, "refill_every out of range");
  if (t.fetch_budget > 1024) return fail(OMR_ERR_INVALID, "fetch_budget out of range");
  if (t.fetch_budget && (t.fetch_min < 1 || t.fetch_min > 64)) return fail(OMR_ERR_INVALID, "fetch_min must be 1..64");
  if (t.stack_short != kShortStackMin && t.stack_short != kShortStackMax)
    return fail(OMR_ERR_INVALID, "stack_short must be 8 or 16");
  if ((uint64_t)t.record_mb << 20 > hbm_) return fail(OMR_ERR_INVALID, "record_mb exceeds the device memory");
  if (t.drain_rounds > 1) return fail(OMR_ERR_INVALID, "drain_rounds must be 0 or 1");
  if (t.round_steps > 16 || (t.round_steps & 1u)) return fail(OMR_ERR_INVALID, "round_steps must be even, 0..16");
  if (t.lanes != 1 && t.lanes != 2) return fail(OMR_ERR_INVALID, "lanes must be 1 or 2");
  std::lock_guard<std::mutex> g(mu);
  hipSetDevice(device_);
  hipStreamSynchronize(stream_);
  drop_graphs();
  swap_lane();
  hipStreamSynchronize(stream_);
  drop_graphs();
  if (t.lanes == 1) {
    // the second pool is not kept when runs no longer use it
    free_lane_pool();
    if (d_gstack_) hipFree(d_gstack_);
    d_gstack_ = nullptr;
    gcap_ = 0;
    gstack_threads_ = 0;
  }
  swap_lane();
  tuning = t;
  tuning.steps_per_sync = (t.steps_per_sync + 1u) & ~1u;
  return OMR_OK;
}

// Sub-queue capacity.  A producer kernel running GW waves in a grid-stride
// loop over at most P entries, C consecutive entries per wave and trip,
// appends at most C * ceil(P / (C GW)) entries per wave, and a shard
// (ceil(GW / kShards) waves) receives at most ceil(GW/kShards) times that.
// Shade (its own grid, C = 64) and refill (the persistent grid, C =
// kRefillSlots) may both append to one queue in one step.
static uint64_t shard_appends(uint32_t P, uint64_t gw, uint64_t chunk) {
  gw = std::max<uint64_t>(1, gw);
  uint64_t per_wave = chunk * ((P + chunk * gw - 1) / (chunk * gw));
  return (gw + kShards - 1) / kShards * per_wave;
}
static uint32_t sub_queue_cap(uint32_t P, uint64_t grid_waves, uint64_t shade_waves, uint64_t shade6_waves) {
  return (uint32_t)std::min<uint64_t>(
      P, shard_appends(P, grid_waves, kRefillSlots) +
             std::max(shard_appends(P, shade_waves, 64), shard_appends(P, shade6_waves, 64)));
}
// Workgroups of the persistent kernels and of shade for a pool of P paths:
// shade runs five waves per SIMD (20 per CU), its single-job form for a
// scene without an environment map six (24 per CU), all resident at once.
static void grids(uint32_t n_cu, const omr_tuning &t, uint32_t P, uint32_t &grid, uint32_t &shade_grid,
                  uint32_t &shade6_grid) {
  uint32_t need = std::max(1u, (P + t.wg_size - 1) / t.wg_size);
  grid = std::min(need, n_cu * t.blocks_per_cu);
  shade_grid = std::min(need, n_cu * (20u * 64u / t.wg_size));
  shade6_grid = std::min(need, n_cu * (24u * 64u / t.wg_size));
}

LaunchShape Engine::shape() const {
  LaunchShape s;
  s.wg = tuning.wg_size;
  grids((uint32_t)n_cu_, tuning, pool_cap_, s.grid, s.shade_grid, s.shade6_grid);
  return s;
}

KernelArgs Engine::args() const {
  KernelArgs a;
  a.ctrl = d_ctrl_;
  a.jobs = d_jobs_;
  a.done = d_done_;
  a.scenes = d_scenes_;
  a.p = pool_;
  a.p.gstack = d_gstack_;
  a.p.gcap = gcap_;
  a.p.df_budget = tuning.fetch_budget;
  a.p.df_fetch_min = tuning.fetch_min;
  a.p.stack_short = tuning.stack_short;
  a.p.node_width = tuning.node_width;
  a.solo = solo_;
  a.any_env = any_env_ ? 1u : 0u;
  return a;
}

void Engine::swap_lane() {
  std::swap(pool_, alt_.pool);
  std::swap(pool_cap_, alt_.pool_cap);
  std::swap(d_ctrl_, alt_.d_ctrl);
  std::swap(h_ctrl_, alt_.h_ctrl);
  std::swap(d_gstack_, alt_.d_gstack);
  std::swap(gcap_, alt_.gcap);
  std::swap(gstack_threads_, alt_.gstack_threads);
  for (int i = 0; i < 8; i++) std::swap(graph_[i], alt_.graph[i]);
  std::swap(graph_steps_, alt_.graph_steps);
  std::swap(graph_round_, alt_.graph_round);
  std::swap(stream_, alt_.stream);
  on_alt_ = !on_alt_;
}

void Engine::free_lane_pool() {
  PoolPtrs &p = pool_;
  void *bufs[] = {p.ray_o, p.ray_d, p.misc, p.rng, p.hit, p.pdf, p.pend, p.queue[0], p.queue[1], p.shadow};
  for (void *b : bufs)
    if (b) hipFree(b);
  p = PoolPtrs();
  pool_cap_ = 0;
}

int Engine::ensure_pool() {
  uint32_t P = tuning.pool_paths;
  uint32_t grid, sgrid, s6grid;
  grids((uint32_t)n_cu_, tuning, P, grid, sgrid, s6grid);
  uint32_t qcap = sub_queue_cap(P, (uint64_t)grid * tuning.wg_size / 64, (uint64_t)sgrid * tuning.wg_size / 64,
                                (uint64_t)s6grid * tuning.wg_size / 64);
  if (pool_cap_ == P && pool_.ray_o && pool_.qcap == qcap) return OMR_OK;
  drop_graphs();
  free_lane_pool();
  PoolPtrs &p = pool_;
  OMR_HIP(hipMalloc(&p.ray_o, (size_t)P * 16));
  OMR_HIP(hipMalloc(&p.ray_d, (size_t)P * 16));
  OMR_HIP(hipMalloc(&p.misc, (size_t)P * 16));
  OMR_HIP(hipMalloc(&p.rng, (size_t)P * 16));
  OMR_HIP(hipMalloc(&p.hit, (size_t)P * 16));
  OMR_HIP(hipMalloc(&p.pdf, (size_t)P * 4));
  OMR_HIP(hipMalloc(&p.pend, (size_t)P * 3 * 48));
  OMR_HIP(hipMalloc(&p.queue[0], (size_t)kShards * qcap * 4));
  OMR_HIP(hipMalloc(&p.queue[1], (size_t)kShards * qcap * 4));
  OMR_HIP(hipMalloc(&p.shadow, (size_t)kShards * qcap * 48));
  OMR_HIP(hipMemset(p.misc, 0, (size_t)P * 16));
  p.pool = P;
  p.qcap = qcap;
  pool_cap_ = P;
  return OMR_OK;
}

int Engine::ensure_gstack(uint32_t need_entries, size_t threads) {
  uint32_t cap = need_entries > tuning.stack_short ? need_entries - tuning.stack_short : 1u;
  if (d_gstack_ && cap <= gcap_ && threads <= gstack_threads_) return OMR_OK;
  drop_graphs();
  cap = std::max(cap, gcap_);
  threads = std::max(threads, gstack_threads_);
  // kernels already on the stream (a served chunk in flight, when a trace
  // call grows the stacks) still use the old stacks
  if (d_gstack_) {
    hipStreamSynchronize(stream_);
    hipFree(d_gstack_);
  }
  d_gstack_ = nullptr;
  OMR_HIP(hipMalloc(&d_gstack_, threads * cap * sizeof(uint2)));
  gcap_ = cap;
  gstack_threads_ = threads;
  return OMR_OK;
}

void Engine::free_scene_mem(SceneHost &s) {
  for (void *p : s.allocs)
    if (p) hipFree(p);
  s.allocs.clear();
}

SceneHost *Engine::scene(omr_scene_id id) {
  if (id < 0) return nullptr;
  SceneHost *s = scenes_[scene_slot(id)].get();
  return s && s->id == id ? s : nullptr;
}

uint64_t Engine::scene_bytes() {
  std::lock_guard<std::mutex> g(scene_mu);
  uint64_t b = 0;
  for (size_t i = 0; i < kMaxScenes; i++)
    if (scenes_[i]) b += scenes_[i]->bytes;
  return b;
}

uint64_t Engine::pool_bytes() const {
  // ray_o, ray_d, misc, rng, hit (16 B each), pdf (4 B), pending branches
  // (144 B) per slot; two queues and the shadow records per queue entry
  auto lane = [](uint64_t P, uint64_t qcap, uint64_t threads, uint64_t gcap) {
    return P * (5 * 16 + 4 + 3 * 48) + (uint64_t)kShards * qcap * (4 + 4 + 48) + threads * gcap * sizeof(uint2);
  };
  return lane(pool_cap_, pool_.qcap, gstack_threads_, gcap_) +
         lane(alt_.pool_cap, alt_.pool.qcap, alt_.gstack_threads, alt_.gcap);
}

int Engine::upload_scene(const omr_scene_desc *d, omr_scene_id *id) {
  if (!d || !id) return fail(OMR_ERR_INVALID, "null argument");
  size_t nt = d->n_triangles, ns = d->n_spheres, n = nt + ns, npl = d->n_planes, nall = n + npl;
  if (n == 0) return fail(OMR_ERR_INVALID, "scene has no triangles or spheres");
  if (nt && !d->triangles) return fail(OMR_ERR_INVALID, "triangles is NULL");
  if (ns && !d->spheres) return fail(OMR_ERR_INVALID, "spheres is NULL");
  if (npl && !d->planes) return fail(OMR_ERR_INVALID, "planes is NULL");
  if (npl > 64) return fail(OMR_ERR_INVALID, "too many planes (max 64)");
  for (size_t i = 0; i < npl; i++)
    if (!std::isfinite(d->planes[i])) return fail(OMR_ERR_INVALID, "plane " + std::to_string(i) + " is not finite");
  if (!d->prim_material || !d->materials || d->n_materials == 0) return fail(OMR_ERR_INVALID, "materials missing");
  if (d->n_materials > 0x00FFFFFFu) return fail(OMR_ERR_INVALID, "too many materials");
  if (nall > 0x7FFFFFFFull) return fail(OMR_ERR_INVALID, "too many primitives");
  for (size_t i = 0; i < d->n_materials; i++) {
    const omr_material &m = d->materials[i];
    if (m.type > 2) return fail(OMR_ERR_INVALID, "material " + std::to_string(i) + ": bad type");
    for (int c = 0; c < 3; c++)
      if (!std::isfinite(m.color[c]) || !std::isfinite(m.emission[c]) || m.emission[c] < 0.0f)
        return fail(OMR_ERR_INVALID, "material " + std::to_string(i) + ": bad color/emission");
    if (m.type == 2 && !(m.ior > 0.0f)) return fail(OMR_ERR_INVALID, "material " + std::to_string(i) + ": bad ior");
    if (m.texture > d->n_textures)
      return fail(OMR_ERR_INVALID, "material " + std::to_string(i) + ": texture index out of range");
  }
  if (d->n_textures && !d->textures) return fail(OMR_ERR_INVALID, "textures is NULL");
  if (d->n_textures > 4096) return fail(OMR_ERR_INVALID, "too many textures (max 4096)");
  size_t texels = 0;
  for (size_t i = 0; i < d->n_textures; i++) {
    const omr_texture &t = d->textures[i];
    if (!t.rgb || t.width == 0 || t.height == 0 || t.width > 32768 || t.height > 32768)
      return fail(OMR_ERR_INVALID, "texture " + std::to_string(i) + ": bad size or NULL pixels");
    texels += (size_t)t.width * t.height;
  }
  if (texels > 0xFFFFFFFFull) return fail(OMR_ERR_INVALID, "textures too large");
  std::vector<uint32_t> pm(nall);
  for (size_t i = 0; i < nall; i++) {
    uint32_t f = d->prim_material[i];
    uint32_t c = d->prim_material_count ? d->prim_material_count[i] : 1u;
    if (c < 1 || c > 256 || (size_t)f + c > d->n_materials)
      return fail(OMR_ERR_INVALID, "primitive " + std::to_string(i) + ": bad material range");
    pm[i] = f | ((c - 1u) << 24);
  }
  if (d->env_rgb && (d->env_width == 0 || d->env_height == 0)) return fail(OMR_ERR_INVALID, "env size is zero");

  auto sh = std::make_unique<SceneHost>();
  for (size_t i = 0; i < d->n_materials; i++) {
    const omr_material &m = d->materials[i];
    if (m.type != 0 && (m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f))
      sh->emissive_specular = true;
  }
  BuildInput bi;
  bi.n_tri = nt;
  bi.tri = d->triangles;
  bi.n_sph = ns;
  bi.sph = d->spheres;
  std::string berr;
  if (!build_wide_bvh(bi, sh->bvh, berr)) return fail(OMR_ERR_INVALID, berr);

  // per-triangle shading normals: n0, n1-n0, n2-n0 padded to 12 floats; the
  // first pad word holds the triangle's material word (prim_mat), so shading
  // a triangle hit needs no separate load for it
  std::vector<float> tn(nt * 12, 0.0f);
  for (size_t i = 0; i < nt; i++) {
    float *o = &tn[i * 12];
    std::memcpy(&o[3], &pm[i], 4);
    if (d->tri_normals) {
      const float *s = d->tri_normals + 9 * i;
      for (int a = 0; a < 3; a++) { o[a] = s[a]; o[4 + a] = s[3 + a]; o[8 + a] = s[6 + a]; }
    } else {
      // Polygon::CalculateNormal (src/renderer/Polygon.h:129-133): (v1-v0)x(v2-v0), normalized
      const float *t = d->triangles + 9 * i;
      double e1[3] = {t[3], t[4], t[5]}, e2[3] = {t[6], t[7], t[8]};
      double c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      double l = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      for (int a = 0; a < 3; a++) o[a] = l > 0 ? (float)(c[a] / l) : 0.0f;
    }
  }
  std::vector<MatRec> mats(d->n_materials);
  for (size_t i = 0; i < d->n_materials; i++) {
    const omr_material &m = d->materials[i];
    MatRec &r = mats[i];
    std::memset(&r, 0, sizeof(r));
    r.type = m.type;
    for (int c = 0; c < 3; c++) { r.color[c] = m.color[c]; r.emission[c] = m.emission[c]; }
    r.ior = m.ior;
    r.rate = m.rate;
    r.tex = m.texture;
  }
  // texels as 4 floats (one 16-byte load) after each other, with a directory
  std::vector<float> tex_px(texels * 4, 0.0f);
  std::vector<TexRec> tex_rec(d->n_textures);
  for (size_t i = 0, first = 0; i < d->n_textures; i++) {
    const omr_texture &t = d->textures[i];
    size_t m = (size_t)t.width * t.height;
    tex_rec[i] = TexRec{(uint32_t)first, t.width, t.height, 0u};
    for (size_t k = 0; k < m; k++)
      for (int c = 0; c < 3; c++) tex_px[(first + k) * 4 + c] = t.rgb[3 * k + c];
    first += m;
  }
  std::vector<float> uv;
  if (d->tri_uv && nt) {
    uv.assign(nt * 8, 0.0f);
    for (size_t i = 0; i < nt; i++)
      for (int c = 0; c < 6; c++) uv[i * 8 + c] = d->tri_uv[i * 6 + c];
  }

  std::lock_guard<std::mutex> g(mu);
  hipSetDevice(device_);
  size_t slot = 0;
  {
    std::lock_guard<std::mutex> sg(scene_mu);
    while (slot < kMaxScenes && scenes_[slot]) slot++;
  }
  if (slot >= kMaxScenes) return fail(OMR_ERR_INVALID, "too many resident scenes (max 64 per context)");
  if (slot_uses_[slot] >= 0x7FFFFFFFu / kMaxScenes) return fail(OMR_ERR_INVALID, "scene ids exhausted");
  SceneHost &S = *sh;
  auto dup = [&](const void *src, size_t bytes, void **dst) -> int {
    *dst = nullptr;
    if (bytes == 0) return OMR_OK;
    OMR_HIP(hipMalloc(dst, bytes));
    S.allocs.push_back(*dst);
    S.bytes += bytes;
    if (src) OMR_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return OMR_OK;
  };
  void *p_nodes, *p_nodes8, *p_prims, *p_tn = nullptr, *p_sph = nullptr, *p_pm, *p_mats, *p_pl = nullptr;
  void *p_tex = nullptr, *p_trec = nullptr, *p_uv = nullptr;
  int rc;
  if ((rc = dup(S.bvh.nodes.data(), S.bvh.nodes.size() * sizeof(WNode4), &p_nodes)) ||
      (rc = dup(S.bvh.nodes8.data(), S.bvh.nodes8.size() * sizeof(WNode8), &p_nodes8)) ||
      (rc = dup(S.bvh.prims.data(), S.bvh.prims.size() * sizeof(PrimRec), &p_prims)) ||
      (rc = dup(tn.data(), tn.size() * sizeof(float), &p_tn)) ||
      (rc = dup(d->spheres, ns * 4 * sizeof(float), &p_sph)) || (rc = dup(pm.data(), nall * 4, &p_pm)) ||
      (rc = dup(mats.data(), mats.size() * sizeof(MatRec), &p_mats)) ||
      (rc = dup(d->planes, npl * sizeof(float), &p_pl)) ||
      (rc = dup(tex_px.data(), tex_px.size() * sizeof(float), &p_tex)) ||
      (rc = dup(tex_rec.data(), tex_rec.size() * sizeof(TexRec), &p_trec)) ||
      (rc = dup(uv.data(), uv.size() * sizeof(float), &p_uv))) {
    free_scene_mem(S);
    return rc;
  }
  SceneDev &D = S.dev;
  D.nodes = (const WNode4 *)p_nodes;
  D.nodes8 = (const WNode8 *)p_nodes8;
  D.prims = (const PrimRec *)p_prims;
  D.tri_n = (const float *)p_tn;
  D.sph = (const float *)p_sph;
  D.prim_mat = (const uint32_t *)p_pm;
  D.mats = (const MatRec *)p_mats;
  D.n_tri = (uint32_t)nt;
  D.n_sph = (uint32_t)ns;
  D.planes = (const float *)p_pl;
  D.n_planes = (uint32_t)npl;
  D.tri_uv = (const float *)p_uv;
  D.tex = (const float *)p_tex;
  D.tex_rec = (const TexRec *)p_trec;
  D.n_tex = (uint32_t)d->n_textures;
  D.n_nodes = (uint32_t)S.bvh.nodes.size();
  D.stack_bound = S.bvh.stack_bound;
  D.n_nodes8 = (uint32_t)S.bvh.nodes8.size();
  D.stack_bound8 = S.bvh.stack_bound8;
  D.o_far = S.bvh.pad * 2097152.0f;  // pad / 8 = o_far * 2^-24
  if (d->env_rgb) {
    uint32_t w = d->env_width, h = d->env_height;
    void *p_env, *p_w, *p_max, *p_q, *p_row, *p_marg, *p_rg, *p_mg, *p_rs, *p_ms, *p_e4;
    if ((rc = dup(d->env_rgb, (size_t)w * h * 3 * sizeof(float), &p_env)) ||
        (rc = dup(nullptr, (size_t)w * h * sizeof(float), &p_w)) || (rc = dup(nullptr, 16, &p_max)) ||
        (rc = dup(nullptr, (size_t)w * h * 8, &p_q)) || (rc = dup(nullptr, (size_t)h * (w + 1) * 8, &p_row)) ||
        (rc = dup(nullptr, (size_t)(h + 1) * 8, &p_marg)) || (rc = dup(nullptr, (size_t)w * h * 8, &p_rg)) ||
        (rc = dup(nullptr, (size_t)h * 8, &p_mg)) || (rc = dup(nullptr, (size_t)h * 8, &p_rs)) ||
        (rc = dup(nullptr, 16, &p_ms)) || (rc = dup(nullptr, (size_t)w * h * 16, &p_e4))) {
      free_scene_mem(S);
      return rc;
    }
    OMR_HIP(hipMemset(p_max, 0, 16));
    hipError_t e = k::env_tables((const float *)p_env, w, h, (float *)p_w, (unsigned int *)p_max,
                                 (unsigned long long *)p_q, (unsigned long long *)p_row, (unsigned long long *)p_marg,
                                 stream_);
    if (e == hipSuccess)
      e = k::env_guides((const unsigned long long *)p_row, (const unsigned long long *)p_marg, w, h, (uint2_t *)p_rg,
                        (uint2_t *)p_mg, (double *)p_rs, (double *)p_ms, stream_);
    if (e == hipSuccess)
      e = k::env_pack((const float *)p_env, (const unsigned long long *)p_q, w * h, (float4 *)p_e4, stream_);
    if (e != hipSuccess) { free_scene_mem(S); return hipfail(e, "env tables"); }
    OMR_HIP(hipStreamSynchronize(stream_));
    D.env = (const float *)p_env;
    D.env_q = (const unsigned long long *)p_q;
    D.env4 = (const float *)p_e4;
    D.env_row = (const unsigned long long *)p_row;
    D.env_marg = (const unsigned long long *)p_marg;
    D.env_rguide = (const uint2_t *)p_rg;
    D.env_mguide = (const uint2_t *)p_mg;
    D.env_rscale = (const double *)p_rs;
    D.env_mscale = (const double *)p_ms;
    D.env_w = w;
    D.env_h = h;
    S.has_env = true;
    S.env_w = w;
    S.env_h = h;
  }
  OMR_HIP(hipMemcpy(d_scenes_ + slot, &D, sizeof(SceneDev), hipMemcpyHostToDevice));
  {
    std::lock_guard<std::mutex> sg(scene_mu);
    sh->id = (omr_scene_id)(slot_uses_[slot]++ * kMaxScenes + slot);
    *id = sh->id;
    scenes_[slot] = std::move(sh);
  }
  return OMR_OK;
}

int Engine::release_scene(omr_scene_id id) {
  std::lock_guard<std::mutex> g(mu);
  std::unique_ptr<SceneHost> s;
  {
    // the check and the removal under the lock a served submission takes
    // its scene under: no request can take the scene in between
    std::lock_guard<std::mutex> sg(scene_mu);
    SceneHost *p = scene(id);
    if (!p) return fail(OMR_ERR_NOTFOUND, "unknown scene");
    if (p->users.load() > 0) return fail(OMR_ERR_STATE, "scene is held by queued or running served jobs");
    s = std::move(scenes_[scene_slot(id)]);
  }
  hipSetDevice(device_);
  hipStreamSynchronize(stream_);
  free_scene_mem(*s);
  SceneDev z{};
  hipMemcpy(d_scenes_ + scene_slot(id), &z, sizeof(SceneDev), hipMemcpyHostToDevice);
  return OMR_OK;
}

int Engine::env_tables_host(omr_scene_id id, uint64_t *row, uint64_t *marg) {
  std::lock_guard<std::mutex> g(mu);
  SceneHost *s = scene(id);
  if (!s) return fail(OMR_ERR_NOTFOUND, "unknown scene");
  if (!s->has_env) return fail(OMR_ERR_STATE, "scene has no environment");
  hipSetDevice(device_);
  OMR_HIP(hipMemcpy(row, s->dev.env_row, (size_t)s->env_h * (s->env_w + 1) * 8, hipMemcpyDeviceToHost));
  OMR_HIP(hipMemcpy(marg, s->dev.env_marg, (size_t)(s->env_h + 1) * 8, hipMemcpyDeviceToHost));
  return OMR_OK;
}

int Engine::trace(omr_scene_id id, const float *d_rays, size_t n, int32_t *d_prim, float *d_t, uint64_t *stats2) {
  SceneHost *s = scene(id);
  if (!s) return fail(OMR_ERR_NOTFOUND, "unknown scene");
  if (n > 0xFFFFFFFFull) return fail(OMR_ERR_INVALID, "too many rays");
  if (n == 0) return OMR_OK;
  hipSetDevice(device_);
  const uint32_t wg = 128;
  uint32_t grid = (uint32_t)std::min<size_t>((n + wg - 1) / wg, (size_t)n_cu_ * 16);
  int rc = ensure_gstack(stack_need(s), (size_t)grid * wg);
  if (rc) return rc;
  // d_cnt[0..1]: node visits, primitive tests; one line on, the dynamic-fetch heads
  const size_t cnt_bytes = kLine + sizeof(Sharded<uint32_t>);
  unsigned long long *d_cnt = nullptr;
  OMR_HIP(hipMalloc(&d_cnt, cnt_bytes));
  hipError_t e = hipMemsetAsync(d_cnt, 0, cnt_bytes, stream_);
  if (e == hipSuccess)
    e = k::trace_rays(d_scenes_ + scene_slot(id), d_rays, (uint32_t)n, d_prim, d_t, stats2 ? d_cnt : nullptr, d_gstack_,
                      gcap_,
                      tuning.stack_short, tuning.node_width, grid, wg,
                      reinterpret_cast<Sharded<uint32_t> *>(d_cnt + kLine / 8),
                      (int)tuning.fetch_budget, tuning.fetch_min, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  unsigned long long c[2] = {0, 0};
  if (e == hipSuccess && stats2) e = hipMemcpy(c, d_cnt, 16, hipMemcpyDeviceToHost);
  hipFree(d_cnt);
  if (e != hipSuccess) return hipfail(e, "trace_rays");
  if (stats2) {
    stats2[0] = c[0];
    stats2[1] = c[1];
  }
  return OMR_OK;
}

int Engine::trace_noalloc(omr_scene_id id, const float *d_rays, size_t n, int32_t *d_prim, float *d_t) {
  if (!noalloc_scratch_) return fail(OMR_ERR_HIP, noalloc_err_);
  SceneHost *s = scene(id);
  if (!s) return fail(OMR_ERR_NOTFOUND, "unknown scene");
  if (n > 0xFFFFFFFFull) return fail(OMR_ERR_INVALID, "too many rays");
  if (n == 0) return OMR_OK;
  hipSetDevice(device_);
  // overflow entries per thread: the scene's exact stack bound past the LDS part
  const uint32_t wg = 128, shrt = tuning.stack_short;
  const uint32_t bound = stack_need(s);
  const uint32_t gcap = bound > shrt ? bound - shrt : 1u;
  const size_t per_group = (size_t)wg * gcap * sizeof(uint2);
  const size_t max_groups = (kNoAllocScratchBytes - kNoAllocHeadBytes) / per_group;
  if (max_groups == 0)
    return fail(OMR_ERR_INVALID, "the scene's traversal stack bound (" + std::to_string(bound) +
                                     " entries) does not fit the allocation-free scratch");
  // persistent grid: every ray is reached whatever the grid (dynamic fetch or
  // grid-stride), so the grid is cut to what the scratch holds
  uint32_t grid = (uint32_t)std::min<size_t>({(n + wg - 1) / wg, (size_t)n_cu_ * 16, max_groups});
  char *base = static_cast<char *>(noalloc_scratch_);
  Sharded<uint32_t> *heads = reinterpret_cast<Sharded<uint32_t> *>(base);
  uint2 *stacks = reinterpret_cast<uint2 *>(base + kNoAllocHeadBytes);
  hipError_t e = k::zero_words(reinterpret_cast<uint32_t *>(base), (uint32_t)(kNoAllocHeadBytes / 4), stream_);
  if (e == hipSuccess)
    e = k::trace_rays(d_scenes_ + scene_slot(id), d_rays, (uint32_t)n, d_prim, d_t, nullptr, stacks, gcap, shrt,
                      tuning.node_width,
                      grid, wg, heads, (int)tuning.fetch_budget, tuning.fetch_min, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  if (e != hipSuccess) return hipfail(e, "trace_rays (allocation-free)");
  return OMR_OK;
}

int Engine::launch_round(bool first, bool with_shadow) {
  const KernelArgs a = args();
  const LaunchShape sh = shape();
  hipError_t e = hipSuccess;
  if (!first) {
    // the last round's drain left queue 0 stale: clear it, then every idle slot takes a sample
    e = k::zero_words(reinterpret_cast<uint32_t *>(&d_ctrl_->qcount[0]), sizeof(d_ctrl_->qcount[0]) / 4, stream_);
    if (e == hipSuccess)
      e = k::zero_words(reinterpret_cast<uint32_t *>(&d_ctrl_->qhead[0]), sizeof(d_ctrl_->qhead[0]) / 4, stream_);
    if (e == hipSuccess) e = k::refill(a, sh, 0, stream_);
    if (e != hipSuccess) return hipfail(e, "round refill");
  }
  if (tuning.round_steps) {
    int rc = launch_steps(tuning.round_steps, with_shadow, true);
    if (rc) return rc;
  }
  // an even number of steps leaves the live paths in queue 0
  if (profile) OMR_HIP(hipEventRecord(ev_c_, stream_));
  e = k::drain(a, sh, 0, profile, stream_);
  if (e != hipSuccess) return hipfail(e, "drain");
  if (profile) OMR_HIP(hipEventRecord(ev_d_, stream_));
  return OMR_OK;
}

void Engine::plan_round(const std::vector<uint64_t> &first, const std::vector<uint64_t> &count,
                        const std::vector<uint32_t> &ring, uint64_t base, uint64_t total, uint32_t *n0,
                        uint32_t *solo_job, bool *env) const {
  const uint64_t P = pool_cap_;
  uint64_t n = std::min<uint64_t>(P, total - base);
  *solo_job = kNoSolo;
  // the job holding sample `base` (ranges ascend)
  auto it = std::upper_bound(first.begin(), first.end(), base);
  const size_t i = (size_t)(it - first.begin()) - 1;
  const uint64_t rem = first[i] + count[i] - base;
  if (rem >= n) {
    *solo_job = ring[i];  // the round lies inside one job
  } else if (rem >= (tuning.solo_round_min ? (uint64_t)tuning.solo_round_min : (P + 1) / 2)) {
    // enough of this job is left: it ends in a round of its own (single-job
    // kernels) rather than sharing one with the next job
    n = rem;
    *solo_job = ring[i];
  }
  *n0 = (uint32_t)n;
  *env = false;
  for (size_t k = i; k < first.size() && first[k] < base + n; k++)
    if (ring_env_[ring[k]]) *env = true;
}

int Engine::launch_placed_round(bool with_shadow, uint64_t base, uint32_t n0, uint32_t solo_job, bool env) {
  const uint32_t saved = solo_;
  const bool saved_env = any_env_;
  solo_ = solo_job == kNoSolo ? kNoSolo : 0u;  // a flag: the kernels read the slot from Ctrl
  any_env_ = env;
  KernelArgs a = args();
  const LaunchShape sh = shape();
  hipError_t e = k::round_setup(d_ctrl_, n0, solo_job == kNoSolo ? 0u : solo_job, stream_);
  if (e == hipSuccess) e = k::fill_fresh(a, base, n0, stream_);
  int rc = e == hipSuccess ? OMR_OK : hipfail(e, "round fill");
  if (!rc && tuning.round_steps) rc = launch_steps(tuning.round_steps, with_shadow, true);
  // an even number of steps leaves the live paths in queue 0
  if (!rc && profile && hipEventRecord(ev_c_, stream_) != hipSuccess) rc = hipfail(hipGetLastError(), "event");
  if (!rc && (e = k::drain(a, sh, 0, profile, stream_)) != hipSuccess) rc = hipfail(e, "drain");
  if (!rc && profile && hipEventRecord(ev_d_, stream_) != hipSuccess) rc = hipfail(hipGetLastError(), "event");
  solo_ = saved;
  any_env_ = saved_env;
  return rc;
}

int Engine::collect_profile() {
  for (size_t i = 0; i < prof_kind_.size(); i++) {
    float ms = 0.0f;
    OMR_HIP(hipEventElapsedTime(&ms, prof_ev_[2 * i], prof_ev_[2 * i + 1]));
    prof_ms_[prof_kind_[i]] += ms;
    if (prof_kind_[i] == 0) prof_extend_launches_++;
    if (prof_kind_[i] == 2) prof_shadow_launches_++;
  }
  prof_kind_.clear();
  return OMR_OK;
}

int Engine::launch_steps(uint32_t steps, bool with_shadow, bool round) {
  KernelArgs a = args();
  LaunchShape sh = shape();
  // refill_every = 0: finished slots regenerate inside shade; k > 0: they go
  // idle and a bulk refill every k steps hands out samples in pixel order.
  // In a drain round finished slots go idle until the next round.
  const bool regen = !round && tuning.refill_every == 0;
  auto refill_now = [&](uint32_t s) {
    return !round && !regen && ((s + 1) % tuning.refill_every == 0 || s + 1 == steps);
  };
  if (profile) {
    // one event pair per launch, every kernel directly on the stream
    size_t need = (size_t)steps * 3 * 2;
    while (prof_ev_.size() < need) {
      hipEvent_t e;
      OMR_HIP(hipEventCreate(&e));
      prof_ev_.push_back(e);
    }
    prof_kind_.clear();
    auto timed = [&](int kind, auto &&fn) -> hipError_t {
      size_t i = prof_kind_.size();
      hipError_t e = hipEventRecord(prof_ev_[2 * i], stream_);
      if (e != hipSuccess) return e;
      e = fn();
      if (e != hipSuccess) return e;
      prof_kind_.push_back(kind);
      return hipEventRecord(prof_ev_[2 * i + 1], stream_);
    };
    for (uint32_t s = 0; s < steps; s++) {
      int parity = (int)(s & 1u);
      hipError_t e = timed(0, [&] { return k::extend(a, sh, parity, true, stream_); });
      if (e == hipSuccess) e = timed(1, [&] { return k::shade(a, sh, parity, regen, true, stream_); });
      if (e == hipSuccess && with_shadow) e = timed(2, [&] { return k::shadow(a, sh, true, stream_); });
      if (e == hipSuccess && refill_now(s)) e = k::refill(a, sh, parity ^ 1, stream_);
      if (e != hipSuccess) return hipfail(e, "profiled bounce step launch");
    }
    return OMR_OK;
  }
  auto enqueue = [&]() -> hipError_t {
    for (uint32_t s = 0; s < steps; s++) {
      int parity = (int)(s & 1u);
      hipError_t e = k::extend(a, sh, parity, false, stream_);
      if (e != hipSuccess) return e;
      e = k::shade(a, sh, parity, regen, false, stream_);
      if (e != hipSuccess) return e;
      if (with_shadow) {
        e = k::shadow(a, sh, false, stream_);
        if (e != hipSuccess) return e;
      }
      if (refill_now(s)) {
        e = k::refill(a, sh, parity ^ 1, stream_);
        if (e != hipSuccess) return e;
      }
    }
    return hipSuccess;
  };
  if (!tuning.use_graph) {
    hipError_t e = enqueue();
    if (e != hipSuccess) return hipfail(e, "bounce step launch");
    return OMR_OK;
  }
  int gi = (with_shadow ? 1 : 0) + (a.solo != kNoSolo ? 2 : 0) + (!a.any_env ? 4 : 0);
  if (!graph_[gi] || graph_steps_ != steps || graph_round_ != round) {
    if (graph_steps_ != steps || graph_round_ != round) drop_graphs();
    hipGraph_t g = nullptr;
    OMR_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    hipError_t e = enqueue();
    hipError_t e2 = hipStreamEndCapture(stream_, &g);
    if (e != hipSuccess) return hipfail(e, "graph capture");
    if (e2 != hipSuccess) return hipfail(e2, "hipStreamEndCapture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hipfail(e, "hipGraphInstantiate");
    graph_[gi] = ex;
    graph_steps_ = steps;
    graph_round_ = round;
  }
  OMR_HIP(hipGraphLaunch(graph_[gi], stream_));
  return OMR_OK;
}

static uint64_t job_samples(const omr_camera &c, const omr_render_params &p) {
  return (uint64_t)c.width * c.height * p.supersamples * p.supersamples * (uint64_t)(p.spp_end - p.spp_begin);
}

// Adds the traced-ray counters of a second lane's control block to `c`.
static void add_counters(Ctrl &c, const Ctrl &o) {
  Sharded<unsigned long long> Ctrl::*f[] = {&Ctrl::segments,       &Ctrl::shadow_rays,    &Ctrl::node_visits,
                                            &Ctrl::prim_tests,     &Ctrl::drain_segments, &Ctrl::drain_paths,
                                            &Ctrl::drain_node_visits, &Ctrl::drain_prim_tests};
  for (auto m : f)
    for (uint32_t k = 0; k < kShards; k++) (c.*m)[k] += (o.*m)[k];
}

int Engine::run_jobs(std::vector<JobRun> &jobs, omr_render_stats *st) {
  auto t0 = std::chrono::steady_clock::now();
  // whatever the way out, lane 0 is current afterwards
  struct LaneBack {
    Engine *e;
    ~LaneBack() {
      if (e->on_alt_) e->swap_lane();
    }
  } lane_back{this};
  if (jobs.size() > kJobRing) return fail(OMR_ERR_INVALID, "too many jobs in one batch (max 1024)");
  hipSetDevice(device_);
  int rc = ensure_pool();
  if (rc) return rc;
  uint32_t need_stack = 1;
  bool any_env_is = false;
  for (auto &j : jobs) {
    SceneHost *s = scene((omr_scene_id)j.scene);
    if (!s) return fail(OMR_ERR_NOTFOUND, "unknown scene");
    need_stack = std::max(need_stack, stack_need(s));
    if (j.params.env_is && s->has_env) any_env_is = true;
    if (j.params.nee && s->emissive_specular)
      return fail(OMR_ERR_INVALID, "next-event estimation on a scene whose mirror or glass materials emit is not "
                                   "supported (the reference adds that emission when their continuation ray hits)");
  }
  LaunchShape sh = shape();
  rc = ensure_gstack(need_stack, (size_t)sh.grid * sh.wg);
  if (rc) return rc;

  std::vector<JobDev> jd(jobs.size());
  unsigned long long total = 0;
  solo_ = jobs.size() == 1 ? 0u : kNoSolo;  // job i takes ring slot i
  for (size_t i = 0; i < jobs.size(); i++) {
    JobRun &j = jobs[i];
    JobDev &D = jd[i];
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
    j.ring = (uint32_t)i;
    ring_env_[j.ring] = scene((omr_scene_id)j.scene)->has_env ? 1u : 0u;
    total += j.n_samples;
  }
  any_env_ = false;
  for (auto &j : jobs)
    if (ring_env_[j.ring]) any_env_ = true;
  Ctrl c;
  std::memset(&c, 0, sizeof(c));
  c.total_samples = total;
  c.job_lo = 0;
  c.job_hi = (uint32_t)jobs.size();
  // First fill of the cleared pool: slot i takes sample i for i < n0, with
  // the cursors and sub-queue counts set here to what the cursor grabs would
  // have left (shard k: the positions below n0); falls back to the refill
  // kernel if a sub-queue could not hold its share.
  const uint32_t n0 = (uint32_t)std::min<uint64_t>(pool_cap_, total);
  bool fresh = true;
  for (uint32_t k = 0; k < kShards; k++)
    if (shard_limit(n0, k) > pool_.qcap) fresh = false;
  if (fresh) {
    for (uint32_t k = 0; k < kShards; k++) {
      c.next_sample[k] = shard_limit(n0, k);
      c.qcount[0][k] = (uint32_t)shard_limit(n0, k);
    }
  }
  OMR_HIP(hipMemcpyAsync(d_jobs_, jd.data(), jd.size() * sizeof(JobDev), hipMemcpyHostToDevice, stream_));
  OMR_HIP(hipMemcpyAsync(d_ctrl_, &c, sizeof(Ctrl), hipMemcpyHostToDevice, stream_));
  OMR_HIP(hipMemsetAsync(d_done_, 0, kJobRing * kShards * sizeof(unsigned long long), stream_));
  // drain rounds placed by the host: every round fills the pool itself
  const bool placed = tuning.drain_rounds && fresh;
  if (fresh) {
    // slots past n0 are read only by the bulk refill (with regeneration
    // they never enter a queue; placed rounds fill the slots they use)
    if (n0 < pool_cap_ && !tuning.drain_rounds && tuning.refill_every != 0)
      OMR_HIP(hipMemsetAsync(pool_.misc + 4ull * n0, 0, (size_t)(pool_cap_ - n0) * 16, stream_));
  } else {
    OMR_HIP(hipMemsetAsync(pool_.misc, 0, (size_t)pool_cap_ * 16, stream_));
  }
  for (auto &j : jobs) {
    // with per-sample records the gather writes every pixel's sums
    if (j.d_recs) OMR_HIP(hipMemsetAsync(j.d_recs, 0, (size_t)j.n_samples * 24, stream_));
    else OMR_HIP(hipMemsetAsync(j.d_accum, 0, (size_t)j.cam.width * j.cam.height * 3 * 8, stream_));
  }
  // two lanes: placed rounds alternate between the pools on two streams, for
  // a single job of more than one round (c2: +2-5 %; a batch's rounds side by
  // side measured 4 % slower on c3, rounds of different scenes sharing the
  // caches; not in profiled runs, which time each launch on one stream, nor
  // while serving, whose HBM budget holds one pool)
  const bool dual =
      placed && tuning.lanes == 2 && !profile && total > pool_cap_ && jobs.size() == 1 && !server_on_.load();
  if (dual) {
    // lane 1: its pool, stacks and control block (the same start as lane
    // 0's, copied from pinned memory on lane 0's stream so the host does not
    // wait), behind lane 0's set-up
    swap_lane();
    if ((rc = ensure_pool()) || (rc = ensure_gstack(need_stack, (size_t)sh.grid * sh.wg))) return rc;
    swap_lane();
    if (pool_.qcap != alt_.pool.qcap) return fail(OMR_ERR_STATE, "the two lanes' pools differ");
    std::memcpy(alt_.h_ctrl, &c, sizeof(Ctrl));
    OMR_HIP(hipMemcpyAsync(alt_.d_ctrl, alt_.h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, stream_));
    OMR_HIP(hipEventRecord(ev_c_, stream_));
    OMR_HIP(hipStreamWaitEvent(alt_.stream, ev_c_, 0));
  }
  OMR_HIP(hipEventRecord(ev_a_, stream_));
  KernelArgs a = args();
  hipError_t e = hipSuccess;
  if (!placed) {
    e = fresh ? k::fill_fresh(a, 0, n0, stream_) : k::refill(a, sh, 0, stream_);
    if (e != hipSuccess) return hipfail(e, "refill");
  }
  uint64_t steps_done = 0;
  const uint32_t chunk = tuning.steps_per_sync;
  uint64_t drained = 0, drain_launches = 0;
  double drain_ms = 0.0;
  prof_ms_[0] = prof_ms_[1] = prof_ms_[2] = 0.0;
  prof_extend_launches_ = prof_shadow_launches_ = 0;
  if (tuning.drain_rounds) {
    // Rounds: every slot holds a sample after the fill, the drain runs them
    // all to their end (and marks the slots idle), the next round fills the
    // pool again.  The host places each round on a range of the global
    // sample order (plan_round), so nothing is read back until the end and a
    // round that lies inside one job runs the single-job kernels; with a pool
    // whose full fill does not fit the sub-queues the refill hands out
    // min(pool, remaining) samples per round instead.
    const uint64_t rounds = (total + pool_cap_ - 1) / pool_cap_;
    std::vector<uint64_t> jfirst, jcount;
    std::vector<uint32_t> jring;
    for (size_t i = 0, b = 0; i < jobs.size(); b += jobs[i].n_samples, i++) {
      jfirst.push_back(b);
      jcount.push_back(jobs[i].n_samples);
      jring.push_back(jobs[i].ring);
    }
    // placed rounds: the host gives each round a range of the global sample
    // order; a round inside one job runs the single-job kernels
    for (uint64_t r = 0, base = 0; placed ? base < total : r < rounds; r++) {
      if (placed) {
        if (dual && r > 0) swap_lane();
        uint32_t n_r = 0, solo_job = kNoSolo;
        bool env = true;
        plan_round(jfirst, jcount, jring, base, total, &n_r, &solo_job, &env);
        rc = launch_placed_round(any_env_is, base, n_r, solo_job, env);
        base += n_r;
      } else {
        rc = launch_round(r == 0, any_env_is);
      }
      if (rc) return rc;
      steps_done += tuning.round_steps;
      drain_launches++;
      if (profile) {
        OMR_HIP(hipEventSynchronize(ev_d_));
        if ((rc = collect_profile())) return rc;
        float ms = 0.0f;
        OMR_HIP(hipEventElapsedTime(&ms, ev_c_, ev_d_));
        drain_ms += ms;
      }
    }
    if (dual) {
      // lane 1's rounds end before anything after them on lane 0's stream
      if (on_alt_) swap_lane();
      OMR_HIP(hipMemcpyAsync(alt_.h_ctrl, alt_.d_ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, alt_.stream));
      OMR_HIP(hipEventRecord(ev_d_, alt_.stream));
      OMR_HIP(hipStreamWaitEvent(stream_, ev_d_, 0));
    }
    OMR_HIP(hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
    OMR_HIP(hipStreamSynchronize(stream_));
    if (dual) add_counters(*h_ctrl_, *alt_.h_ctrl);
    drained = sum_shards(h_ctrl_->drain_paths);
  }
  for (uint64_t it = 0; !tuning.drain_rounds; it++) {
    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (it > 2000000ull || el > 600.0) return fail(OMR_ERR_STATE, "bounce loop did not terminate");
    rc = launch_steps(chunk, any_env_is);
    if (rc) return rc;
    steps_done += chunk;
    OMR_HIP(hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
    OMR_HIP(hipStreamSynchronize(stream_));
    if (profile && (rc = collect_profile())) return rc;
    const uint64_t queued = queue_len(*h_ctrl_, 0);
    const bool exhausted = samples_exhausted(*h_ctrl_);
    if (queued == 0 && exhausted) break;
    if (exhausted && queued <= tuning.drain_below) {
      // the live paths (queue 0 after an even number of steps) run to the end in one launch
      drained = queued;
      if (profile) OMR_HIP(hipEventRecord(ev_c_, stream_));
      e = k::drain(a, sh, 0, profile, stream_);
      if (e != hipSuccess) return hipfail(e, "drain");
      drain_launches++;
      if (profile) OMR_HIP(hipEventRecord(ev_d_, stream_));
      OMR_HIP(hipMemcpyAsync(h_ctrl_, d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
      OMR_HIP(hipStreamSynchronize(stream_));
      if (profile) {
        float ms = 0.0f;
        OMR_HIP(hipEventElapsedTime(&ms, ev_c_, ev_d_));
        drain_ms = ms;
      }
      break;
    }
  }
  for (auto &j : jobs) {
    if (!j.d_recs) continue;
    e = k::gather_records(j.d_recs, j.params.supersamples * j.params.supersamples * (j.params.spp_end - j.params.spp_begin),
                          j.cam.width * j.cam.height, j.d_accum, stream_);
    if (e != hipSuccess) return hipfail(e, "gather_records");
  }
  OMR_HIP(hipEventRecord(ev_b_, stream_));
  for (auto &j : jobs) {
    if (j.d_out) {
      double nspp = (double)j.params.supersamples * j.params.supersamples * (double)(j.params.spp_end - j.params.spp_begin);
      e = k::resolve(j.d_accum, j.d_out, j.cam.width * j.cam.height, nspp, stream_);
      if (e != hipSuccess) return hipfail(e, "resolve");
    }
  }
  OMR_HIP(hipMemcpyAsync(h_done_, d_done_, kJobRing * kShards * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                         stream_));
  OMR_HIP(hipStreamSynchronize(stream_));
  for (size_t i = 0; i < jobs.size(); i++) {
    unsigned long long done = job_done(h_done_, (uint32_t)i);
    if (done != jobs[i].n_samples)
      return fail(OMR_ERR_STATE, "job " + std::to_string(i) + " completed " + std::to_string(done) + " of " +
                                     std::to_string(jobs[i].n_samples) + " samples");
    jobs[i].done = true;
  }
  if (st) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, ev_a_, ev_b_);
    st->segments = sum_shards(h_ctrl_->segments);
    st->shadow_rays = sum_shards(h_ctrl_->shadow_rays);
    st->bounce_steps = steps_done;
    st->gpu_ms = ms;
    st->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st->extend_ms = prof_ms_[0];
    st->shade_ms = prof_ms_[1];
    st->shadow_ms = prof_ms_[2];
    st->extend_launches = prof_extend_launches_;
    st->shadow_launches = prof_shadow_launches_;
    st->node_visits = sum_shards(h_ctrl_->node_visits);
    st->prim_tests = sum_shards(h_ctrl_->prim_tests);
    st->drained_paths = drained;
    st->drain_segments = sum_shards(h_ctrl_->drain_segments);
    st->drain_ms = drain_ms;
    st->drain_node_visits = sum_shards(h_ctrl_->drain_node_visits);
    st->drain_prim_tests = sum_shards(h_ctrl_->drain_prim_tests);
    st->drain_launches = drain_launches;
  }
  return OMR_OK;
}

uint64_t job_samples_u64(const omr_camera &c, const omr_render_params &p) { return job_samples(c, p); }

}  // namespace omr
