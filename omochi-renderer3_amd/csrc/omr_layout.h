// omr_layout.h -- data layout shared by the host runtime (g++) and the HIP
// kernels (hipcc).  Plain structs only; every device-side record is a whole
// number of 16-byte vectors so a lane fetches it with dwordx4 loads.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define OMR_HD __host__ __device__
#else
#define OMR_HD
#endif

namespace omr {

// Binary32 constants of the reference (src/tools/Constant.h:4-8).
constexpr float kInf = 0x1.fffffep+126f;  // (FLT_MAX-1)/2
constexpr float kEps = 0.0005f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kTwoPi = 6.28318530717958647692f;
constexpr float kInvPi = 0.31830988618379067154f;
// Fixed-point accumulation: 2^24 units per radiance unit, per-event clamp.
constexpr float kFixScale = 16777216.0f;
constexpr float kFixClamp = 1.0e8f;
// Slab-test far-distance inflation, 1 + 2^-20 (conservative box test).
constexpr float kSlabInflate = 1.00000095f;

// 4-wide BVH node, 128 bytes = one cache line.  Children bounds stored
// component-major so one float4 holds the same bound of all four children.
// child: >= 0 inner node index; -1 empty; leaf = 0x80000000 | (count-1)<<27 | first.
struct alignas(128) WNode4 {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t meta[4];  // meta[0] = depth of this node (diagnostics)
};
static_assert(sizeof(WNode4) == 128, "node must be one cache line");

// 8-wide BVH node, 256 bytes = two cache lines, the same component-major
// layout over eight children (a bound of children 0-3, then of 4-7, is one
// float4 each).
struct alignas(256) WNode8 {
  float lox[8], hix[8], loy[8], hiy[8], loz[8], hiz[8];
  int32_t child[8];
  int32_t meta[8];  // meta[0] = depth of this node (diagnostics)
};
static_assert(sizeof(WNode8) == 256, "8-wide node is two cache lines");

OMR_HD inline bool child_is_leaf(int32_t c) { return ((uint32_t)c & 0xC0000000u) == 0x80000000u; }
OMR_HD inline uint32_t leaf_first(int32_t c) { return (uint32_t)c & 0x07FFFFFFu; }
OMR_HD inline uint32_t leaf_count(int32_t c) { return (((uint32_t)c >> 27) & 7u) + 1u; }
OMR_HD inline int32_t make_leaf(uint32_t first, uint32_t count) {
  return (int32_t)(0x80000000u | ((count - 1u) << 27) | first);
}
constexpr uint32_t kMaxLeafPrims = 8;
constexpr uint32_t kMaxLeafFirst = 0x07FFFFFFu;

// Primitive record in leaf order, 48 bytes.
//  triangle: a = v0.xyz | id,  b = e1.xyz | kind 0,  c = e2.xyz | 0
//  sphere:   a = c.xyz  | id,  b = r,0,0  | kind 1,  c = 0
struct alignas(16) PrimRec {
  float a[4], b[4], c[4];
};

// Material record, 48 bytes (Material, src/renderer/Material.h:11-36, plus the
// normalized multi-material rate, src/renderer/SceneObject.h:17-30).
struct alignas(16) MatRec {
  uint32_t type;
  float color[3];
  float emission[3];
  float ior;
  float rate;
  uint32_t tex;  // 0 = none, k = texture k-1
  uint32_t pad[2];
};

// Texture directory entry: index of the first texel and the size.
struct TexRec {
  uint32_t first, width, height, pad;
};

// Two 32-bit words, one 8-byte load (an entry range of a bucket directory).
struct alignas(8) uint2_t {
  uint32_t lo, hi;
};

// Device-side view of one resident scene.
struct SceneDev {
  const WNode4 *nodes;
  const WNode8 *nodes8;        // the same tree collapsed to 8-wide nodes (same leaves)
  const PrimRec *prims;        // leaf order
  const float *tri_n;          // per original triangle: 12 floats (n0,pm, dn1,_, dn2,_), pm = its prim_mat word
  const float *sph;            // per original sphere: 4 floats (c, r)
  const uint32_t *prim_mat;    // per original primitive: first | (count-1) << 24
  const MatRec *mats;
  const float *env;            // env_h x env_w x 3 or null
  const unsigned long long *env_q;     // env_h x env_w
  // env_h x env_w texel records for shading, 4 words each: radiance r, g, b
  // and the quantized weight q (< 2^32) as the fourth word, so the radiance
  // and the pdf of a texel come with one 16-byte load
  const float *env4;
  const unsigned long long *env_row;   // env_h x (env_w+1)
  const unsigned long long *env_marg;  // env_h+1
  // bucket directories of the inversions: for key k of a table with total T
  // and n entries, bucket floor(k n / T) (binary64) holds the first and last
  // entry any key of the bucket can select
  const uint2_t *env_rguide;           // env_h x env_w
  const uint2_t *env_mguide;           // env_h
  const double *env_rscale;            // env_h: env_w / row total (0 for an empty row)
  // 2: env_h / total, and the pdf factor (w h) / (2 pi^2 total) (0 without weights)
  const double *env_mscale;
  uint32_t n_tri, n_sph;
  const float *planes;         // infinite floors y = planes[i]; ids n_tri + n_sph + i
  uint32_t n_planes;
  const float *tri_uv;         // per original triangle: 8 floats (uv0, duv1, duv2, pad) or null
  const float *tex;            // texels of every texture: 4 floats each (linear RGB, pad)
  const TexRec *tex_rec;       // n_tex entries
  uint32_t n_tex;
  uint32_t env_w, env_h;
  uint32_t n_nodes;
  uint32_t stack_bound;        // max traversal stack entries (4-wide nodes)
  uint32_t n_nodes8;
  uint32_t stack_bound8;       // max traversal stack entries (8-wide nodes)
  // largest ray-origin coordinate magnitude whose rounding in a fused slab
  // distance (2^-24 of it) the child boxes' padding covers eight times over
  float o_far;
};

// One render job in the device job ring.
struct JobDev {
  unsigned long long sample_base;  // global index of the job's first sample
  uint32_t n_samples;              // W*H*ss^2*(spp_end-spp_begin)
  uint32_t width, height, ss;
  uint32_t nspp, spp_begin;
  uint32_t max_depth, rr_depth;
  uint32_t flags;                  // bit0: environment importance sampling; bit1: next-event estimation mode
  uint32_t scene;
  unsigned long long seed;
  float cam_pos[3], cam_c[3], cam_x[3], cam_y[3];
  float aperture, focal;
  unsigned long long *accum;       // W*H*3 fixed-point sums
  // n_samples*3 fixed-point sums, one record per sample, or null.  With
  // records every contribution goes to its sample's record by a plain
  // read-modify-write (a sample has one owner at a time) and a gather kernel
  // sums each pixel's records into accum at the end; without them
  // contributions go to accum by atomics.  Integer sums: the same totals.
  unsigned long long *recs;
};

// Every device-wide counter that waves update in the hot loop is split into
// kShards copies; wave w uses copy w % kShards.  Device-scope atomics on one
// cache line are serialized at the memory side (tens of ns each), so each copy
// sits on its own kLine-byte line: copies packed into one line would still
// serialize.
constexpr uint32_t kShards = 64;
constexpr uint32_t kLine = 256;

template <class T>
struct Sharded {
  struct alignas(kLine) Cell {
    T v;
  };
  Cell c[kShards];
  OMR_HD T &operator[](uint32_t k) { return c[k].v; }
  OMR_HD const T &operator[](uint32_t k) const { return c[k].v; }
};

// Cursor shard k hands out runs of kSampleRun consecutive samples: its j-th
// sample is in run j / kSampleRun of the shard, and run q of shard k starts
// at (q kShards + k) kSampleRun.  A wave that takes 64 cursor positions gets
// consecutive samples (the sub-samples of one pixel, then its neighbours):
// rays that start alike and per-sample sums that share cache lines.
constexpr unsigned long long kSampleRun = 64;
OMR_HD inline unsigned long long shard_sample(unsigned long long j, uint32_t k) {
  return (j / kSampleRun) * (kShards * kSampleRun) + (unsigned long long)k * kSampleRun + j % kSampleRun;
}
// Number of cursor positions of shard k whose samples are below `total`
// (shard_sample is increasing in j).
OMR_HD inline unsigned long long shard_limit(unsigned long long total, uint32_t k) {
  const unsigned long long round = kShards * kSampleRun, rem = total % round, lo = (unsigned long long)k * kSampleRun;
  const unsigned long long part = rem > lo ? (rem - lo < kSampleRun ? rem - lo : kSampleRun) : 0ull;
  return total / round * kSampleRun + part;
}

// Control block in device memory (one per engine).
struct Ctrl {
  Sharded<unsigned long long> next_sample;                // sample cursors (position j -> sample shard_sample(j, k))
  alignas(kLine) unsigned long long total_samples;        // samples admitted so far (host-written)
  alignas(kLine) uint32_t job_lo;                         // live job window [job_lo, job_hi) (logical ring
  uint32_t job_hi;                                        // indices); the pair is copied as 8 bytes
  uint32_t solo_job;                                      // ring slot of the only job of single-job launches
  Sharded<unsigned long long> segments;     // closest-hit queries traced
  Sharded<unsigned long long> shadow_rays;  // environment shadow rays traced
  Sharded<unsigned long long> node_visits;  // stats kernels only
  Sharded<unsigned long long> prim_tests;   // stats kernels only
  Sharded<unsigned long long> drain_segments;  // closest-hit queries traced by the drain kernel
  Sharded<unsigned long long> drain_paths;     // paths the drain kernel took from its queue
  Sharded<unsigned long long> drain_node_visits;  // drain, stats variant only
  Sharded<unsigned long long> drain_prim_tests;   // drain, stats variant only
  Sharded<uint32_t> qcount[2];              // ping-pong ray sub-queue counts
  Sharded<uint32_t> scount;                 // shadow sub-queue counts
  Sharded<uint32_t> qhead[2];               // dynamic-fetch heads of the ray sub-queues
  Sharded<uint32_t> shead;                  // dynamic-fetch heads of the shadow sub-queues
};

constexpr uint32_t kJobRing = 1024;
// Resident scenes per context (slots of the device scene table).
constexpr uint32_t kMaxSceneSlots = 64;
// Completed-sample counters: kShards x kJobRing, shard-major, so the waves of
// one shard that finish samples of one job update one line.
OMR_HD inline size_t done_index(uint32_t shard, uint32_t ring) { return (size_t)shard * kJobRing + ring; }

// What the served loop reads back after each chunk of bounce steps: the
// length of ray queue 0, whether every admitted sample has been handed out,
// the traced-ray totals, and the completed samples of each live job (entry i:
// ring slot (lo + i) % kJobRing).  Written by one small kernel, so a chunk
// costs one copy of 32 + 8 x live-jobs bytes instead of the whole control
// block and counter array.
struct ServeStatus {
  unsigned long long queued;
  unsigned long long exhausted;
  unsigned long long segments, shadow_rays;
  unsigned long long done[kJobRing];
};

// Per-slot packed path state word (misc.y).
//  bits 0-7 depth, 8-15 branch, 16-17 pending count, 18 prev-lambert (MIS), 19 active
OMR_HD inline uint32_t ps_depth(uint32_t w) { return w & 0xFFu; }
OMR_HD inline uint32_t ps_branch(uint32_t w) { return (w >> 8) & 0xFFu; }
OMR_HD inline uint32_t ps_pending(uint32_t w) { return (w >> 16) & 3u; }
OMR_HD inline uint32_t ps_mis(uint32_t w) { return (w >> 18) & 1u; }
OMR_HD inline uint32_t ps_pack(uint32_t depth, uint32_t branch, uint32_t pend, uint32_t mis) {
  return depth | (branch << 8) | (pend << 16) | (mis << 18) | (1u << 19);
}

}  // namespace omr
