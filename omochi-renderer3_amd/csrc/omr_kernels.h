// omr_kernels.h -- host-side declarations of the HIP kernel launchers
// (implemented in omr_kernels.hip).  Host code sees plain pointers only.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "omr_layout.h"

namespace omr {

// Entries of the per-lane LDS traversal stack (8 or 16, a launch choice);
// deeper entries spill to the per-thread global area (PoolPtrs::gstack, gcap
// entries per thread).
constexpr uint32_t kShortStackMin = 8, kShortStackMax = 16;
// Slots one wave of the refill kernel covers per trip (4 per lane).
constexpr uint32_t kRefillSlots = 256;

// Device pointers of the wavefront path pool (capacity `pool` slots).
struct PoolPtrs {
  float *ray_o;        // float4: o.xyz, beta.x
  float *ray_d;        // float4: d.xyz, beta.y
  uint32_t *misc;      // uint4: beta.z bits, state word, sample index in job, job ring slot
  uint32_t *rng;       // uint4: xorshift128 state
  float *hit;          // float4: t, prim id bits, u, v
  float *pdf;          // float: MIS pdf of the last Lambert bounce
  float *pend;         // 3 pending branches x pool x 3 float4
  uint32_t *queue[2];  // ping-pong ray queues: kShards sub-queues of qcap entries
  float *shadow;       // shadow queue: kShards sub-queues of qcap records (3 float4)
  uint32_t qcap;       // capacity of one sub-queue
  uint2 *gstack;       // per-thread overflow traversal stacks
  uint32_t pool;
  uint32_t gcap;       // entries per thread in gstack
  uint32_t df_budget;     // dynamic fetch: traversal visits between fetches (0 = grid-stride)
  uint32_t df_fetch_min;  // dynamic fetch: idle lanes of a wave that trigger a fetch
  uint32_t stack_short;   // LDS traversal stack entries per lane (8 or 16)
  uint32_t node_width;    // BVH nodes the traversal reads: 4 (SceneDev::nodes) or 8 (SceneDev::nodes8)
};

// KernelArgs::solo when a run may hold several jobs
constexpr uint32_t kNoSolo = 0xFFFFFFFFu;

struct KernelArgs {
  Ctrl *ctrl;
  JobDev *jobs;          // ring of kJobRing entries
  unsigned long long *done;  // completed-sample counters, done_index(shard, ring)
  const SceneDev *scenes;
  PoolPtrs p;
  uint32_t solo;         // kNoSolo, or single-job kernels (the job's ring slot is Ctrl::solo_job)
  uint32_t any_env;      // 0 when no job of the launch has a scene with an environment map
};

struct LaunchShape {
  uint32_t grid;        // workgroups of the persistent kernels
  uint32_t wg;          // threads per workgroup: 64, 128 or 256
  uint32_t shade_grid;       // workgroups of the shade kernel: five waves per SIMD
  uint32_t shade6_grid;  // ... of its single-job form for a scene without an environment map: six
};

// Working memory of the allocation-free closest-hit entry point: a fixed
// device array per GPU (held until the process ends).  The first
// kNoAllocHeadBytes hold the dynamic-fetch heads (cleared by a fill kernel at
// every call), the rest the per-thread overflow traversal stacks.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocHeadBytes = sizeof(Sharded<uint32_t>);
static_assert(kNoAllocHeadBytes % 256 == 0, "the stacks start on a 256-byte boundary");

namespace k {
// Device address of the allocation-free scratch array on the current device.
hipError_t noalloc_scratch(void **ptr);
// p[0..n) = 0 (a kernel, not a memset: callers may capture it in a graph).
hipError_t zero_words(uint32_t *p, uint32_t n, hipStream_t st);
// Fill of an idle pool (queue 0): slot i takes sample base + i for i < n0,
// without atomics; the queue-0 counts of shard k are shard_limit(n0, k)
// (set by the host or by round_setup).
hipError_t fill_fresh(const KernelArgs &a, unsigned long long base, uint32_t n0, hipStream_t st);
// Before a host-placed round of n0 samples: queue-0 counts for fill_fresh,
// zero fetch heads, and the single-job launches' job slot.
hipError_t round_setup(Ctrl *ctrl, uint32_t n0, uint32_t solo_job, hipStream_t st);
hipError_t refill(const KernelArgs &a, const LaunchShape &s, int parity, hipStream_t st);
// stats: count node visits / primitive tests into Ctrl (profiling runs)
hipError_t extend(const KernelArgs &a, const LaunchShape &s, int parity, bool stats, hipStream_t st);
// regen: finished slots take the next sample inline; otherwise they go idle
// until the next refill.
// tag: a profiled step (single-job runs at 64-thread groups launch a
// separately named copy of the kernel, so profiler traces tell them apart)
hipError_t shade(const KernelArgs &a, const LaunchShape &s, int parity, bool regen, bool tag, hipStream_t st);
hipError_t shadow(const KernelArgs &a, const LaunchShape &s, bool tag, hipStream_t st);
// Runs every path of ray queue `parity` to its end in one launch (closest
// hit, shading, visibility, pending branches); only once no samples are left.
hipError_t drain(const KernelArgs &a, const LaunchShape &s, int parity, bool stats, hipStream_t st);
// Zeroes the completed-sample counters of ring slot `ring` (every shard).
hipError_t clear_done(unsigned long long *done, uint32_t ring, hipStream_t st);
// Served mode, between chunks of bounce steps (stream-ordered, no kernel of
// the run in flight): admits the samples [old_total, new_total) -- a cursor
// that ran past its shard's end of the old range moves back to that end, so
// the new samples start exactly at old_total -- and sets the live job window
// [lo, hi).
hipError_t admit(Ctrl *ctrl, unsigned long long old_total, unsigned long long new_total, uint32_t lo, uint32_t hi,
                 hipStream_t st);
// Served mode: the ServeStatus of the run after the work before it on the
// stream, for the n live jobs starting at logical ring index lo.
hipError_t serve_status(const Ctrl *ctrl, const unsigned long long *done, uint32_t lo, uint32_t n, ServeStatus *out,
                        hipStream_t st);
// accum[3p..3p+2] = sums of the records of pixel p's spp_px samples
hipError_t gather_records(const unsigned long long *recs, uint32_t spp_px, uint32_t n_px, unsigned long long *accum,
                          hipStream_t st);
hipError_t resolve(const unsigned long long *accum, float *out, uint32_t n_px, double n_samples, hipStream_t st);
// Closest hit for n rays (n x 8 floats).  counters (optional, device) gets
// node visits and primitive tests.
// head: zeroed fetch heads (used when budget > 0, dynamic fetch).
hipError_t trace_rays(const SceneDev *scene, const float *rays, uint32_t n, int32_t *prim, float *t,
                      unsigned long long *counters, uint2 *gstack, uint32_t gcap, uint32_t stack_short,
                      uint32_t node_width, uint32_t grid, uint32_t wg, Sharded<uint32_t> *head, int budget,
                      uint32_t fetch_min, hipStream_t st);
// Environment importance-sampling tables: weights, per-row scan, marginal scan.
hipError_t env_tables(const float *env, uint32_t w, uint32_t h, float *wbuf, unsigned int *wmax_bits,
                      unsigned long long *q, unsigned long long *row, unsigned long long *marg, hipStream_t st);
// Bucket directories of the row and marginal inversions (after env_tables).
hipError_t env_guides(const unsigned long long *row, const unsigned long long *marg, uint32_t w, uint32_t h,
                      uint2_t *rguide, uint2_t *mguide, double *rscale, double *mscale, hipStream_t st);
// Texel records for shading (after env_tables): radiance and quantized weight
// of n texels, 16 bytes each.
hipError_t env_pack(const float *env, const unsigned long long *q, uint32_t n, float4 *env4, hipStream_t st);
}  // namespace k
}  // namespace omr
