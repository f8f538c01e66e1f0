// omr_kernels.hip -- wavefront path-tracing kernels for gfx950 (CDNA4).
//
// Stages of one bounce step (all persistent, grid-stride over device-side
// queue counts so a fixed sequence can be captured into a hipGraph):
//   refill : idle pool slots take new samples (camera rays)      PathTracer.cpp:143-167
//   extend : closest hit through the 4-wide BVH                  Scene.cpp:51-85, QBVH.cpp:30-162
//   shade  : material roulette, Russian roulette, Lambert /
//            mirror / glass, emission and environment           PathTracer.cpp:180-591
//            accumulation, path regeneration, queue compaction
//   shadow : any-hit visibility for environment light samples   (no reference counterpart)
// Queues are compacted with a wave ballot + mbcnt prefix and one atomic per
// wave; traversal stacks live in LDS (circular short stack) with a per-thread
// global overflow area.
//
// Built with -ffp-contract=off: every floating-point expression below is
// evaluated in the order written so results match the CPU restatement bit
// for bit.
#include <hip/hip_runtime.h>

#include "omr_kernels.h"

namespace omr {
namespace {

// ------------------------------------------------------------ vector math
struct f3 {
  float x, y, z;
};
__device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline f3 scl(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
// Vector3::dot / cross / normalize (src/tools/Vector.h:66-86)
__device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 cross(f3 a, f3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// normalize: one correctly rounded reciprocal of the length and three
// products (the reference divides each component by the length; the two
// differ by at most an ulp per component, and the CPU restatement uses the
// same form)
__device__ inline f3 normalize(f3 a) {
  float inv = 1.0f / sqrtf(dot(a, a));
  return mk(a.x * inv, a.y * inv, a.z * inv);
}
__device__ inline f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

// 16-byte load through the global address space.  Scene pointers are read
// from device memory, so the compiler cannot prove where they point and would
// emit flat loads (which also count against the LDS wait counter).
typedef float v4f __attribute__((ext_vector_type(4)));
__device__ inline float4 ld4(const float4 *p) {
  v4f v = *(__attribute__((address_space(1))) const v4f *)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
// The same view for the other records that scene and job descriptors point
// to, and for the pixel-sum atomics: flat atomics in flight would also hold
// up every later LDS and scalar-load wait of the wave.
template <class T>
__device__ inline const __attribute__((address_space(1))) T *gptr(const T *p) {
  return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ inline T gld(const T *p) {
  return *gptr(p);
}
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
__device__ inline TexRec gld(const TexRec *p) {
  v4u v = *(__attribute__((address_space(1))) const v4u *)p;
  return TexRec{v.x, v.y, v.z, v.w};
}
__device__ inline uint2_t gld(const uint2_t *p) {
  v2u v = *(__attribute__((address_space(1))) const v2u *)p;
  return uint2_t{v.x, v.y};
}
// Copy of a job or scene record read with scalar loads (p wave-uniform, in
// uniform control flow): the fields then sit in scalar registers.
template <class T>
__device__ inline T cload(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "records are whole words");
  const __attribute__((address_space(4))) uint32_t *q = (const __attribute__((address_space(4))) uint32_t *)p;
  uint32_t w[sizeof(T) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); i++) w[i] = q[i];
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

__device__ inline void gadd(unsigned long long *p, unsigned long long v) {
  __hip_atomic_fetch_add((__attribute__((address_space(1))) unsigned long long *)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------ wave helpers
__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ inline uint32_t rank_in(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Wave-aggregated append: returns this lane's position in the queue (valid
// when pred).  Must be reached by every active lane of the wave.
__device__ inline uint32_t wave_append(uint32_t *counter, bool pred) {
  unsigned long long m = __ballot(pred);
  if (m == 0ull) return 0u;
  uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, (int)leader);
  return base + rank_in(m);
}

// Shard of the calling wave (global wave index mod kShards).
// (read through the first lane: the value is uniform across the wave, and as
// a scalar the counter addresses built from it stay out of vector registers)
__device__ inline uint32_t wave_shard() {
  uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)(w % kShards));
}

// A queue is kShards sub-queues of `cap` entries.  s_pre (LDS, kShards+1)
// receives the exclusive prefix of the sub-queue counts; the flat index i of
// an entry maps to its sub-queue by binary search.  Whole block must call.
__device__ inline void load_prefix(uint32_t *s_pre, const Sharded<uint32_t> &counts) {
  if (threadIdx.x < kShards) s_pre[threadIdx.x + 1] = counts[threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    s_pre[0] = 0u;
    for (uint32_t k = 1; k <= kShards; k++) s_pre[k] += s_pre[k - 1];
  }
  __syncthreads();
}
__device__ inline uint32_t flat_slot(const uint32_t *s_pre, uint32_t i, uint32_t cap) {
  uint32_t lo = 0, hi = kShards;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (s_pre[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo * cap + (i - s_pre[lo]);
}

// ------------------------------------------------------------ RNG
// Random (src/tools/Random.h:9-29): MT-style seeding recurrence + xorshift128.
__device__ inline uint4 rng_init(uint32_t seed) {
  uint32_t s = seed, st[4];
#pragma unroll
  for (uint32_t i = 1; i <= 4; i++) {
    s = 1812433253u * (s ^ (s >> 30)) + i;
    st[i - 1] = s;
  }
  return make_uint4(st[0], st[1], st[2], st[3]);
}
__device__ inline uint32_t rng_next(uint4 &st) {
  uint32_t t = st.x ^ (st.x << 11);
  st.x = st.y;
  st.y = st.z;
  st.z = st.w;
  st.w = (st.w ^ (st.w >> 19)) ^ (t ^ (t >> 8));
  return st.w;
}
__device__ inline float rng_float(uint4 &st) { return (float)rng_next(st) / 4294967295.0f; }

__device__ inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
__device__ inline uint32_t combine(uint32_t h, uint32_t v) { return mix32(h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2))); }
// Per-path seed: hash(seed, pixel, sub-pixel, sample number, branch).
__device__ inline uint32_t path_seed(unsigned long long seed, uint32_t pixel, uint32_t sub, uint32_t sample,
                                     uint32_t branch) {
  uint32_t h = mix32((uint32_t)seed ^ 0x6f6d7233u);
  h = combine(h, (uint32_t)(seed >> 32));
  h = combine(h, pixel);
  h = combine(h, sub);
  h = combine(h, sample);
  h = combine(h, branch);
  return h;
}

// ------------------------------------------------------------ elementary functions
// sin/cos of 2*pi*x: quadrant reduction, Taylor polynomials on [-pi/4, pi/4].
__device__ inline void sincos_turns(float x, float &s_out, float &c_out) {
  float q = floorf(x * 4.0f + 0.5f);
  float f = x - q * 0.25f;
  float a = f * kTwoPi;
  float a2 = a * a;
  float sp = a + a * a2 * (-0.16666667f + a2 * (0.0083333333f + a2 * (-0.00019841270f + a2 * 2.7557319e-6f)));
  float cp = 1.0f + a2 * (-0.5f + a2 * (0.041666668f + a2 * (-0.0013888889f + a2 * 2.4801587e-5f)));
  int qi = ((int)q) & 3;
  float s, c;
  if (qi == 0) { s = sp; c = cp; }
  else if (qi == 1) { s = cp; c = -sp; }
  else if (qi == 2) { s = -sp; c = -cp; }
  else { s = -cp; c = sp; }
  s_out = s;
  c_out = c;
}
__device__ inline float asin_small(float z) {
  float z2 = z * z;
  float p = 0.0083903358f;
  p = 0.0097616095f + z2 * p;
  p = 0.011551800f + z2 * p;
  p = 0.013964844f + z2 * p;
  p = 0.017352764f + z2 * p;
  p = 0.022372159f + z2 * p;
  p = 0.030381944f + z2 * p;
  p = 0.044642858f + z2 * p;
  p = 0.075f + z2 * p;
  p = 0.16666667f + z2 * p;
  return z + z * z2 * p;
}
__device__ inline float acos_f(float x) {
  if (!(x > -1.0f)) x = -1.0f;
  if (x > 1.0f) x = 1.0f;
  float ax = fabsf(x);
  float r;
  if (ax <= 0.5f) {
    r = 1.5707964f - asin_small(ax);
    if (x < 0.0f) r = kPi - r;
  } else {
    float s = sqrtf((1.0f - ax) * 0.5f);
    r = 2.0f * asin_small(s);
    if (x < 0.0f) r = kPi - r;
  }
  return r;
}

// ------------------------------------------------------------ primitives
// Polygon::CheckIntersection (src/renderer/Polygon.h:63-106), single-sided.
// u_out, v_out are the unnormalized barycentrics and det_out their divisor:
// the traversal divides once, for the closest hit only (u/det is the same
// value whenever it is computed).
__device__ inline bool hit_triangle(f3 o, f3 d, f3 p0, f3 e1, f3 e2, float &t_out, float &u_out, float &v_out,
                                    float &det_out) {
  f3 P = cross(d, e2);
  float det = dot(P, e1);
  if (!(det > kEps)) return false;
  f3 T = sub(o, p0);
  float u = dot(P, T);
  if (!(u >= 0.0f && u <= det)) return false;
  f3 Q = cross(T, e1);
  float v = dot(Q, d);
  if (!(v >= 0.0f && u + v <= det)) return false;
  float t = dot(Q, e2) / det;
  if (!(t >= kEps)) return false;
  t_out = t;
  u_out = u;
  v_out = v;
  det_out = det;
  return true;
}
// Sphere::CheckIntersection (src/renderer/Sphere.h:21-48) with the
// cancellation-free discriminant r^2 - |(c-x) - b v|^2 (|v| = 1).
__device__ inline bool hit_sphere(f3 x, f3 v, f3 c, float r, float &t_out) {
  f3 oc = sub(c, x);
  float b = dot(v, oc);
  f3 perp = sub(oc, scl(v, b));
  float D = r * r - dot(perp, perp);
  if (D < 0.0f) return false;
  float sD = sqrtf(D);
  float t1 = sD + b;
  float t2 = -sD + b;
  if (t1 <= kEps && t2 <= kEps) return false;
  t_out = (t2 > kEps) ? t2 : t1;
  return true;
}

// InfiniteFloor::CheckIntersection (src/renderer/InfiniteFloor.h:20-38):
// the plane y = const, both sides.
__device__ inline bool hit_plane(f3 o, f3 d, float y, float &t_out) {
  if (fabsf(d.y) < kEps) return false;
  float t = (y - o.y) / d.y;
  if (t < kEps) return false;
  t_out = t;
  return true;
}

// ------------------------------------------------------------ traversal
struct Hit {
  float t, u, v;
  int32_t prim;
};

// Circular LDS short stack of SHORT entries per lane, spilling the oldest
// entries to a per-thread global area.  Entry = (child code, entry distance).
template <int WG, int SHORT>
struct Stack {
  uint2 *lds;
  uint2 *g;
  uint32_t sp;
  __device__ inline void push(uint2 e) {
    uint32_t slot = sp % SHORT;
    uint2 *cell = &lds[slot * WG + threadIdx.x];
    if (sp >= SHORT) g[sp - SHORT] = *cell;
    *cell = e;
    sp++;
  }
  __device__ inline uint2 pop() {
    sp--;
    uint32_t slot = sp % SHORT;
    uint2 *cell = &lds[slot * WG + threadIdx.x];
    uint2 e = *cell;
    if (sp >= SHORT) *cell = g[sp - SHORT];
    return e;
  }
};

__device__ inline void cswap(float &ka, int32_t &ca, float &kb, int32_t &cb) {
  if (ka > kb) {
    float tk = ka; ka = kb; kb = tk;
    int32_t tc = ca; ca = cb; cb = tc;
  }
}

// Closest (ANY = false) or any (ANY = true) hit in [EPS, best.t).  Restates
// QBVH::CheckIntersection (src/renderer/QBVH.cpp:30-144): slab test of four
// boxes per node (BoundingBox.h:127-172), nearest-first order, leaves tested
// as they are reached, pruning by the closest distance so far; ties on t go
// to the smaller primitive id so the result does not depend on the order.
// The traversal is resumable: init() sets up a ray, step(k) performs at most
// k node or leaf visits and reports whether the ray is finished.  traverse()
// runs one ray to the end.  FMA: slab distances as one fused operation per
// plane (see init); without it as (bound - o) * inverse, for kernels whose
// registers cannot hold the three more values per ray that the fused form
// keeps.
template <int WG, int SHORT, bool ANY, bool STATS, int NW, bool FMA = true>
struct Trav {
  const float4 *nodes, *prims;
  f3 o, d;
  float ix, iy, iz;
  float oix, oiy, oiz;  // o * inverse direction: a slab distance is fma(bound, ix, -oix)
  float slk;            // absolute slack of a box's far distance (see init)
  uint32_t nx, ny, nz;  // 1 when the ray runs towards -axis: the near slab is the upper bound
  int32_t cur;   // current entry: inner node (>= 0), leaf code or kNoEntry
  int32_t pend;  // postponed leaf or kNoEntry
  Hit best;   // u, v unnormalized until the traversal ends
  float det;  // divisor of best.u, best.v (1 for spheres)
  const float *planes;  // infinite floors, ids plane_base + k, tested after the BVH
  uint32_t n_planes, plane_base;

  __device__ inline void finish() {
    best.u = best.u / det;
    best.v = best.v / det;
    // objects outside the BVH come after it (Scene.cpp:72-80)
    for (uint32_t k = 0; k < n_planes; k++) {
      float t;
      if (!hit_plane(o, d, gld(planes + k), t)) continue;
      const int32_t id = (int32_t)(plane_base + k);
      if (t < best.t || (t == best.t && best.prim >= 0 && id < best.prim)) {
        best.t = t;
        best.u = best.v = 0.0f;
        best.prim = id;
      }
    }
  }

  __device__ inline void init(const SceneDev &S, f3 o_, f3 d_, float tmax, Stack<WG, SHORT> &stk) {
    o = o_;
    d = d_;
    nx = !(d.x >= 0.0f) ? 1u : 0u;
    ny = !(d.y >= 0.0f) ? 1u : 0u;
    nz = !(d.z >= 0.0f) ? 1u : 0u;
    if (!FMA) {
      // QBVH.cpp:56-60: zero direction component -> INF
      ix = (d.x == 0.0f) ? kInf : 1.0f / d.x;
      iy = (d.y == 0.0f) ? kInf : 1.0f / d.y;
      iz = (d.z == 0.0f) ? kInf : 1.0f / d.z;
      oix = oiy = oiz = slk = 0.0f;
    } else {
    // Slab distances are fma(bound, ix, -o.x * ix): one operation where
    // (bound - o.x) * ix takes two, and one more rounding, that of o.x * ix:
    // the distances of an origin moved by at most 2^-24 |o.x| along x.  The
    // builder pads every child box by 2^-17 of the scene's largest
    // coordinate magnitude, which covers that move eight times over while
    // every |o| <= o_far (16 times that magnitude): then a box the ray enters
    // is never culled and the hit does not change (the tie rule decides it,
    // whatever boxes are visited).  A farther origin adds the absolute slack
    // slk = 2^-21 max |o ix| to the far distances and to best.t where keys
    // are compared with it.  QBVH.cpp:56-60 takes INF for a zero direction
    // component; here a component below 2^-64 takes +-2^100 with the sign of
    // its slab choice: (bound - o.x) * 2^100 with one rounding, o.x * 2^100
    // exact, and no product overflows.
    const bool tx = fabsf(d.x) < 0x1p-64f, ty = fabsf(d.y) < 0x1p-64f, tz = fabsf(d.z) < 0x1p-64f;
    ix = tx ? (nx ? -0x1p100f : 0x1p100f) : 1.0f / d.x;
    iy = ty ? (ny ? -0x1p100f : 0x1p100f) : 1.0f / d.y;
    iz = tz ? (nz ? -0x1p100f : 0x1p100f) : 1.0f / d.z;
    oix = o.x * ix;
    oiy = o.y * iy;
    oiz = o.z * iz;
    const float om = __builtin_fmaxf(__builtin_fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z));
    slk = om <= S.o_far ? 0.0f
                        : __builtin_fmaxf(__builtin_fmaxf(tx ? 0.0f : fabsf(oix), ty ? 0.0f : fabsf(oiy)),
                                          tz ? 0.0f : fabsf(oiz)) *
                              0x1p-21f;
    }
    nodes = NW == 8 ? reinterpret_cast<const float4 *>(S.nodes8) : reinterpret_cast<const float4 *>(S.nodes);
    prims = reinterpret_cast<const float4 *>(S.prims);
    planes = S.planes;
    n_planes = S.n_planes;
    plane_base = S.n_tri + S.n_sph;
    best.t = tmax;
    best.u = best.v = 0.0f;
    best.prim = -1;
    det = 1.0f;
    cur = 0;
    pend = -1;
    stk.sp = 0;
  }

  __device__ inline bool step(Stack<WG, SHORT> &stk, int budget, uint32_t &nvis, uint32_t &ntest);
};

// Lanes of a wave move in one of two phases per trip, chosen for the whole
// wave: a node phase, where every lane whose current entry is an inner node
// visits it, or a leaf phase, where every lane holding a postponed leaf tests
// its primitives.  A lane that reaches a leaf postpones it (one slot) and
// keeps descending, so node trips run with most lanes busy and leaf trips
// gather many leaves at once; the wave switches to a leaf trip when no lane
// can visit a node or at least kLeafBatch/64 of its working lanes hold a
// leaf.  The
// primitives a ray tests and its result do not depend on the trip order (the
// tie rule "smaller t, then smaller id" and conservative boxes).
constexpr uint32_t kLeafBatch = 32;
constexpr int32_t kNoEntry = -1;

template <int WG, int SHORT, bool ANY, bool STATS, int NW, bool FMA>
__device__ inline bool Trav<WG, SHORT, ANY, STATS, NW, FMA>::step(Stack<WG, SHORT> &stk, int budget, uint32_t &nvis,
                                                          uint32_t &ntest) {
  const float ox = o.x, oy = o.y, oz = o.z;
  // next stack entry that can still hold a closer hit, or kNoEntry (an
  // entry's key is its near distance; in the fused form best.t gets the
  // slack of a far origin)
  auto pop_next = [&]() -> int32_t {
    const float lim = FMA ? best.t + slk : best.t;
    while (stk.sp > 0) {
      uint2 e = stk.pop();
      if (__uint_as_float(e.y) > lim) continue;
      return (int32_t)e.x;
    }
    return kNoEntry;
  };
  // a leaf reached as the current entry goes to the free postponed slot
  auto postpone = [&]() {
    if (pend == kNoEntry && child_is_leaf(cur)) {
      pend = cur;
      cur = pop_next();
    }
  };
  for (int it = 0; it < budget; it++) {
    const bool can_visit = cur >= 0;
    const unsigned long long vm = __ballot(can_visit);
    const uint32_t leaves = (uint32_t)__popcll(__ballot(pend != kNoEntry));
    const uint32_t lanes = (uint32_t)__popcll(__ballot(true));
    if (vm != 0ull && leaves * 64u < kLeafBatch * lanes) {
      if (can_visit) {
        // node = lo.x, hi.x, lo.y, hi.y, lo.z, hi.z (NW children each), child
        // codes: the near and far slab of each axis are picked by address;
        // an 8-wide node holds each bound as two float4 (children 0-3, 4-7)
        constexpr uint32_t Q = NW / 4;  // float4 per bound
        const float4 *np = nodes + (size_t)cur * (8u * Q);
        if (STATS) nvis++;
        float key[NW];
        int32_t code[NW];
#pragma unroll
        for (uint32_t h = 0; h < Q; h++) {
          float4 nX = ld4(np + Q * nx + h), fX = ld4(np + Q * (nx ^ 1u) + h);
          float4 nY = ld4(np + Q * (2u + ny) + h), fY = ld4(np + Q * (2u + (ny ^ 1u)) + h);
          float4 nZ = ld4(np + Q * (4u + nz) + h), fZ = ld4(np + Q * (4u + (nz ^ 1u)) + h);
          float4 chf = ld4(np + Q * 6u + h);
          float near_x[4] = {nX.x, nX.y, nX.z, nX.w}, far_x[4] = {fX.x, fX.y, fX.z, fX.w};
          float near_y[4] = {nY.x, nY.y, nY.z, nY.w}, far_y[4] = {fY.x, fY.y, fY.z, fY.w};
          float near_z[4] = {nZ.x, nZ.y, nZ.z, nZ.w}, far_z[4] = {fZ.x, fZ.y, fZ.z, fZ.w};
          int32_t c4[4] = {__float_as_int(chf.x), __float_as_int(chf.y), __float_as_int(chf.z),
                           __float_as_int(chf.w)};
          // tmin = n > tmin ? n : tmin (and tmax likewise) as one v_max / v_min:
          // equal for every input here (a NaN operand leaves the other one in
          // both forms; n, f are never signalling NaNs; +0 and -0 compare equal
          // in every later test)
#pragma unroll
          for (int k = 0; k < 4; k++) {
            float tmin = 0.0f, tmax = best.t;
            auto dist = [&](float b, float i, float o_, float oi) {
              return FMA ? __builtin_fmaf(b, i, -oi) : (b - o_) * i;
            };
            float n = dist(near_x[k], ix, ox, oix);
            float f = dist(far_x[k], ix, ox, oix);
            tmin = __builtin_fmaxf(n, tmin);
            tmax = __builtin_fminf(f, tmax);
            n = dist(near_y[k], iy, oy, oiy);
            f = dist(far_y[k], iy, oy, oiy);
            tmin = __builtin_fmaxf(n, tmin);
            tmax = __builtin_fminf(f, tmax);
            n = dist(near_z[k], iz, oz, oiz);
            f = dist(far_z[k], iz, oz, oiz);
            tmin = __builtin_fmaxf(n, tmin);
            tmax = __builtin_fminf(f, tmax);
            const float lim = FMA ? __builtin_fmaf(tmax, kSlabInflate, slk) : tmax * kSlabInflate;
            bool hit = (c4[k] != -1) && (tmin <= lim);
            key[4 * h + k] = hit ? tmin : __builtin_huge_valf();
            code[4 * h + k] = c4[k];
          }
          // sorting network, nearest first: each group of four, then (8-wide)
          // Batcher's odd-even merge of the two sorted groups
          float *kk = key + 4 * h;
          int32_t *cc = code + 4 * h;
          cswap(kk[0], cc[0], kk[1], cc[1]);
          cswap(kk[2], cc[2], kk[3], cc[3]);
          cswap(kk[0], cc[0], kk[2], cc[2]);
          cswap(kk[1], cc[1], kk[3], cc[3]);
          cswap(kk[1], cc[1], kk[2], cc[2]);
          // 8-wide: the second group's bounds are loaded only after the
          // first group is sorted, so one group's 28 words are live at a time
          if (NW == 8 && h == 0) __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (NW == 8) {
          cswap(key[0], code[0], key[4], code[4]);
          cswap(key[1], code[1], key[5], code[5]);
          cswap(key[2], code[2], key[6], code[6]);
          cswap(key[3], code[3], key[7], code[7]);
          cswap(key[2], code[2], key[4], code[4]);
          cswap(key[3], code[3], key[5], code[5]);
          cswap(key[1], code[1], key[2], code[2]);
          cswap(key[3], code[3], key[4], code[4]);
          cswap(key[5], code[5], key[6], code[6]);
        }
        const float kmiss = __builtin_huge_valf();
#pragma unroll
        for (int k = NW - 1; k >= 1; k--)
          if (key[k] != kmiss) stk.push(make_uint2((uint32_t)code[k], __float_as_uint(key[k])));
        cur = (key[0] != kmiss && !(key[0] > (FMA ? best.t + slk : best.t))) ? code[0] : pop_next();
        postpone();
      }
    } else if (pend != kNoEntry) {
      // primitive records are 3 x 16 bytes; the next record is loaded while
      // the current one is tested
      uint32_t first = leaf_first(pend), cnt = leaf_count(pend);
      const float4 *pr = prims + (size_t)first * 3;
      float4 a = ld4(pr), b = ld4(pr + 1), c = ld4(pr + 2);
      for (uint32_t k = 0; k < cnt; k++) {
        float4 na = a, nb = b, nc = c;
        if (k + 1u < cnt) {
          const float4 *pn = pr + (size_t)(k + 1u) * 3;
          na = ld4(pn);
          nb = ld4(pn + 1);
          nc = ld4(pn + 2);
        }
        uint32_t id = __float_as_uint(a.w);
        uint32_t kind = __float_as_uint(b.w);
        float t, u = 0.0f, v = 0.0f, dt = 1.0f;
        bool ok;
        if (STATS) ntest++;
        if (kind == 0u) {
          ok = hit_triangle(o, d, xyz(a), xyz(b), xyz(c), t, u, v, dt);
        } else {
          ok = hit_sphere(o, d, xyz(a), b.x, t);
        }
        if (ok && (t < best.t || (t == best.t && best.prim >= 0 && (int32_t)id < best.prim))) {
          best.t = t;
          best.u = u;
          best.v = v;
          det = dt;
          best.prim = (int32_t)id;
        }
        a = na;
        b = nb;
        c = nc;
      }
      if (ANY && best.prim >= 0) {
        finish();
        return true;
      }
      pend = kNoEntry;
      postpone();
    }
    if (cur == kNoEntry && pend == kNoEntry) {
      finish();
      return true;
    }
  }
  return false;
}

template <int WG, int SHORT, bool ANY, bool STATS, int NW, bool FMA = true>
__device__ inline void traverse(const SceneDev &S, f3 o, f3 d, Hit &best, Stack<WG, SHORT> &stk, uint32_t &nvis,
                                uint32_t &ntest) {
  Trav<WG, SHORT, ANY, STATS, NW, FMA> tr;
  tr.init(S, o, d, best.t, stk);
  tr.step(stk, 0x7FFFFFFF, nvis, ntest);
  best = tr.best;
}

// Dynamic fetch: every lane keeps its own ray; after each slice of `budget`
// visits, lanes whose ray finished take new rays (one atomic per wave) as soon
// as at least `fetch_min` lanes are idle, so a wave does not idle behind its
// slowest ray.  Work comes from kShards sub-queues: a wave starts on its own
// shard's sub-queue (heads[k] counts what was taken) and, when that one runs
// dry, reads every head at once (one lane per shard, plain loads: heads only
// grow, so a shard seen used up is used up) and moves to the next shard that
// still has entries -- no atomics on drained sub-queues at the end of a
// launch.  Must be called by whole waves.  count(k): entries of sub-queue k;
// load(k, off, S, o, d, tmax, tag) reads one entry; store(tag, hit) writes a
// finished ray's result.
static_assert(kShards <= 64, "the sub-queue scan gives one lane of a wave to each shard");
template <int WG, int SHORT, bool ANY, bool STATS, int NW, class Count, class Load, class Store>
__device__ inline uint32_t dynamic_fetch(Sharded<uint32_t> &heads, Count count, Stack<WG, SHORT> &stk, int budget,
                                         uint32_t fetch_min, Load load, Store store, uint32_t &nvis,
                                         uint32_t &ntest) {
  Trav<WG, SHORT, ANY, STATS, NW> tr;
  bool has = false, exhausted = false;
  uint32_t tag = 0, done = 0, shard = wave_shard();
  const uint32_t active = (uint32_t)__popcll(__ballot(true));
  for (;;) {
    unsigned long long m = __ballot(!has);
    uint32_t idle = (uint32_t)__popcll(m);
    if (!exhausted && idle != 0u && (idle >= fetch_min || idle == active)) {
      uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&heads[shard], idle);
      base = __shfl(base, (int)leader);
      uint32_t n = count(shard);
      uint32_t avail = base < n ? n - base : 0u;
      uint32_t r = rank_in(m);
      if (!has && r < avail) {
        const SceneDev *S;
        f3 o, d;
        float tmax;
        load(shard, base + r, S, o, d, tmax, tag);
        tr.init(*S, o, d, tmax, stk);
        has = true;
      }
      if (base + idle >= n) {  // this sub-queue is used up: find one with entries left
        const uint32_t lane = lane_id();
        bool left = false;
        if (lane < kShards) left = __hip_atomic_load(&heads[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < count(lane);
        unsigned long long lm = __ballot(left);
        if (lm == 0ull) {
          exhausted = true;
        } else {
          unsigned long long after = lm & ~((2ull << shard) - 1ull);  // shards past the current one first
          shard = (uint32_t)(__ffsll((long long)(after ? after : lm)) - 1);
        }
      }
    }
    unsigned long long busy = __ballot(has);
    if (busy == 0ull && exhausted) break;
    if (busy == 0ull) continue;
    if (has && tr.step(stk, budget, nvis, ntest)) {
      store(tag, tr.best);
      has = false;
      done++;
    }
  }
  return done;
}

// ------------------------------------------------------------ environment
// IBL::Sample(dir) (src/renderer/IBL.h:30-45) + PhiThetaToUV (:67-71).
__device__ inline void env_texel(f3 d, uint32_t w, uint32_t h, uint32_t &ix, uint32_t &iy) {
  float theta = acos_f(d.y);
  float rxz = sqrtf(d.x * d.x + d.z * d.z);
  float phi = (rxz > 0.0f) ? acos_f(d.x / rxz) : 0.0f;
  if (d.z < 0.0f) phi = kTwoPi - phi;
  float u = phi / kTwoPi;
  float v = theta / kPi;
  ix = (uint32_t)((int)(u * (float)w) % (int)w);
  iy = (uint32_t)((int)(v * (float)h) % (int)h);
}
// Texel index of direction d and the importance-sampling pdf there:
// pdf = q (w h) / (2 pi^2 total) / sin(theta) for the texel's quantized
// weight q, sin(theta) = |d.xz|: the product q times the scene's pdf factor in
// binary64, rounded to binary32, then one binary32 division (0 without
// weights); env_at(i) is its radiance.
__device__ inline size_t env_index(const SceneDev &S, f3 d) {
  uint32_t ix, iy;
  env_texel(d, S.env_w, S.env_h, ix, iy);
  return (size_t)iy * S.env_w + ix;
}
// Both read the texel's 16-byte record, so a pdf and a radiance lookup of one
// texel are one load.
__device__ inline float env_pdf_at(const SceneDev &S, size_t i, f3 d) {
  const double kf = gld(S.env_mscale + 1);
  const uint32_t q = __float_as_uint(ld4(reinterpret_cast<const float4 *>(S.env4) + i).w);
  float st = sqrtf(d.x * d.x + d.z * d.z);
  if (kf == 0.0 || q == 0u || !(st > 0.0f)) return 0.0f;
  return (float)((double)q * kf) / st;
}
__device__ inline f3 env_at(const SceneDev &S, size_t i) {
  return xyz(ld4(reinterpret_cast<const float4 *>(S.env4) + i));
}
// ------------------------------------------------------------ textures
// Surface uv of a hit: interpolated vertex uv on triangles (Polygon.h:86-96,
// duv1 * u + duv2 * v + uv0), the latlong uv of the normal on spheres
// (Sphere.h:52-75: u = (atan2(n.z, n.x) + pi) / 2pi and 1 - (asin(n.y) +
// pi/2) / pi = acos(n.y) / pi), 0 on planes (InfiniteFloor sets none).
__device__ inline void hit_uv(const SceneDev &S, uint32_t pid, float bu, float bv, f3 nrm, float &u, float &v) {
  u = 0.0f;
  v = 0.0f;
  if (pid < S.n_tri) {
    if (S.tri_uv) {
      const float4 *q = reinterpret_cast<const float4 *>(S.tri_uv) + (size_t)pid * 2;
      float4 a = ld4(q), b = ld4(q + 1);
      u = (a.z * bu + b.x * bv) + a.x;
      v = (a.w * bu + b.y * bv) + a.y;
    }
  } else if (pid < S.n_tri + S.n_sph) {
    float rxz = sqrtf(nrm.x * nrm.x + nrm.z * nrm.z);
    float phi = (rxz > 0.0f) ? acos_f(nrm.x / rxz) : 0.0f;
    u = (nrm.z >= 0.0f ? kPi + phi : kPi - phi) / kTwoPi;
    v = acos_f(nrm.y) / kPi;
  }
}
// Image::GetPixelByUV (src/tools/Image.h:21-30): u, v clamped to [0, 1]
// (NaN to 0), scaled by 0.99999, texel (int(u W), int(v H)), row 0 on top.
__device__ inline f3 texel(const SceneDev &S, uint32_t k, float u, float v) {
  const TexRec r = gld(S.tex_rec + k);
  u = u > 0.0f ? (u < 1.0f ? u : 1.0f) : 0.0f;
  v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
  uint32_t x = (uint32_t)(int)((u * 0.99999f) * (float)r.width);
  uint32_t y = (uint32_t)(int)((v * 0.99999f) * (float)r.height);
  x = x < r.width ? x : r.width - 1u;
  y = y < r.height ? y : r.height - 1u;
  float4 p = ld4(reinterpret_cast<const float4 *>(S.tex) + ((size_t)r.first + (size_t)y * r.width + x));
  return mk(p.x, p.y, p.z);
}

// Bucket of key k in a directory of n buckets: floor(k * scale) in binary64,
// scale = n / total; the same expression builds and reads the directories.
__device__ inline uint32_t env_bucket(unsigned long long k, double scale, uint32_t n) {
  uint32_t b = (uint32_t)((double)k * scale);
  return b < n ? b : n - 1u;
}
// search_le(a, n, key) restricted to the entries the key's bucket can select:
// the same index, found in a few steps.
__device__ inline uint32_t search_guided(const unsigned long long *a, uint32_t n, unsigned long long key,
                                         const uint2_t *guide, double scale) {
  const uint2_t r = gld(guide + env_bucket(key, scale, n));
  uint32_t lo = r.lo, hi = r.hi + 1u;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (gld(a + mid) <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}
// A direction from the tables with four uniforms in [0,1]: a texel by the
// marginal and row CDFs (its index into `texel`), then a point of it uniform
// in (phi, theta).
__device__ inline bool env_sample4(const SceneDev &S, float u1, float u2, float u3, float u4, f3 &dir,
                                   size_t &texel) {
  uint32_t W = S.env_w, H = S.env_h;
  unsigned long long tot = gld(S.env_marg + H);
  if (tot == 0ull) return false;
  unsigned long long t = (unsigned long long)((double)u1 * (double)tot);
  if (t >= tot) t = tot - 1;
  uint32_t iy = search_guided(S.env_marg, H, t, S.env_mguide, gld(S.env_mscale));
  const unsigned long long *row = S.env_row + (size_t)iy * (W + 1);
  unsigned long long rt = gld(row + W);
  unsigned long long t2 = (unsigned long long)((double)u2 * (double)rt);
  if (t2 >= rt) t2 = rt - 1;
  uint32_t ix = search_guided(row, W, t2, S.env_rguide + (size_t)iy * W, gld(S.env_rscale + iy));
  float fx = ((float)ix + u3) / (float)W;
  float fy = ((float)iy + u4) / (2.0f * (float)H);
  if (fx >= 1.0f) fx = 0.99999994f;
  if (fy >= 0.5f) fy = 0.49999997f;
  float sp, cp, sth, cth;
  sincos_turns(fx, sp, cp);
  sincos_turns(fy, sth, cth);
  dir = mk(sth * cp, cth, sth * sp);
  texel = (size_t)iy * W + ix;
  return true;
}

// ------------------------------------------------------------ accumulation
__device__ inline unsigned long long fix(float v) {
  if (!(v > 0.0f)) return 0ull;
  if (v > kFixClamp) v = kFixClamp;
  return (unsigned long long)(long long)rintf(v * kFixScale);
}
// Adds one contribution of sample sidx (pixel `pixel`) to the job's sums.
// Device-scope atomics on the pixel sums cost about as much as the rest of
// a shading step on an image-based-lighting render, so jobs that have
// per-sample records take a plain read-modify-write there instead: no two
// lanes ever hold the same sample (its paths and pending branches run one at
// a time in one slot, and a shadow record exists only for the step that made
// it).
// Without records the pixel of the sample is computed here, where it is
// needed (pixel = sample / (ss^2 spp): samples are stored pixel-major).
__device__ inline void deposit(const JobDev &J, uint32_t sidx, unsigned long long a, unsigned long long b,
                               unsigned long long c) {
  if ((a | b | c) == 0ull) return;
  if (J.recs) {
    __attribute__((address_space(1))) unsigned long long *r =
        (__attribute__((address_space(1))) unsigned long long *)J.recs + 3ull * sidx;
    const unsigned long long ra = r[0], rb = r[1], rc = r[2];
    r[0] = ra + a;
    r[1] = rb + b;
    r[2] = rc + c;
    return;
  }
  const uint32_t pixel = sidx / (J.ss * J.ss * J.nspp);
  unsigned long long *acc = J.accum + 3ull * pixel;
  if (a) gadd(acc + 0, a);
  if (b) gadd(acc + 1, b);
  if (c) gadd(acc + 2, c);
}
__device__ inline void add_px(const JobDev &J, uint32_t sidx, f3 beta, f3 L, float s) {
  deposit(J, sidx, fix(beta.x * L.x * s), fix(beta.y * L.y * s), fix(beta.z * L.z * s));
}

// ------------------------------------------------------------ job lookup
__device__ inline uint32_t find_job(const Ctrl *ctrl, const JobDev *jobs, unsigned long long g) {
  uint32_t lo = ctrl->job_lo, hi = ctrl->job_hi;  // answer in [lo, hi)
  while (hi - lo > 1) {
    uint32_t mid = lo + ((hi - lo) >> 1);
    if (jobs[mid % kJobRing].sample_base <= g) lo = mid;
    else hi = mid;
  }
  return lo % kJobRing;
}

// Camera::SampleRayForPixel (src/renderer/Camera.h:65-93) + CircleAperture
// (src/renderer/Aperture.h:23-35); sub-pixel centres (PathTracer.cpp:158-161).
struct Decomp {
  uint32_t pixel, sub, sample_no;
};
__device__ inline Decomp decompose(const JobDev &J, uint32_t s) {
  uint32_t spp_px = J.ss * J.ss * J.nspp;
  Decomp r;
  r.pixel = s / spp_px;
  uint32_t rem = s - r.pixel * spp_px;
  r.sub = rem / J.nspp;
  r.sample_no = J.spp_begin + 1u + (rem - r.sub * J.nspp);
  return r;
}
__device__ inline void camera_ray(const JobDev &J, const Decomp &dc, f3 &o, f3 &d) {
  uint32_t W = J.width, H = J.height, ss = J.ss;
  uint32_t r = dc.pixel / W, x = dc.pixel - r * W;
  uint32_t y = H - 1u - r;
  uint32_t sxi = dc.sub % ss, syi = dc.sub / ss;
  float rx = (2.0f * (float)sxi + 1.0f) / (2.0f * (float)ss);
  float ry = (2.0f * (float)syi + 1.0f) / (2.0f * (float)ss);
  float xs = (float)x + rx, ys = (float)y + ry;
  f3 pos = mk(J.cam_pos[0], J.cam_pos[1], J.cam_pos[2]);
  f3 C = mk(J.cam_c[0], J.cam_c[1], J.cam_c[2]);
  f3 X = mk(J.cam_x[0], J.cam_x[1], J.cam_x[2]);
  f3 Y = mk(J.cam_y[0], J.cam_y[1], J.cam_y[2]);
  float a = xs / (float)W - 0.5f;
  float b = ys / (float)H - 0.5f;
  f3 target = add(add(C, scl(X, a)), scl(Y, b));
  f3 dir = normalize(sub(target, pos));
  f3 org = pos;
  if (J.aperture > 0.0f) {
    uint4 lr = rng_init(path_seed(J.seed, dc.pixel, dc.sub, J.spp_begin, 0xFFFFFFFFu));
    float l0 = rng_float(lr), l1 = rng_float(lr);
    float st, ct;
    sincos_turns(l0, st, ct);
    float rr = J.aperture * sqrtf(l1);
    float dx = rr * ct, dy = rr * st;
    float ft = J.focal / fabsf(dir.z);
    f3 focus = add(pos, scl(dir, ft));
    org = add(pos, mk(dx, dy, 0.0f));
    dir = normalize(sub(focus, org));
  }
  o = org;
  d = dir;
}

// Wave-level grab of up to `want` lanes' worth of samples: one fetch-and-add
// per wave on a cursor shard k, whose j-th sample is shard_sample(j, k).  A wave
// starts on its own shard; when that one is used up it reads every cursor at
// once (one lane per shard) and moves to the next shard with samples left, so
// no shard's samples wait for waves that have no work.  The admitted total is
// read once per launch (it changes only between launches); once every shard
// is used up the wave stops asking.  A cursor may move past its limit; the
// host resets it before it admits more work (no kernel runs then).  Returns
// this lane's global sample index, or ~0ull when none is left for it (lanes
// past a shard's last sample go without until the next call).
struct SampleCursor {
  unsigned long long total;
  uint32_t shard;
  bool open;
};
__device__ inline SampleCursor sample_cursor(const Ctrl *ctrl) {
  SampleCursor c;
  c.total = __atomic_load_n(&ctrl->total_samples, __ATOMIC_RELAXED);
  c.shard = wave_shard();
  c.open = true;
  return c;
}
// Moves sc to the next shard (after the current one, cyclically) that still
// has samples; false, and sc closed, when none has.  Whole wave.
__device__ inline bool next_sample_shard(const Ctrl *ctrl, SampleCursor &sc) {
  const uint32_t lane = lane_id();
  bool left = false;
  if (lane < kShards)
    left = __hip_atomic_load(&ctrl->next_sample[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           shard_limit(sc.total, lane);
  unsigned long long lm = __ballot(left);
  if (lm == 0ull) {
    sc.open = false;
    return false;
  }
  unsigned long long after = lm & ~((2ull << sc.shard) - 1ull);
  sc.shard = (uint32_t)(__ffsll((long long)(after ? after : lm)) - 1);
  return true;
}
__device__ inline unsigned long long grab_sample(Ctrl *ctrl, SampleCursor &sc, bool want) {
  unsigned long long m = __ballot(want);
  if (m == 0ull || !sc.open) return ~0ull;
  const uint32_t cnt = (uint32_t)__popcll(m);
  const uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
  const uint32_t r = rank_in(m);
  for (;;) {
    const uint32_t k = sc.shard;
    unsigned long long base = 0ull, take = 0ull;
    if (lane_id() == leader) {
      unsigned long long lim = shard_limit(sc.total, k);
      unsigned long long old = atomicAdd(&ctrl->next_sample[k], (unsigned long long)cnt);
      unsigned long long avail = lim > old ? lim - old : 0ull;
      take = avail < cnt ? avail : (unsigned long long)cnt;
      base = old;
    }
    base = __shfl(base, (int)leader);
    take = __shfl(take, (int)leader);
    if (take > 0ull) {
      if (take < cnt) next_sample_shard(ctrl, sc);  // shard k is used up now
      if (!want || (unsigned long long)r >= take) return ~0ull;
      return shard_sample(base + r, k);
    }
    if (!next_sample_shard(ctrl, sc)) return ~0ull;
  }
}

// Per-wave completion counting: a wave keeps a running count for one job and
// adds it to that job's counter in the wave's shard when the job changes and
// when the kernel ends, so a launch costs about one atomic per wave.  Called
// with every lane of the wave active.
struct DoneAcc {
  uint32_t job = ~0u;
  uint32_t n = 0u;
};
__device__ inline void done_flush(unsigned long long *done, DoneAcc &acc) {
  if (acc.n != 0u && lane_id() == 0u) atomicAdd(&done[done_index(wave_shard(), acc.job)], (unsigned long long)acc.n);
  acc.n = 0u;
}
__device__ inline void count_done(unsigned long long *done, DoneAcc &acc, bool fin, uint32_t job) {
  bool need = fin;
  for (;;) {
    unsigned long long m = __ballot(need);
    if (m == 0ull) break;
    uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
    uint32_t lj = __shfl(job, (int)leader);
    bool mine = need && job == lj;
    uint32_t c = (uint32_t)__popcll(__ballot(mine));
    if (lj != acc.job) {
      done_flush(done, acc);
      acc.job = lj;
    }
    acc.n += c;
    if (mine) need = false;
  }
}

// One atomic per wave into the wave's shard of a counter.
__device__ inline void wave_count(Sharded<unsigned long long> &shards, unsigned long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if (lane_id() == 0 && v) atomicAdd(&shards[wave_shard()], v);
}

// A ray-queue entry: the path's slot (pool <= 2^26 slots) and, in the top
// six bits, its scene's slot in the scene table, so a traversal can read the
// scene record without first loading the path's job.
constexpr uint32_t kEntrySlotBits = 26;
static_assert(kMaxSceneSlots <= (1u << (32 - kEntrySlotBits)), "scene slot must fit the entry's top bits");
__device__ inline uint32_t q_entry(uint32_t slot, uint32_t scene) { return slot | (scene << kEntrySlotBits); }
__device__ inline uint32_t q_slot(uint32_t e) { return e & ((1u << kEntrySlotBits) - 1u); }
__device__ inline uint32_t q_scene(uint32_t e) { return e >> kEntrySlotBits; }

// Writes a fresh camera path for global sample g into slot; returns the
// queue entry for it.
__device__ inline uint32_t start_sample(const KernelArgs &A, uint32_t slot, unsigned long long g) {
  uint32_t js = find_job(A.ctrl, A.jobs, g);
  const JobDev &J = A.jobs[js];
  uint32_t s = (uint32_t)(g - J.sample_base);
  Decomp dc = decompose(J, s);
  f3 o, d;
  camera_ray(J, dc, o, d);
  uint4 rng = rng_init(path_seed(J.seed, dc.pixel, dc.sub, dc.sample_no, 0u));
  reinterpret_cast<float4 *>(A.p.ray_o)[slot] = make_float4(o.x, o.y, o.z, 1.0f);
  reinterpret_cast<float4 *>(A.p.ray_d)[slot] = make_float4(d.x, d.y, d.z, 1.0f);
  reinterpret_cast<uint4 *>(A.p.misc)[slot] = make_uint4(__float_as_uint(1.0f), ps_pack(0, 0, 0, 0), s, js);
  reinterpret_cast<uint4 *>(A.p.rng)[slot] = rng;
  return q_entry(slot, J.scene);
}

// ------------------------------------------------------------ kernels
// refill: every inactive slot takes a new sample while samples remain.  A
// wave covers kRefillSlots consecutive slots per trip, as groups of 64 (lane
// l holds slot l of each group: coalesced accesses), and takes samples for
// all its idle ones with one cursor atomic and one queue append (one atomic
// per 64 slots serialized on the cursor lines and held a 2^26-slot refill to
// 1.5 ms).  Idle slots are numbered in slot order, so consecutive slots take
// consecutive samples and enter the queue in slot order, as with one group.
constexpr uint32_t kRefillPerLane = kRefillSlots / 64;
template <int WG>
__global__ __launch_bounds__(WG) void k_refill(KernelArgs A, int parity) {
  const uint32_t P = A.p.pool;
  const uint32_t waves = gridDim.x * (WG / 64);
  const uint32_t wave = blockIdx.x * (WG / 64) + (threadIdx.x >> 6);
  const uint32_t lane = lane_id();
  SampleCursor sc = sample_cursor(A.ctrl);
  for (uint32_t base = wave * kRefillSlots; base < P; base += waves * kRefillSlots) {
    bool idle[kRefillPerLane];
    uint32_t req[kRefillPerLane];  // the slot's position among the wave's idle slots
    uint32_t want = 0;
#pragma unroll
    for (uint32_t u = 0; u < kRefillPerLane; u++) {
      const uint32_t slot = base + u * 64u + lane;
      idle[u] = slot < P && ((reinterpret_cast<const uint4 *>(A.p.misc)[slot].y >> 19) & 1u) == 0u;
      const unsigned long long m = __ballot(idle[u]);
      req[u] = want + rank_in(m);
      want += (uint32_t)__popcll(m);
    }
    if (want == 0u) continue;
    // positions [done, done + take) come from cursor shard k starting at `old`
    unsigned long long got[kRefillPerLane];
#pragma unroll
    for (uint32_t u = 0; u < kRefillPerLane; u++) got[u] = ~0ull;
    uint32_t done = 0;
    while (done < want && sc.open) {
      const uint32_t k = sc.shard, ask = want - done;
      unsigned long long old = 0ull, take = 0ull;
      if (lane == 0u) {
        const unsigned long long lim = shard_limit(sc.total, k);
        old = atomicAdd(&A.ctrl->next_sample[k], (unsigned long long)ask);
        const unsigned long long avail = lim > old ? lim - old : 0ull;
        take = avail < ask ? avail : (unsigned long long)ask;
      }
      old = __shfl(old, 0);
      take = __shfl(take, 0);
#pragma unroll
      for (uint32_t u = 0; u < kRefillPerLane; u++)
        if (idle[u] && req[u] >= done && req[u] < done + (uint32_t)take) got[u] = shard_sample(old + (req[u] - done), k);
      done += (uint32_t)take;
      if (take < ask && !next_sample_shard(A.ctrl, sc)) break;  // shard k is used up
    }
    // the started slots enter sub-queue k in slot order
    uint32_t qpos[kRefillPerLane], n_all = 0;
#pragma unroll
    for (uint32_t u = 0; u < kRefillPerLane; u++) {
      const unsigned long long m = __ballot(got[u] != ~0ull);
      qpos[u] = n_all + rank_in(m);
      n_all += (uint32_t)__popcll(m);
    }
    if (n_all == 0u) continue;
    const uint32_t k = wave_shard();
    uint32_t qbase = 0;
    if (lane == 0u) qbase = atomicAdd(&A.ctrl->qcount[parity][k], n_all);
    qbase = __shfl(qbase, 0);
#pragma unroll
    for (uint32_t u = 0; u < kRefillPerLane; u++) {
      if (got[u] == ~0ull) continue;
      const uint32_t slot = base + u * 64u + lane;
      A.p.queue[parity][k * A.p.qcap + qbase + qpos[u]] = start_sample(A, slot, got[u]);
    }
  }
}

// A round of samples [base, base + n0) placed by the host: queue 0 holds
// the n0 entries k_fill_fresh writes (sub-queue k: shard_limit(n0, k)), its
// fetch heads start at zero, and solo_job names the round's only job when
// it has one.  One block of kShards threads.
__global__ void k_round_setup(Ctrl *ctrl, uint32_t n0, uint32_t solo_job) {
  const uint32_t k = threadIdx.x;
  if (k < kShards) {
    ctrl->qcount[0][k] = (uint32_t)shard_limit(n0, k);
    ctrl->qhead[0][k] = 0u;
  }
  if (k == 0) ctrl->solo_job = solo_job;
}

// First fill of a cleared pool: slot i < n0 takes sample base + i.  Sample i sits
// at cursor position (r / kShards) * 64 + i % 64 of shard r % kShards, r =
// i / 64, so run r goes to that sub-queue at the same entries, and the host
// sets the cursors and counts to match: the same state the cursor grabs
// would leave, without an atomic.
__global__ __launch_bounds__(256) void k_fill_fresh(KernelArgs A, unsigned long long base, uint32_t n0) {
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n0; i += stride) {
    const uint32_t e = start_sample(A, i, base + i);
    const uint32_t r = i / 64u;
    A.p.queue[0][(r % kShards) * A.p.qcap + (r / kShards) * 64u + i % 64u] = e;
  }
}

// Five waves per SIMD with 4-wide nodes and 64-lane groups (96 VGPRs; the
// fused slab test's three more values per ray would otherwise take the
// kernel to 98-102 and four waves, which measured 5 % slower on c2).
template <int WG, int SHORT, int NW, bool STATS, bool DF, bool SOLO>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(NW == 4 && WG == 64 ? 5 : 4))) void k_extend(
    KernelArgs A, int parity) {
  __shared__ uint2 lds[SHORT * WG];
  __shared__ uint32_t s_pre[kShards + 1];
  const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
  // this step's shade appends to the other queue; clear its counts and the
  // next extend's fetch heads (nobody else touches them during this kernel)
  if (blockIdx.x == 0 && threadIdx.x < kShards) {
    A.ctrl->qcount[parity ^ 1][threadIdx.x] = 0u;
    A.ctrl->qhead[parity ^ 1][threadIdx.x] = 0u;
  }
  load_prefix(s_pre, A.ctrl->qcount[parity]);
  const uint32_t n = s_pre[kShards];
  const uint32_t stride = gridDim.x * WG;
  const uint32_t *queue = A.p.queue[parity];
  const uint32_t qcap = A.p.qcap;
  Stack<WG, SHORT> stk;
  stk.lds = lds;
  stk.g = A.p.gstack + (size_t)gtid * A.p.gcap;
  uint32_t segs = 0, nv = 0, nt = 0;
  SceneDev Ss{};  // SOLO: the only job's scene, in scalar registers
  if (SOLO) Ss = cload(A.scenes + cload(A.jobs + cload(&A.ctrl->solo_job)).scene);
  if (DF) {
    auto count = [&](uint32_t k) { return s_pre[k + 1] - s_pre[k]; };
    auto load = [&](uint32_t k, uint32_t off, const SceneDev *&S, f3 &o, f3 &d, float &tmax, uint32_t &tag) {
      const uint32_t e = queue[k * qcap + off], slot = q_slot(e);
      float4 ro = reinterpret_cast<const float4 *>(A.p.ray_o)[slot];
      float4 rd = reinterpret_cast<const float4 *>(A.p.ray_d)[slot];
      if (SOLO) S = &Ss;
      else S = &A.scenes[q_scene(e)];
      o = xyz(ro);
      d = xyz(rd);
      tmax = kInf;
      tag = slot;
    };
    auto store = [&](uint32_t slot, const Hit &h) {
      reinterpret_cast<float4 *>(A.p.hit)[slot] = make_float4(h.t, __int_as_float(h.prim), h.u, h.v);
    };
    segs = dynamic_fetch<WG, SHORT, false, STATS, NW>(A.ctrl->qhead[parity], count, stk, (int)A.p.df_budget,
                                                 A.p.df_fetch_min, load, store, nv, nt);
  }
  for (uint32_t base = blockIdx.x * WG; !DF && base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    if (i < n) {
      const uint32_t e = queue[flat_slot(s_pre, i, qcap)], slot = q_slot(e);
      float4 ro = reinterpret_cast<const float4 *>(A.p.ray_o)[slot];
      float4 rd = reinterpret_cast<const float4 *>(A.p.ray_d)[slot];
      const SceneDev &S = SOLO ? Ss : A.scenes[q_scene(e)];
      Hit h;
      h.t = kInf;
      h.u = h.v = 0.0f;
      h.prim = -1;
      traverse<WG, SHORT, false, STATS, NW>(S, xyz(ro), xyz(rd), h, stk, nv, nt);
      reinterpret_cast<float4 *>(A.p.hit)[slot] = make_float4(h.t, __int_as_float(h.prim), h.u, h.v);
      segs++;
    }
  }
  wave_count(A.ctrl->segments, segs);
  if (STATS) {
    wave_count(A.ctrl->node_visits, nv);
    wave_count(A.ctrl->prim_tests, nt);
  }
}

// Path state between segments (kept in the slot arrays between launches).
struct PathState {
  f3 o, d, beta;
  uint32_t word;  // ps_pack(depth, branch, pending, mis)
  uint32_t sidx;  // sample index within the job
  uint32_t js;    // job ring slot
  uint4 rng;
  float pdf;      // MIS: BSDF pdf of the last Lambert bounce (when ps_mis(word))
};
// An environment light sample waiting for its visibility test.
struct ShadowRay {
  f3 o, d, c;
};

// Shades the segment of slot `slot` that ended at hr (t, primitive id bits,
// u, v): environment and emission contributions go to the job's pixel sums
// and st becomes the next segment's state.  Mirrors the CPU restatement's
// per-segment loop body.  pdf_mem: where the slot's last Lambert pdf is
// stored (null: st.pdf).  Returns true while the path, or one of its pending
// glass branches, continues; `shadow` is set when an environment sample needs
// a visibility ray.
// SPEC: the environment light sample of a Lambert vertex is drawn at the
// start of the hit branch, before the material is known, so its chain of
// dependent table loads overlaps the normal and material loads instead of
// following them (the sample is used only where the vertex turns out to be a
// non-emissive Lambert one that samples the environment; its uniforms are the
// same draws: one for the material, one for Russian roulette past rr_depth,
// two for the Lambert direction come first).  Costs registers, so only the
// shading kernel uses it.  ENV = false: the scene has no environment map, so
// the code for misses into it, its sampling and MIS is left out (it is the
// peak of register pressure: the single-job kernels of such a scene keep
// every register and the shading kernel runs five waves per SIMD instead of
// four).
template <bool SPEC = false, bool ENV = true>
__device__ inline bool shade_segment(const KernelArgs &A, uint32_t slot, PathState &st, float4 hr,
                                     const float *pdf_mem, bool &shadow, ShadowRay &sh, const JobDev &J,
                                     const SceneDev &S) {
  float4 *pend = reinterpret_cast<float4 *>(A.p.pend);
  const uint32_t P = A.p.pool;
  bool cont = false;
  f3 o = st.o, d = st.d, beta = st.beta;
  uint4 rng = st.rng;
  uint32_t word = st.word;
  const uint32_t sidx = st.sidx;
  float pdf_next = 0.0f;
  uint32_t depth = ps_depth(word), branch = ps_branch(word), npend = ps_pending(word);
  const bool env_on = ENV && S.env != nullptr;
  bool nee = env_on && (J.flags & 1u) && gld(S.env_marg + S.env_h) > 0ull;
  int32_t prim = __float_as_int(hr.y);
  uint32_t mis_next = 0;
  if (prim < 0) {
    // PathTracer.cpp:357-363: miss -> IBL lookup or black background
    if (env_on) {
      const size_t ei = env_index(S, d);
      float w = 1.0f;
      if (ps_mis(word)) {
        float pp = pdf_mem ? pdf_mem[slot] : st.pdf;
        float pe = env_pdf_at(S, ei, d);
        w = (pp * pp) / (pp * pp + pe * pe);
      }
      add_px(J, sidx, beta, env_at(S, ei), w);
    }
  } else {
    uint32_t pid = (uint32_t)prim;
    bool s_ok = false;  // SPEC: the environment sample, its texel's pdf and radiance
    f3 s_ld = mk(0.0f, 0.0f, 0.0f), s_L = s_ld;
    float s_pe = 0.0f;
    if (SPEC && nee) {
      uint4 r2 = rng;
      (void)rng_next(r2);                             // material roulette
      if (depth > J.rr_depth) (void)rng_next(r2);     // Russian roulette
      (void)rng_next(r2);                             // Lambert direction (two draws)
      (void)rng_next(r2);
      float e1 = rng_float(r2), e2 = rng_float(r2), e3 = rng_float(r2), e4 = rng_float(r2);
      size_t ei;
      s_ok = env_sample4(S, e1, e2, e3, e4, s_ld, ei);
      if (s_ok) {
        s_pe = env_pdf_at(S, ei, s_ld);
        s_L = env_at(S, ei);
      }
    }
    f3 x = add(o, scl(d, hr.x));
    f3 nrm;
    uint32_t pm;  // material word: in the normal record of a triangle
    if (pid < S.n_tri) {
      const float4 *tn = reinterpret_cast<const float4 *>(S.tri_n) + (size_t)pid * 3;
      float4 n0 = ld4(tn), n1 = ld4(tn + 1), n2 = ld4(tn + 2);
      nrm = normalize(add(add(scl(xyz(n1), hr.z), scl(xyz(n2), hr.w)), xyz(n0)));
      pm = __float_as_uint(n0.w);
    } else if (pid < S.n_tri + S.n_sph) {
      float4 sc = ld4(reinterpret_cast<const float4 *>(S.sph) + (pid - S.n_tri));
      nrm = normalize(sub(x, xyz(sc)));
      pm = gld(S.prim_mat + pid);
    } else {
      nrm = mk(0.0f, d.y < 0.0f ? 1.0f : -1.0f, 0.0f);  // InfiniteFloor.h:34-36: facing the ray
      pm = gld(S.prim_mat + pid);
    }
    f3 nl = dot(nrm, d) < 0.0f ? nrm : scl(nrm, -1.0f);
    // material roulette (PathTracer.cpp:368-386)
    uint32_t mfirst = pm & 0x00FFFFFFu, mcount = (pm >> 24) + 1u;
    float value = rng_float(rng);
    uint32_t sel = mcount - 1u;
    for (uint32_t q = 0; q < mcount; q++) {
      float r = gptr(S.mats)[mfirst + q].rate;
      if (value <= r) { sel = q; break; }
      value -= r;
    }
    const __attribute__((address_space(1))) MatRec &m = gptr(S.mats)[mfirst + sel];
    // hit colour: material colour times the texel (PathTracer.cpp:327-346,388-389)
    const f3 col0 = mk(m.color[0], m.color[1], m.color[2]);
    f3 col = col0;
    if (m.tex) {
      float tu, tv;
      hit_uv(S, pid, hr.z, hr.w, nrm, tu, tv);
      f3 tx = texel(S, m.tex - 1u, tu, tv);
      col = mk(col0.x * tx.x, col0.y * tx.y, col0.z * tx.z);
    }
    f3 em = mk(m.emission[0], m.emission[1], m.emission[2]);
    // Russian roulette (PathTracer.cpp:398-416)
    float prr = col.x > (col.y > col.z ? col.y : col.z) ? col.x : (col.y > col.z ? col.y : col.z);
    if (J.max_depth == 0u && depth > 64u) prr *= ldexpf(1.0f, -(int)(depth - 64u));
    bool killed = false;
    if (depth > J.rr_depth) {
      if (rng_float(rng) >= prr) killed = true;
    } else {
      prr = 1.0f;
    }
    // with the reference's next-event estimation (flags bit 1) emission counts
    // only where the camera ray hits (PathTracer.cpp:406-439); its direct
    // term is zero (:312-314)
    // (a surface that emits nothing adds zero: skipped)
    if ((!(J.flags & 2u) || depth == 0u) && (em.x != 0.0f || em.y != 0.0f || em.z != 0.0f))
      add_px(J, sidx, beta, em, 1.0f);
    uint32_t dmax = J.max_depth == 0u ? 255u : J.max_depth;
    bool can_continue = !killed && depth + 1u < dmax;
    if (!killed) {
      // Each material branch only chooses the continuation: the unnormalized
      // direction `raw` (normalized once after the branches, where the lanes
      // of every material run it together, instead of once per material
      // branch) or the refracted direction, and the throughput factor
      // num / den on the colour col.  Operations and their order per lane are
      // those of PathTracer.cpp:445-591.
      bool go = false, use_refr = false, lam = false;
      f3 raw = mk(0.0f, 0.0f, 0.0f);
      float num = 1.0f, den = prr;
      if (m.type == 0u) {
        // Radiance_Lambert (PathTracer.cpp:445-500), unnormalized basis
        if (dot(em, em) == 0.0f) {
          f3 w = nl;
          f3 u = fabsf(nl.x) > kEps ? cross(mk(0.0f, 1.0f, 0.0f), w) : cross(mk(1.0f, 0.0f, 0.0f), w);
          f3 vv = cross(w, u);
          float u1 = rng_float(rng), u2 = rng_float(rng);
          float s1, c1;
          sincos_turns(u1, s1, c1);
          float r2 = sqrtf(u2), r3 = sqrtf(1.0f - u2);
          raw = add(add(scl(scl(u, r3), c1), scl(scl(vv, r3), s1)), scl(w, r2));
          lam = true;
          go = can_continue;
        }
      } else {
        // Radiance_Specular (PathTracer.cpp:503-520) and the reflected ray of
        // Radiance_Refraction (:523-591) are the same mirror direction
        raw = sub(d, scl(scl(nl, 2.0f), dot(d, nl)));
        if (m.type == 1u) {
          go = can_continue;
        } else {
          bool into = dot(nrm, nl) > 0.0f;
          float nr = into ? 1.0f / m.ior : m.ior / 1.0f;
          float dd = dot(d, nl);
          float cos2t = 1.0f - nr * nr * (1.0f - dd * dd);
          if (cos2t < 0.0f) {
            // total internal reflection weighs by the untextured colour (PathTracer.cpp:543-548)
            col = col0;
            go = can_continue;
          } else {
            const f3 refr = normalize(sub(scl(d, nr), scl(scl(nrm, into ? 1.0f : -1.0f), dd * nr + sqrtf(cos2t))));
            float F0 = ((m.ior - 1.0f) * (m.ior - 1.0f)) / ((m.ior + 1.0f) * (m.ior + 1.0f));
            float c = 1.0f - (into ? -dd : -dot(refr, nl));
            float c5 = c * c * c * c * c;
            float Fr = F0 + (1.0f - F0) * c5;
            float Tr = (1.0f - Fr) * (nr * nr);
            if (depth > 2u) {
              float prob = 0.1f + 0.8f * Fr;
              if (rng_float(rng) < prob) {
                num = Fr;
                den = prr * prob;
              } else {
                num = Tr;
                den = prr * (1.0f - prob);
                raw = refr;  // already normalized
                use_refr = true;
              }
              go = can_continue;
            } else if (can_continue) {
              // depth <= 2: both branches; the refracted one waits in the
              // slot's pending stack, the reflected one continues now.
              float kt = Tr / prr;
              f3 bt = mk(beta.x * (col.x * kt), beta.y * (col.y * kt), beta.z * (col.z * kt));
              float4 *pe = pend + ((size_t)npend * P + slot) * 3;
              pe[0] = make_float4(x.x, x.y, x.z, bt.x);
              pe[1] = make_float4(refr.x, refr.y, refr.z, bt.y);
              pe[2] = make_float4(bt.z, __uint_as_float(depth + 1u), __uint_as_float(branch | (1u << depth)), 0.0f);
              npend++;
              num = Fr;
              go = true;
            }
          }
        }
      }
      const float k = num / den;
      const f3 nd = normalize(raw);
      if (lam && nee && can_continue) {
        // one environment light sample, weighed by the power heuristic
        float e1 = rng_float(rng), e2 = rng_float(rng), e3 = rng_float(rng), e4 = rng_float(rng);
        f3 ld = s_ld;
        float pe = s_pe;
        bool ok = s_ok;
        size_t ei = 0;
        if (!SPEC) {
          ok = env_sample4(S, e1, e2, e3, e4, ld, ei);
          if (ok) pe = env_pdf_at(S, ei, ld);
        }
        if (ok) {
          // the sampled texel's pdf and radiance
          float c = dot(nl, ld);
          if (c > 0.0f && pe > 0.0f) {
            float pb = c * kInvPi;
            float wl = (pe * pe) / (pe * pe + pb * pb);
            float s = ((pb / pe) * wl) * k;
            const f3 L = SPEC ? s_L : env_at(S, ei);
            f3 bc = mk(beta.x * col.x, beta.y * col.y, beta.z * col.z);
            sh.c = mk(bc.x * L.x * s, bc.y * L.y * s, bc.z * L.z * s);
            sh.o = x;
            sh.d = ld;
            shadow = true;
          }
        }
      }
      if (go) {
        beta = mk(beta.x * (col.x * k), beta.y * (col.y * k), beta.z * (col.z * k));
        if (lam && nee) {
          mis_next = 1u;
          pdf_next = dot(nl, nd) * kInvPi;
        }
        o = x;
        d = mk(use_refr ? raw.x : nd.x, use_refr ? raw.y : nd.y, use_refr ? raw.z : nd.z);
        depth++;
        cont = true;
      }
    }
  }
  if (cont) {
    word = ps_pack(depth, branch, npend, mis_next);
  } else if (npend > 0u) {
    // resume the most recent pending branch with its own RNG stream
    npend--;
    const float4 *pe = pend + ((size_t)npend * P + slot) * 3;
    float4 a = pe[0], b = pe[1], c = pe[2];
    o = xyz(a);
    d = xyz(b);
    beta = mk(a.w, b.w, c.x);
    uint32_t nd = __float_as_uint(c.y), nb = __float_as_uint(c.z);
    Decomp dc = decompose(J, sidx);
    rng = rng_init(path_seed(J.seed, dc.pixel, dc.sub, dc.sample_no, nb));
    word = ps_pack(nd, nb, npend, 0u);
    cont = true;
  }
  st.o = o;
  st.d = d;
  st.beta = beta;
  st.rng = rng;
  st.word = word;
  st.pdf = pdf_next;
  return cont;
}

// One shading step over the ray queue `parity`: shade each hit, then
// regenerate finished slots (regen) and compact the continuing paths into the
// other queue.
// Five waves per SIMD (96 VGPRs, no spills), six in the forms for launches
// whose jobs have no environment map (76-78 VGPRs): shade waits on memory most of
// the time, and another wave hides more of it than reading further ahead
// with the registers (c2: 5 % faster, c1 4 %).  Each lane
// reads the queue entry of its next item one item ahead, so that load is off
// the front of the next item's chain of dependent loads.  SOLO: the run has
// one job (A.solo), whose job and scene records are read once, into scalar
// registers, instead of per lane and segment (two dependent loads ahead of
// every other load of the segment).  TAG: the launch of a profiled step (the
// same code under its own name, so a profiler can tell those launches apart).
template <int WG, bool SOLO, bool TAG = false, bool ENV = true>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(!ENV ? 6 : 5))) void k_shade(KernelArgs A, int parity,
                                                                                      int regen) {
  __shared__ uint32_t s_pre[kShards + 1];
  load_prefix(s_pre, A.ctrl->qcount[parity]);
  const uint32_t n = s_pre[kShards];
  const uint32_t stride = gridDim.x * WG;
  const uint32_t qcap = A.p.qcap;
  float4 *ray_o = reinterpret_cast<float4 *>(A.p.ray_o);
  float4 *ray_d = reinterpret_cast<float4 *>(A.p.ray_d);
  uint4 *misc = reinterpret_cast<uint4 *>(A.p.misc);
  uint4 *rngs = reinterpret_cast<uint4 *>(A.p.rng);
  SampleCursor sc = sample_cursor(A.ctrl);
  DoneAcc dacc;
  JobDev Js{};
  SceneDev Ss{};
  if (SOLO) {
    Js = cload(A.jobs + cload(&A.ctrl->solo_job));
    Ss = cload(A.scenes + Js.scene);
  }
  const float4 *hits = reinterpret_cast<const float4 *>(A.p.hit);
  // Read-ahead: while a lane shades item i, the queue entry of item
  // i + stride is in flight.  (Reading the next item's path state ahead as
  // well costs 20 VGPRs; every form spends its registers on a fifth or sixth
  // wave instead, which hides more of the memory latency.)
  auto entry_at = [&](uint32_t j) { return j < n ? A.p.queue[parity][flat_slot(s_pre, j, qcap)] : 0u; };
  struct State {
    float4 ro, rd, hr;
    uint4 ms, rng;
  };
  auto state_of = [&](uint32_t j, uint32_t e, State &to) {
    if (j < n) {
      const uint32_t sl = q_slot(e);
      to.ro = ray_o[sl];
      to.rd = ray_d[sl];
      to.ms = misc[sl];
      to.hr = hits[sl];
      to.rng = rngs[sl];
    }
  };
  uint32_t next = entry_at(blockIdx.x * WG + threadIdx.x);  // entry of the lane's next item
  for (uint32_t base = blockIdx.x * WG; base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    const uint32_t slot = q_slot(next);
    State cur{};
    next = entry_at(i + stride);
    state_of(i, q_entry(slot, 0u), cur);  // (in a block of its own: loading it in the shading
                                          // block below spilled three VGPRs)
    bool cont = false, fin = false, shadow = false;
    PathState st;
    st.js = 0;
    ShadowRay sh;
    if (valid) {
      const float4 ro = cur.ro, rd = cur.rd, hr = cur.hr;
      const uint4 ms = cur.ms;
      st.rng = cur.rng;
      st.o = xyz(ro);
      st.d = xyz(rd);
      st.beta = mk(ro.w, rd.w, __uint_as_float(ms.x));
      st.word = ms.y;
      st.sidx = ms.z;
      st.js = ms.w;
      st.pdf = 0.0f;
      if (SOLO) {
        cont = shade_segment<ENV, ENV>(A, slot, st, hr, A.p.pdf, shadow, sh, Js, Ss);
      } else {
        const JobDev &J = A.jobs[st.js];
        cont = shade_segment<false, ENV>(A, slot, st, hr, A.p.pdf, shadow, sh, J, A.scenes[J.scene]);
      }
      fin = !cont;
    }
    count_done(A.done, dacc, fin, st.js);
    // regenerate finished slots while samples remain
    unsigned long long g = grab_sample(A.ctrl, sc, fin && regen);
    bool fresh = g != ~0ull;
    uint32_t entry = 0;
    if (fresh) {
      entry = start_sample(A, slot, g);
    } else if (cont) {
      // a continuing path keeps its scene (read again rather than kept live)
      entry = q_entry(slot, SOLO ? Js.scene : A.jobs[st.js].scene);
    } else if (fin) {
      misc[slot].y = 0u;  // inactive
    }
    if (valid && cont) {
      ray_o[slot] = make_float4(st.o.x, st.o.y, st.o.z, st.beta.x);
      ray_d[slot] = make_float4(st.d.x, st.d.y, st.d.z, st.beta.y);
      misc[slot] = make_uint4(__float_as_uint(st.beta.z), st.word, st.sidx, st.js);
      rngs[slot] = st.rng;
      if (ps_mis(st.word)) A.p.pdf[slot] = st.pdf;
    }
    bool enq = valid && (cont || fresh);
    const uint32_t k = wave_shard();
    uint32_t pos = wave_append(&A.ctrl->qcount[parity ^ 1][k], enq);
    if (enq) A.p.queue[parity ^ 1][k * qcap + pos] = entry;
    uint32_t spos = wave_append(&A.ctrl->scount[k], shadow);
    if (shadow) {
      float4 *sq = reinterpret_cast<float4 *>(A.p.shadow) + ((size_t)k * qcap + spos) * 3;
      sq[0] = make_float4(sh.o.x, sh.o.y, sh.o.z, __uint_as_float(st.js));
      sq[1] = make_float4(sh.d.x, sh.d.y, sh.d.z, 0.0f);
      sq[2] = make_float4(sh.c.x, sh.c.y, sh.c.z, __uint_as_float(st.sidx));
    }
  }
  done_flush(A.done, dacc);
}

// Accumulates the contribution of shadow record r if its ray was unoccluded.
// SOLO: Js is the run's only job.
template <bool SOLO>
__device__ inline void add_shadow(const KernelArgs &A, uint32_t r, const JobDev &Js) {
  const float4 *sq = reinterpret_cast<const float4 *>(A.p.shadow) + (size_t)r * 3;
  const float4 a = sq[0], c = sq[2];
  const JobDev &J = SOLO ? Js : A.jobs[__float_as_uint(a.w)];
  deposit(J, __float_as_uint(c.w), fix(c.x), fix(c.y), fix(c.z));
}

// Five waves per SIMD with 4-wide nodes and 64-lane groups, as k_extend.
template <int WG, int SHORT, int NW, bool DF, bool SOLO, bool TAG = false>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(NW == 4 && WG == 64 ? 5 : 4))) void k_shadow(
    KernelArgs A) {
  __shared__ uint2 lds[SHORT * WG];
  __shared__ uint32_t s_pre[kShards + 1];
  const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
  load_prefix(s_pre, A.ctrl->scount);
  const uint32_t n = s_pre[kShards];
  const uint32_t stride = gridDim.x * WG;
  const uint32_t qcap = A.p.qcap;
  Stack<WG, SHORT> stk;
  stk.lds = lds;
  stk.g = A.p.gstack + (size_t)gtid * A.p.gcap;
  uint32_t cnt = 0;
  JobDev Js{};  // SOLO: the only job and its scene, in scalar registers
  SceneDev Ss{};
  if (SOLO) {
    Js = cload(A.jobs + cload(&A.ctrl->solo_job));
    Ss = cload(A.scenes + Js.scene);
  }
  auto load = [&](uint32_t k, uint32_t off, const SceneDev *&S, f3 &o, f3 &d, float &tmax, uint32_t &tag) {
    uint32_t r = k * qcap + off;
    const float4 *sq = reinterpret_cast<const float4 *>(A.p.shadow) + (size_t)r * 3;
    float4 a = sq[0], b = sq[1];
    if (SOLO) S = &Ss;
    else S = &A.scenes[A.jobs[__float_as_uint(a.w)].scene];
    o = xyz(a);
    d = xyz(b);
    tmax = kInf;
    tag = r;
  };
  if (DF) {
    auto count = [&](uint32_t k) { return s_pre[k + 1] - s_pre[k]; };
    auto store = [&](uint32_t r, const Hit &h) {
      if (h.prim < 0) add_shadow<SOLO>(A, r, Js);
    };
    uint32_t nv = 0, nt = 0;
    cnt = dynamic_fetch<WG, SHORT, true, false, NW>(A.ctrl->shead, count, stk, (int)A.p.df_budget, A.p.df_fetch_min,
                                               load, store, nv, nt);
  }
  for (uint32_t base = blockIdx.x * WG; !DF && base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    if (i < n) {
      uint32_t r = flat_slot(s_pre, i, qcap);
      const SceneDev *S;
      f3 o, d;
      float tmax;
      uint32_t tag;
      load(r / qcap, r % qcap, S, o, d, tmax, tag);
      Hit h;
      h.t = tmax;
      h.u = h.v = 0.0f;
      h.prim = -1;
      uint32_t nv = 0, nt = 0;
      traverse<WG, SHORT, true, false, NW>(*S, o, d, h, stk, nv, nt);
      if (h.prim < 0) add_shadow<SOLO>(A, r, Js);
      cnt++;
    }
  }
  wave_count(A.ctrl->shadow_rays, cnt);
}

// Drain: once every admitted sample has been handed out and the ray queue
// `parity` is short, each lane carries its paths to the end -- closest hit,
// shading, environment visibility, pending glass branches -- in one launch,
// instead of one extend/shade pair per remaining bounce, each as long as its
// slowest ray.  Same per-path arithmetic and RNG streams as the step
// kernels, and integer accumulation, so the result bits are the same.  The
// queues are left stale: the run ends after a drain.
// Held to 128 VGPRs (four waves per SIMD; nine registers spill, which costs
// less than the fourth wave gains: 6.8 -> 5.9 ms on the Cornell + icosphere
// render).
// Four waves per SIMD (128 VGPRs) with 4-wide nodes; the 8-wide node visit
// holds twice the sort keys and spilled 38-75 VGPRs there, so it runs three.
template <int WG, int SHORT, int NW, bool STATS, bool SOLO, bool ENV = true>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(NW == 8 ? 3 : 4))) void k_drain(KernelArgs A, int parity) {
  __shared__ uint2 lds[SHORT * WG];
  __shared__ uint32_t s_pre[kShards + 1];
  const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
  load_prefix(s_pre, A.ctrl->qcount[parity]);
  const uint32_t n = s_pre[kShards];
  const uint32_t qcap = A.p.qcap;
  Stack<WG, SHORT> stk;
  stk.lds = lds;
  stk.g = A.p.gstack + (size_t)gtid * A.p.gcap;
  const float4 *ray_o = reinterpret_cast<const float4 *>(A.p.ray_o);
  const float4 *ray_d = reinterpret_cast<const float4 *>(A.p.ray_d);
  uint4 *misc = reinterpret_cast<uint4 *>(A.p.misc);
  const uint4 *rngs = reinterpret_cast<const uint4 *>(A.p.rng);
  uint32_t segs = 0, shadows = 0, nv = 0, nt = 0, taken = 0;  // node/primitive counts of the closest-hit traversals
  uint32_t djob = ~0u, dcnt = 0;  // this lane's completions of job djob not yet added
  // Lanes take a new path as soon as theirs ends: a wave-level fetch on the
  // sub-queue heads (zero here: the extend before the last shade cleared
  // them), moving on to other sub-queues when its own is used up.
  Sharded<uint32_t> &heads = A.ctrl->qhead[parity];
  auto count = [&](uint32_t k) { return s_pre[k + 1] - s_pre[k]; };
  uint32_t shard = wave_shard();
  bool exhausted = n == 0u, have = false;
  // a lane's path is either tracing its next segment (a traversal sliced into
  // `budget` visits per trip, as in dynamic fetch) or ready to be shaded; the
  // wave shades once half of its paths are ready or none is tracing
  bool tracing = false, ready = false;
  const int budget = A.p.df_budget > 0 ? (int)A.p.df_budget : 8;
  // fused slab form only where its registers fit (no sky)
  constexpr bool FMA = !ENV;
  Trav<WG, SHORT, false, STATS, NW, FMA> tr;
  uint32_t slot = 0;
  PathState st;
  st.js = 0;
  // SOLO: the run's only job and its scene, in scalar registers
  JobDev Js{};
  SceneDev Ss{};
  if (SOLO) {
    Js = cload(A.jobs + cload(&A.ctrl->solo_job));
    Ss = cload(A.scenes + Js.scene);
  }
  for (;;) {
    unsigned long long m = __ballot(!have);
    uint32_t idle = (uint32_t)__popcll(m);
    if (!exhausted && idle != 0u) {
      uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&heads[shard], idle);
      base = __shfl(base, (int)leader);
      uint32_t cnt = count(shard);
      uint32_t r = rank_in(m);
      if (!have && base + r < cnt) {
        const uint32_t e = A.p.queue[parity][shard * qcap + base + r];
        slot = q_slot(e);
        float4 ro = ray_o[slot], rd = ray_d[slot];
        uint4 ms = misc[slot];
        st.rng = rngs[slot];
        st.o = xyz(ro);
        st.d = xyz(rd);
        st.beta = mk(ro.w, rd.w, __uint_as_float(ms.x));
        st.word = ms.y;
        st.sidx = ms.z;
        st.js = ms.w;
        st.pdf = ps_mis(st.word) ? A.p.pdf[slot] : 0.0f;
        have = true;
        tracing = true;
        taken++;
        tr.init(SOLO ? Ss : A.scenes[q_scene(e)], st.o, st.d, kInf, stk);
      }
      if (base + idle >= cnt) {  // this sub-queue is used up: find one with entries left
        const uint32_t lane = lane_id();
        bool left = false;
        if (lane < kShards) left = __hip_atomic_load(&heads[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < count(lane);
        unsigned long long lm = __ballot(left);
        if (lm == 0ull) {
          exhausted = true;
        } else {
          unsigned long long after = lm & ~((2ull << shard) - 1ull);
          shard = (uint32_t)(__ffsll((long long)(after ? after : lm)) - 1);
        }
      }
    }
    const unsigned long long hm = __ballot(have);
    if (hm == 0ull) {
      if (exhausted) break;
      continue;
    }
    if (tracing && tr.step(stk, budget, nv, nt)) {
      tracing = false;
      ready = true;
    }
    const unsigned long long rm = __ballot(ready);
    if (rm == 0ull || (__ballot(tracing) != 0ull && 2u * (uint32_t)__popcll(rm) < (uint32_t)__popcll(hm))) continue;
    if (!ready) continue;
    ready = false;
    const JobDev &J = SOLO ? Js : A.jobs[st.js];
    const SceneDev &S = SOLO ? Ss : A.scenes[J.scene];
    segs++;
    bool shadow = false;
    ShadowRay sh;
    // the segment's origin and direction live in the traversal only, so one
    // copy of them stays in registers across the loop
    st.o = tr.o;
    st.d = tr.d;
    bool cont = shade_segment<false, ENV>(A, slot, st, make_float4(tr.best.t, __int_as_float(tr.best.prim), tr.best.u, tr.best.v),
                              nullptr, shadow, sh, J, S);
    if (shadow) {
      Hit hs;
      hs.t = kInf;
      hs.u = hs.v = 0.0f;
      hs.prim = -1;
      uint32_t snv = 0, snt = 0;
      traverse<WG, SHORT, true, false, NW, FMA>(S, sh.o, sh.d, hs, stk, snv, snt);
      shadows++;
      if (hs.prim < 0) deposit(J, st.sidx, fix(sh.c.x), fix(sh.c.y), fix(sh.c.z));
    }
    if (cont) {
      tracing = true;
      tr.init(S, st.o, st.d, kInf, stk);
    } else {
      misc[slot].y = 0u;  // inactive
      if (st.js != djob) {
        if (dcnt) atomicAdd(&A.done[done_index(wave_shard(), djob)], (unsigned long long)dcnt);
        djob = st.js;
        dcnt = 0;
      }
      dcnt++;
      have = false;
    }
  }
  // the remaining completion counts, one atomic per job present in the wave
  bool need = dcnt != 0u;
  for (;;) {
    unsigned long long m = __ballot(need);
    if (m == 0ull) break;
    uint32_t leader = (uint32_t)(__ffsll((long long)m) - 1);
    uint32_t lj = __shfl(djob, (int)leader);
    bool mine = need && djob == lj;
    unsigned long long c = mine ? dcnt : 0u;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if (lane_id() == leader) atomicAdd(&A.done[done_index(wave_shard(), lj)], c);
    if (mine) need = false;
  }
  wave_count(A.ctrl->segments, segs);
  wave_count(A.ctrl->drain_segments, segs);
  wave_count(A.ctrl->drain_paths, taken);
  wave_count(A.ctrl->shadow_rays, shadows);
  if (STATS) {
    wave_count(A.ctrl->drain_node_visits, nv);
    wave_count(A.ctrl->drain_prim_tests, nt);
  }
}

// Clears the shadow sub-queue counts and fetch heads at the start of a step.
__global__ void k_clear_shadow(Ctrl *ctrl) {
  if (threadIdx.x < kShards) {
    ctrl->scount[threadIdx.x] = 0u;
    ctrl->shead[threadIdx.x] = 0u;
  }
}

__global__ void k_clear_done(unsigned long long *done, uint32_t ring) {
  if (threadIdx.x < kShards) done[done_index(threadIdx.x, ring)] = 0ull;
}

// One wave: lane k clamps cursor shard k to the end of the old range.
__global__ void k_admit(Ctrl *ctrl, unsigned long long old_total, unsigned long long new_total, uint32_t lo,
                        uint32_t hi) {
  const uint32_t k = threadIdx.x;
  if (k < kShards) {
    const unsigned long long lim = shard_limit(old_total, k);
    if (ctrl->next_sample[k] > lim) ctrl->next_sample[k] = lim;
  }
  if (k == 0) {
    ctrl->total_samples = new_total;
    ctrl->job_lo = lo;
    ctrl->job_hi = hi;
  }
}

// First wave: the queue length, cursor state and ray totals (lane k reads
// shard k); every thread: the completion sums of live jobs t, t + 1024, ...
__global__ __launch_bounds__(1024) void k_serve_status(const Ctrl *ctrl, const unsigned long long *done, uint32_t lo,
                                                        uint32_t n, ServeStatus *out) {
  const uint32_t t = threadIdx.x;
  if (t < 64u) {
    unsigned long long q = 0ull, seg = 0ull, sh = 0ull;
    bool left = false;
    if (t < kShards) {
      q = ctrl->qcount[0][t];
      seg = ctrl->segments[t];
      sh = ctrl->shadow_rays[t];
      left = ctrl->next_sample[t] < shard_limit(ctrl->total_samples, t);
    }
    const unsigned long long any_left = __ballot(left);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      q += __shfl_xor(q, off);
      seg += __shfl_xor(seg, off);
      sh += __shfl_xor(sh, off);
    }
    if (t == 0u) {
      out->queued = q;
      out->exhausted = any_left == 0ull ? 1ull : 0ull;
      out->segments = seg;
      out->shadow_rays = sh;
    }
  }
  for (uint32_t i = t; i < n; i += 1024u) {
    const uint32_t ring = (lo + i) % kJobRing;
    unsigned long long s = 0ull;
    for (uint32_t k = 0; k < kShards; k++) s += done[done_index(k, ring)];
    out->done[i] = s;
  }
}

// Pixel sums from per-sample records: sample s belongs to pixel s / spp_px,
// so each pixel's records are contiguous; one wave sums a pixel's records
// (lanes stride through them) and writes its three sums.
__global__ __launch_bounds__(256) void k_gather_records(const unsigned long long *recs, uint32_t spp_px,
                                                        uint32_t n_px, unsigned long long *accum) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nw = gridDim.x * 4u;
  for (uint32_t p = blockIdx.x * 4u + (threadIdx.x >> 6); p < n_px; p += nw) {
    const __attribute__((address_space(1))) unsigned long long *r = gptr(recs) + 3ull * spp_px * p;
    unsigned long long a = 0ull, b = 0ull, c = 0ull;
    for (uint32_t k = lane; k < spp_px; k += 64u) {
      a += r[3ull * k];
      b += r[3ull * k + 1ull];
      c += r[3ull * k + 2ull];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
      c += __shfl_xor(c, off);
    }
    if (lane == 0u) {
      accum[3ull * p] = a;
      accum[3ull * p + 1ull] = b;
      accum[3ull * p + 2ull] = c;
    }
  }
}

// mean = sum * 2^-24 / (ss^2 * samples), in double, as the CPU restatement
__global__ void k_resolve(const unsigned long long *accum, float *out, uint32_t n, double n_samples) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)((double)(long long)accum[i] * (1.0 / 16777216.0) / n_samples);
}

template <int WG, int SHORT, int NW, bool STATS, bool DF>
__global__ __launch_bounds__(WG) void k_trace_rays(const SceneDev *scene, const float4 *rays, uint32_t n, int32_t *prim,
                                                   float *tout, unsigned long long *counters, uint2 *gstack,
                                                   uint32_t gcap, Sharded<uint32_t> *head, int budget, uint32_t fetch_min) {
  __shared__ uint2 lds[SHORT * WG];
  const uint32_t gtid = blockIdx.x * WG + threadIdx.x;
  const uint32_t stride = gridDim.x * WG;
  Stack<WG, SHORT> stk;
  stk.lds = lds;
  stk.g = gstack + (size_t)gtid * gcap;
  const SceneDev S = *scene;
  uint32_t nv = 0, nt = 0;
  if (DF) {
    // sub-queue k = rays [k n / kShards, (k+1) n / kShards)
    auto first = [&](uint32_t k) { return (uint32_t)(((unsigned long long)n * k) / kShards); };
    auto count = [&](uint32_t k) { return first(k + 1) - first(k); };
    auto load = [&](uint32_t k, uint32_t off, const SceneDev *&Sp, f3 &o, f3 &d, float &tmax, uint32_t &tag) {
      uint32_t i = first(k) + off;
      float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
      Sp = scene;
      o = xyz(a);
      d = xyz(b);
      tmax = a.w;
      tag = i;
    };
    auto store = [&](uint32_t i, const Hit &h) {
      prim[i] = h.prim;
      tout[i] = h.prim >= 0 ? h.t : kInf;
    };
    dynamic_fetch<WG, SHORT, false, STATS, NW>(*head, count, stk, budget, fetch_min, load, store, nv, nt);
  }
  for (uint32_t base = blockIdx.x * WG; !DF && base < n; base += stride) {
    uint32_t i = base + threadIdx.x;
    if (i < n) {
      float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
      Hit h;
      h.t = a.w;
      h.u = h.v = 0.0f;
      h.prim = -1;
      traverse<WG, SHORT, false, STATS, NW>(S, xyz(a), xyz(b), h, stk, nv, nt);
      prim[i] = h.prim;
      tout[i] = h.prim >= 0 ? h.t : kInf;
    }
  }
  if (STATS) {
    unsigned long long a = nv, b = nt;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane_id() == 0) {
      atomicAdd(&counters[0], a);
      atomicAdd(&counters[1], b);
    }
  }
}

// ------------------------------------------------------------ environment tables
// Per-texel weight = luminance * sin(theta at the row centre); max by integer
// atomicMax on the (non-negative) float bits.
__global__ void k_env_weight(const float *env, uint32_t w, uint32_t h, float *wbuf, unsigned int *wmax_bits) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float wv = 0.0f;
  if (i < w * h) {
    uint32_t iy = i / w;
    float s, c;
    sincos_turns(((float)iy + 0.5f) / (2.0f * (float)h), s, c);
    const float *p = env + 3 * (size_t)i;
    float lum = (0.2126f * p[0] + 0.7152f * p[1]) + 0.0722f * p[2];
    wv = lum * s;
    if (!(wv > 0.0f)) wv = 0.0f;
    wbuf[i] = wv;
  }
  float m = wv;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    float o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m > 0.0f) atomicMax(wmax_bits, __float_as_uint(m));
}

// One workgroup (256 threads) per row: quantize, inclusive scan in LDS.
__global__ __launch_bounds__(256) void k_env_rowscan(const float *wbuf, uint32_t w, const unsigned int *wmax_bits,
                                                     unsigned long long *q, unsigned long long *row) {
  __shared__ unsigned long long part[256];
  uint32_t iy = blockIdx.x;
  float wmax = __uint_as_float(*wmax_bits);
  double scale = wmax > 0.0f ? 4294967295.0 / (double)wmax : 0.0;
  uint32_t per = (w + 255u) / 256u;
  uint32_t b = threadIdx.x * per, e = b + per < w ? b + per : w;
  unsigned long long local = 0ull;
  for (uint32_t ix = b; ix < e; ix++) {
    unsigned long long v = (unsigned long long)((double)wbuf[(size_t)iy * w + ix] * scale);
    q[(size_t)iy * w + ix] = v;
    local += v;
  }
  part[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t off = 1; off < 256u; off <<= 1) {
    unsigned long long add_v = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += add_v;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0ull;
  unsigned long long *r = row + (size_t)iy * (w + 1);
  if (threadIdx.x == 0) r[0] = 0ull;
  for (uint32_t ix = b; ix < e; ix++) {
    run += q[(size_t)iy * w + ix];
    r[ix + 1] = run;
  }
}

// Single workgroup: prefix over the row totals.
__global__ __launch_bounds__(256) void k_env_margscan(const unsigned long long *row, uint32_t w, uint32_t h,
                                                      unsigned long long *marg) {
  __shared__ unsigned long long part[256];
  uint32_t per = (h + 255u) / 256u;
  uint32_t b = threadIdx.x * per, e = b + per < h ? b + per : h;
  unsigned long long local = 0ull;
  for (uint32_t iy = b; iy < e; iy++) local += row[(size_t)iy * (w + 1) + w];
  part[threadIdx.x] = local;
  __syncthreads();
  for (uint32_t off = 1; off < 256u; off <<= 1) {
    unsigned long long add_v = threadIdx.x >= off ? part[threadIdx.x - off] : 0ull;
    __syncthreads();
    part[threadIdx.x] += add_v;
    __syncthreads();
  }
  unsigned long long run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0ull;
  if (threadIdx.x == 0) marg[0] = 0ull;
  for (uint32_t iy = b; iy < e; iy++) {
    run += row[(size_t)iy * (w + 1) + w];
    marg[iy + 1] = run;
  }
}

// Bucket directory of one inclusive prefix table a[0..n] (a[0] = 0): entry
// i serves the keys [a[i], a[i+1]), whose buckets are env_bucket(a[i]) ..
// env_bucket(a[i+1] - 1); each bucket records the first and last such entry.
// One thread per table: the rows, then the marginal (index h).
__global__ void k_env_guides(const unsigned long long *row, const unsigned long long *marg, uint32_t w, uint32_t h,
                             uint2_t *rguide, uint2_t *mguide, double *rscale, double *mscale) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > h) return;
  const bool m = t == h;
  const unsigned long long *a = m ? marg : row + (size_t)t * (w + 1);
  const uint32_t n = m ? h : w;
  uint2_t *g = m ? mguide : rguide + (size_t)t * w;
  const unsigned long long total = a[n];
  const double scale = total > 0ull ? (double)n / (double)total : 0.0;
  if (m) {
    mscale[0] = scale;
    mscale[1] = total > 0ull ? ((double)w * (double)h) / (19.739208802178716 * (double)total) : 0.0;
  } else {
    rscale[t] = scale;
  }
  for (uint32_t b = 0; b < n; b++) g[b] = uint2_t{0xFFFFFFFFu, 0u};
  if (total == 0ull) return;
  for (uint32_t i = 0; i < n; i++) {
    if (a[i + 1] <= a[i]) continue;
    uint32_t b0 = env_bucket(a[i], scale, n), b1 = env_bucket(a[i + 1] - 1ull, scale, n);
    for (uint32_t b = b0; b <= b1; b++) {
      if (i < g[b].lo) g[b].lo = i;
      g[b].hi = i;  // i only grows
    }
  }
}

// The launchers pick the template instance of the run's launch choices:
// group size, LDS stack depth (a.p.stack_short), node width (a.p.node_width),
// fetch mode, statistics, and the single-job form.
template <int WG, int SHORT, int NW>
hipError_t launch_extend_n(const KernelArgs &a, uint32_t grid, int parity, bool stats, hipStream_t st) {
  bool df = a.p.df_budget > 0;
  if (WG == 64 && df && a.solo != kNoSolo) {  // single-job runs at the default group size and fetch mode
    if (stats) hipLaunchKernelGGL((k_extend<WG, SHORT, NW, true, true, true>), dim3(grid), dim3(WG), 0, st, a, parity);
    else hipLaunchKernelGGL((k_extend<WG, SHORT, NW, false, true, true>), dim3(grid), dim3(WG), 0, st, a, parity);
    return hipGetLastError();
  }
  if (stats && df)
    hipLaunchKernelGGL((k_extend<WG, SHORT, NW, true, true, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  else if (stats)
    hipLaunchKernelGGL((k_extend<WG, SHORT, NW, true, false, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  else if (df)
    hipLaunchKernelGGL((k_extend<WG, SHORT, NW, false, true, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  else hipLaunchKernelGGL((k_extend<WG, SHORT, NW, false, false, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  return hipGetLastError();
}
template <int WG, int SHORT>
hipError_t launch_extend_ws(const KernelArgs &a, uint32_t grid, int parity, bool stats, hipStream_t st) {
  if (a.p.node_width == 8) return launch_extend_n<WG, SHORT, 8>(a, grid, parity, stats, st);
  return launch_extend_n<WG, SHORT, 4>(a, grid, parity, stats, st);
}
template <int WG>
hipError_t launch_extend_wg(const KernelArgs &a, uint32_t grid, int parity, bool stats, hipStream_t st) {
  if (a.p.stack_short == kShortStackMin) return launch_extend_ws<WG, (int)kShortStackMin>(a, grid, parity, stats, st);
  return launch_extend_ws<WG, (int)kShortStackMax>(a, grid, parity, stats, st);
}

template <int WG, int SHORT, int NW>
hipError_t launch_shadow_n(const KernelArgs &a, uint32_t grid, bool tag, hipStream_t st) {
  if (WG == 64 && a.p.df_budget > 0 && a.solo != kNoSolo && tag)  // profiled single-job step
    hipLaunchKernelGGL((k_shadow<WG, SHORT, NW, true, true, true>), dim3(grid), dim3(WG), 0, st, a);
  else if (WG == 64 && a.p.df_budget > 0 && a.solo != kNoSolo)  // single-job runs, default group size and fetch mode
    hipLaunchKernelGGL((k_shadow<WG, SHORT, NW, true, true>), dim3(grid), dim3(WG), 0, st, a);
  else if (a.p.df_budget > 0) hipLaunchKernelGGL((k_shadow<WG, SHORT, NW, true, false>), dim3(grid), dim3(WG), 0, st, a);
  else hipLaunchKernelGGL((k_shadow<WG, SHORT, NW, false, false>), dim3(grid), dim3(WG), 0, st, a);
  return hipGetLastError();
}
template <int WG>
hipError_t launch_shadow_wg(const KernelArgs &a, uint32_t grid, bool tag, hipStream_t st) {
  const bool s8 = a.p.stack_short == kShortStackMin, w8 = a.p.node_width == 8;
  constexpr int S8 = (int)kShortStackMin, S16 = (int)kShortStackMax;
  if (s8) return w8 ? launch_shadow_n<WG, S8, 8>(a, grid, tag, st) : launch_shadow_n<WG, S8, 4>(a, grid, tag, st);
  return w8 ? launch_shadow_n<WG, S16, 8>(a, grid, tag, st) : launch_shadow_n<WG, S16, 4>(a, grid, tag, st);
}

template <int WG, int SHORT, int NW>
hipError_t launch_drain_n(const KernelArgs &a, uint32_t grid, int parity, bool stats, hipStream_t st) {
  if (WG == 64 && a.solo != kNoSolo) {  // single-job runs at the default group size
    if (a.any_env) {
      if (stats) hipLaunchKernelGGL((k_drain<WG, SHORT, NW, true, true, true>), dim3(grid), dim3(WG), 0, st, a, parity);
      else hipLaunchKernelGGL((k_drain<WG, SHORT, NW, false, true, true>), dim3(grid), dim3(WG), 0, st, a, parity);
    } else {
      if (stats) hipLaunchKernelGGL((k_drain<WG, SHORT, NW, true, true, false>), dim3(grid), dim3(WG), 0, st, a, parity);
      else hipLaunchKernelGGL((k_drain<WG, SHORT, NW, false, true, false>), dim3(grid), dim3(WG), 0, st, a, parity);
    }
    return hipGetLastError();
  }
  if (WG == 64 && !a.any_env) {  // no job of the launch has an environment map
    if (stats) hipLaunchKernelGGL((k_drain<WG, SHORT, NW, true, false, false>), dim3(grid), dim3(WG), 0, st, a, parity);
    else hipLaunchKernelGGL((k_drain<WG, SHORT, NW, false, false, false>), dim3(grid), dim3(WG), 0, st, a, parity);
    return hipGetLastError();
  }
  if (stats) hipLaunchKernelGGL((k_drain<WG, SHORT, NW, true, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  else hipLaunchKernelGGL((k_drain<WG, SHORT, NW, false, false>), dim3(grid), dim3(WG), 0, st, a, parity);
  return hipGetLastError();
}
template <int WG>
hipError_t launch_drain_wg(const KernelArgs &a, uint32_t grid, int parity, bool stats, hipStream_t st) {
  const bool s8 = a.p.stack_short == kShortStackMin, w8 = a.p.node_width == 8;
  constexpr int S8 = (int)kShortStackMin, S16 = (int)kShortStackMax;
  if (s8)
    return w8 ? launch_drain_n<WG, S8, 8>(a, grid, parity, stats, st) : launch_drain_n<WG, S8, 4>(a, grid, parity, stats, st);
  return w8 ? launch_drain_n<WG, S16, 8>(a, grid, parity, stats, st)
            : launch_drain_n<WG, S16, 4>(a, grid, parity, stats, st);
}

template <int WG, int SHORT, int NW, bool STATS>
hipError_t launch_trace_n(const SceneDev *scene, const float4 *r4, uint32_t n, int32_t *prim, float *t,
                          unsigned long long *counters, uint2 *gstack, uint32_t gcap, uint32_t grid,
                          Sharded<uint32_t> *head, int budget, uint32_t fetch_min, hipStream_t st) {
  if (budget > 0)
    hipLaunchKernelGGL((k_trace_rays<WG, SHORT, NW, STATS, true>), dim3(grid), dim3(WG), 0, st, scene, r4, n, prim, t,
                       counters, gstack, gcap, head, budget, fetch_min);
  else
    hipLaunchKernelGGL((k_trace_rays<WG, SHORT, NW, STATS, false>), dim3(grid), dim3(WG), 0, st, scene, r4, n, prim,
                       t, counters, gstack, gcap, head, budget, fetch_min);
  return hipGetLastError();
}
template <int WG, bool STATS>
hipError_t launch_trace_wg(const SceneDev *scene, const float4 *r4, uint32_t n, int32_t *prim, float *t,
                           unsigned long long *counters, uint2 *gstack, uint32_t gcap, uint32_t stack_short,
                           uint32_t node_width, uint32_t grid, Sharded<uint32_t> *head, int budget, uint32_t fetch_min,
                           hipStream_t st) {
  constexpr int S8 = (int)kShortStackMin, S16 = (int)kShortStackMax;
  const bool s8 = stack_short == kShortStackMin, w8 = node_width == 8;
#define OMR_TRACE(S, W) \
  launch_trace_n<WG, S, W, STATS>(scene, r4, n, prim, t, counters, gstack, gcap, grid, head, budget, fetch_min, st)
  if (s8) return w8 ? OMR_TRACE(S8, 8) : OMR_TRACE(S8, 4);
  return w8 ? OMR_TRACE(S16, 8) : OMR_TRACE(S16, 4);
#undef OMR_TRACE
}

__global__ void k_zero_words(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

// The allocation-free entry point's scratch (fetch heads, then overflow
// stacks).  Host code reaches it only through its device address
// (k::noalloc_scratch); the name itself is a host-side stand-in.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace k {

hipError_t noalloc_scratch(void **ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

hipError_t zero_words(uint32_t *p, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (n + 255u) / 256u;
  hipLaunchKernelGGL(k_zero_words, dim3(blocks < 1024u ? blocks : 1024u), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t fill_fresh(const KernelArgs &a, unsigned long long base, uint32_t n0, hipStream_t st) {
  if (n0 == 0) return hipSuccess;
  const uint32_t need = (n0 + 255u) / 256u;
  hipLaunchKernelGGL(k_fill_fresh, dim3(need < 16384u ? need : 16384u), dim3(256), 0, st, a, base, n0);
  return hipGetLastError();
}

hipError_t round_setup(Ctrl *ctrl, uint32_t n0, uint32_t solo_job, hipStream_t st) {
  hipLaunchKernelGGL(k_round_setup, dim3(1), dim3(kShards), 0, st, ctrl, n0, solo_job);
  return hipGetLastError();
}

hipError_t refill(const KernelArgs &a, const LaunchShape &s, int parity, hipStream_t st) {
  switch (s.wg) {
    case 64: hipLaunchKernelGGL(k_refill<64>, dim3(s.grid), dim3(64), 0, st, a, parity); break;
    case 128: hipLaunchKernelGGL(k_refill<128>, dim3(s.grid), dim3(128), 0, st, a, parity); break;
    default: hipLaunchKernelGGL(k_refill<256>, dim3(s.grid), dim3(256), 0, st, a, parity); break;
  }
  return hipGetLastError();
}

hipError_t extend(const KernelArgs &a, const LaunchShape &s, int parity, bool stats, hipStream_t st) {
  switch (s.wg) {
    case 64: return launch_extend_wg<64>(a, s.grid, parity, stats, st);
    case 128: return launch_extend_wg<128>(a, s.grid, parity, stats, st);
    default: return launch_extend_wg<256>(a, s.grid, parity, stats, st);
  }
}

hipError_t shade(const KernelArgs &a, const LaunchShape &s, int parity, bool regen, bool tag, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_shadow, dim3(1), dim3(kShards), 0, st, a.ctrl);
  int r = regen ? 1 : 0;
  const bool solo = a.solo != kNoSolo;
  switch (s.wg) {
    case 64:
      if (solo && a.any_env && tag)
        hipLaunchKernelGGL((k_shade<64, true, true>), dim3(s.shade_grid), dim3(64), 0, st, a, parity, r);
      else if (solo && a.any_env)
        hipLaunchKernelGGL((k_shade<64, true>), dim3(s.shade_grid), dim3(64), 0, st, a, parity, r);
      else if (solo && tag)
        hipLaunchKernelGGL((k_shade<64, true, true, false>), dim3(s.shade6_grid), dim3(64), 0, st, a, parity, r);
      else if (solo)
        hipLaunchKernelGGL((k_shade<64, true, false, false>), dim3(s.shade6_grid), dim3(64), 0, st, a, parity, r);
      else if (!a.any_env)
        hipLaunchKernelGGL((k_shade<64, false, false, false>), dim3(s.shade6_grid), dim3(64), 0, st, a, parity, r);
      else hipLaunchKernelGGL((k_shade<64, false>), dim3(s.shade_grid), dim3(64), 0, st, a, parity, r);
      break;
    case 128:
      if (solo) hipLaunchKernelGGL((k_shade<128, true>), dim3(s.shade_grid), dim3(128), 0, st, a, parity, r);
      else hipLaunchKernelGGL((k_shade<128, false>), dim3(s.shade_grid), dim3(128), 0, st, a, parity, r);
      break;
    default:
      if (solo) hipLaunchKernelGGL((k_shade<256, true>), dim3(s.shade_grid), dim3(256), 0, st, a, parity, r);
      else hipLaunchKernelGGL((k_shade<256, false>), dim3(s.shade_grid), dim3(256), 0, st, a, parity, r);
      break;
  }
  return hipGetLastError();
}

hipError_t shadow(const KernelArgs &a, const LaunchShape &s, bool tag, hipStream_t st) {
  switch (s.wg) {
    case 64: return launch_shadow_wg<64>(a, s.grid, tag, st);
    case 128: return launch_shadow_wg<128>(a, s.grid, false, st);
    default: return launch_shadow_wg<256>(a, s.grid, false, st);
  }
}

hipError_t drain(const KernelArgs &a, const LaunchShape &s, int parity, bool stats, hipStream_t st) {
  switch (s.wg) {
    case 64: return launch_drain_wg<64>(a, s.grid, parity, stats, st);
    case 128: return launch_drain_wg<128>(a, s.grid, parity, stats, st);
    default: return launch_drain_wg<256>(a, s.grid, parity, stats, st);
  }
}

hipError_t clear_done(unsigned long long *done, uint32_t ring, hipStream_t st) {
  hipLaunchKernelGGL(k_clear_done, dim3(1), dim3(kShards), 0, st, done, ring);
  return hipGetLastError();
}
hipError_t admit(Ctrl *ctrl, unsigned long long old_total, unsigned long long new_total, uint32_t lo, uint32_t hi,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_admit, dim3(1), dim3(kShards), 0, st, ctrl, old_total, new_total, lo, hi);
  return hipGetLastError();
}
hipError_t serve_status(const Ctrl *ctrl, const unsigned long long *done, uint32_t lo, uint32_t n, ServeStatus *out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_serve_status, dim3(1), dim3(1024), 0, st, ctrl, done, lo, n, out);
  return hipGetLastError();
}
hipError_t gather_records(const unsigned long long *recs, uint32_t spp_px, uint32_t n_px, unsigned long long *accum,
                          hipStream_t st) {
  if (n_px == 0) return hipSuccess;
  const uint32_t grid = n_px < 262144u ? (n_px + 3u) / 4u : 65536u;
  hipLaunchKernelGGL(k_gather_records, dim3(grid), dim3(256), 0, st, recs, spp_px, n_px, accum);
  return hipGetLastError();
}

hipError_t resolve(const unsigned long long *accum, float *out, uint32_t n_px, double n_samples, hipStream_t st) {
  uint32_t n = n_px * 3u;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_resolve, dim3((n + 255u) / 256u), dim3(256), 0, st, accum, out, n, n_samples);
  return hipGetLastError();
}

hipError_t trace_rays(const SceneDev *scene, const float *rays, uint32_t n, int32_t *prim, float *t,
                      unsigned long long *counters, uint2 *gstack, uint32_t gcap, uint32_t stack_short,
                      uint32_t node_width, uint32_t grid, uint32_t wg, Sharded<uint32_t> *head, int budget,
                      uint32_t fetch_min, hipStream_t st) {
  const float4 *r4 = reinterpret_cast<const float4 *>(rays);
  const uint32_t ss = stack_short, nw = node_width;
  if (counters)
    return launch_trace_wg<128, true>(scene, r4, n, prim, t, counters, gstack, gcap, ss, nw, grid, head, budget,
                                      fetch_min, st);
  switch (wg) {
    case 64:
      return launch_trace_wg<64, false>(scene, r4, n, prim, t, nullptr, gstack, gcap, ss, nw, grid, head, budget,
                                        fetch_min, st);
    case 256:
      return launch_trace_wg<256, false>(scene, r4, n, prim, t, nullptr, gstack, gcap, ss, nw, grid, head, budget,
                                         fetch_min, st);
    default:
      return launch_trace_wg<128, false>(scene, r4, n, prim, t, nullptr, gstack, gcap, ss, nw, grid, head, budget,
                                         fetch_min, st);
  }
}

hipError_t env_tables(const float *env, uint32_t w, uint32_t h, float *wbuf, unsigned int *wmax_bits,
                      unsigned long long *q, unsigned long long *row, unsigned long long *marg, hipStream_t st) {
  uint32_t n = w * h;
  hipLaunchKernelGGL(k_env_weight, dim3((n + 255u) / 256u), dim3(256), 0, st, env, w, h, wbuf, wmax_bits);
  hipLaunchKernelGGL(k_env_rowscan, dim3(h), dim3(256), 0, st, wbuf, w, wmax_bits, q, row);
  hipLaunchKernelGGL(k_env_margscan, dim3(1), dim3(256), 0, st, row, w, h, marg);
  return hipGetLastError();
}

__global__ void k_env_pack(const float *env, const unsigned long long *q, uint32_t n, float4 *env4) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float *p = env + 3 * (size_t)i;
    env4[i] = make_float4(p[0], p[1], p[2], __uint_as_float((uint32_t)q[i]));  // q < 2^32 (k_env_rowscan)
  }
}
hipError_t env_pack(const float *env, const unsigned long long *q, uint32_t n, float4 *env4, hipStream_t st) {
  hipLaunchKernelGGL(k_env_pack, dim3((n + 255u) / 256u), dim3(256), 0, st, env, q, n, env4);
  return hipGetLastError();
}

hipError_t env_guides(const unsigned long long *row, const unsigned long long *marg, uint32_t w, uint32_t h,
                      uint2_t *rguide, uint2_t *mguide, double *rscale, double *mscale, hipStream_t st) {
  hipLaunchKernelGGL(k_env_guides, dim3((h + 1u + 63u) / 64u), dim3(64), 0, st, row, marg, w, h, rguide, mguide,
                     rscale, mscale);
  return hipGetLastError();
}

}  // namespace k
}  // namespace omr
