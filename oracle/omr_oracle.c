/*
 * omr_oracle.c -- CPU restatement of omochi-renderer3's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Never linked into
 * or called by the product library.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math (see oracle/Makefile).  All
 * arithmetic is IEEE binary32 with the operation order spelled out below; the
 * only double-precision steps are the ones marked "(double)".
 *
 * Pinning of each function (what checks it besides the GPU comparison):
 *   orc_rng_*            reference Random (src/tools/Random.h:9-29) compiled
 *                        from the reference sources -> tests/golden/rng_kat.json
 *   orc_path_seed        project-defined per-path seeding (the reference seeds
 *                        one stream per row, PathTracer.cpp:151, so no
 *                        reference vector exists); pinned by what the hash
 *                        must satisfy: every argument changes the seed, and
 *                        the streams of neighbouring pixels, sub-pixels,
 *                        samples, branches and the lens stream are distinct
 *                        and decorrelated (tests/test_oracle_math.py).
 *   orc_sincos_turns     closed form: sin^2+cos^2, libm comparison < 2e-7
 *   orc_acos             closed form: libm comparison < 4e-7
 *   orc_intersect_*      compiled reference Polygon/Sphere (FP64) golden hits
 *                        + hand-built closed-form cases
 *   orc_trace            brute force == own BVH == external wide BVH
 *   orc_intersect_plane  compiled reference InfiniteFloor golden hits
 *                        + closed-form cases
 *   orc_sphere_uv        compiled reference Sphere uv golden vectors + axes
 *   orc_texel_xy         compiled reference Image::GetPixelByUV golden vectors
 *   orc_env_texel        compiled reference IBL::Sample texel golden vectors
 *   orc_env_cdf/sample   chi^2 histogram + pdf integrates to 1 (closed form)
 *   orc_render           white furnace (closed form), compiled reference
 *                        PathTracer statistical comparison (z-score) on the
 *                        config-0 box, the reference's CornellBoxScene (with
 *                        and without next-event estimation), a Lambert sphere
 *                        under the reference's HDR and textured.scene
 */
#include "omr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* src/tools/Constant.h:4-8 -- INF = (FLT_MAX-1)/2, EPS = 0.0005 */
#define ORC_INF 0x1.fffffep+126f
#define ORC_EPS 0.0005f
#define ORC_PI 3.14159265358979323846f
#define ORC_TWO_PI 6.28318530717958647692f
#define ORC_INV_PI 0.31830988618379067154f
/* Fixed-point scale of the per-pixel accumulators: 2^24 units per radiance unit. */
#define ORC_FIX_SCALE 16777216.0f
#define ORC_FIX_CLAMP 1.0e8f

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
/* Vector3::dot (src/tools/Vector.h:66-68): x*r.x + y*r.y + z*r.z, left to right */
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* Vector3::cross (src/tools/Vector.h:70-76) */
static inline v3 cross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* Vector3::normalize (src/tools/Vector.h:83-86) divides each component by the
 * length; here one correctly rounded reciprocal of the length and three
 * products (at most an ulp apart per component; the GPU path uses the same
 * form, so the two stay bit-identical) */
static inline v3 normalize(v3 a) {
  float inv = 1.0f / sqrtf(dot(a, a));
  return mk(a.x * inv, a.y * inv, a.z * inv);
}
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }

/* ------------------------------------------------------------------ RNG */
/* Random(seed) (src/tools/Random.h:9-15): s = 1812433253*(s^(s>>30)) + i */
void orc_rng_init(uint32_t seed, uint32_t st[4]) {
  uint32_t s = seed;
  for (uint32_t i = 1; i <= 4; i++) {
    s = 1812433253u * (s ^ (s >> 30)) + i;
    st[i - 1] = s;
  }
}
/* Random::next (src/tools/Random.h:17-23): xorshift128 */
uint32_t orc_rng_next(uint32_t st[4]) {
  uint32_t t = st[0] ^ (st[0] << 11);
  st[0] = st[1];
  st[1] = st[2];
  st[2] = st[3];
  st[3] = (st[3] ^ (st[3] >> 19)) ^ (t ^ (t >> 8));
  return st[3];
}
/* Random::nextDouble (src/tools/Random.h:25-27) = next()/UINT_MAX in [0,1];
 * restated in binary32: (float)next() / 4294967295.0f. */
float orc_rng_float(uint32_t st[4]) { return (float)orc_rng_next(st) / 4294967295.0f; }

/* Per-path seeding (replaces the reference's per-row seed y+1+prev*H,
 * src/renderer/PathTracer.cpp:151).  Integer hash chain, lowbias32 mixer. */
static inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16; h *= 0x7feb352du;
  h ^= h >> 15; h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}
static inline uint32_t combine(uint32_t h, uint32_t v) { return mix32(h ^ (v + 0x9e3779b9u + (h << 6) + (h >> 2))); }
uint32_t orc_path_seed(uint64_t seed, uint32_t pixel, uint32_t sub, uint32_t sample, uint32_t branch) {
  uint32_t h = mix32((uint32_t)seed ^ 0x6f6d7233u);
  h = combine(h, (uint32_t)(seed >> 32));
  h = combine(h, pixel);
  h = combine(h, sub);
  h = combine(h, sample);
  h = combine(h, branch);
  return h;
}

/* ------------------------------------------------------- elementary math */
/* sin and cos of 2*pi*x, x in [0,1]: quadrant reduction + Taylor polynomials
 * on [-pi/4, pi/4] (replaces libm cos/sin at src/renderer/PathTracer.cpp:489). */
void orc_sincos_turns(float x, float *s_out, float *c_out) {
  float q = floorf(x * 4.0f + 0.5f);
  float f = x - q * 0.25f;
  float a = f * ORC_TWO_PI;
  float a2 = a * a;
  float sp = a + a * a2 * (-0.16666667f + a2 * (0.0083333333f + a2 * (-0.00019841270f + a2 * 2.7557319e-6f)));
  float cp = 1.0f + a2 * (-0.5f + a2 * (0.041666668f + a2 * (-0.0013888889f + a2 * 2.4801587e-5f)));
  int qi = ((int)q) & 3;
  float s, c;
  if (qi == 0) { s = sp; c = cp; }
  else if (qi == 1) { s = cp; c = -sp; }
  else if (qi == 2) { s = -sp; c = -cp; }
  else { s = -cp; c = sp; }
  *s_out = s;
  *c_out = c;
}

/* asin on [0, 0.5] by its Taylor series (10 terms). */
static float asin_small(float z) {
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
/* acos (replaces libm acos at src/renderer/IBL.h:34-35). */
float orc_acos(float x) {
  if (!(x > -1.0f)) x = -1.0f;  /* also maps NaN to -1 */
  if (x > 1.0f) x = 1.0f;
  float ax = fabsf(x);
  float r;
  if (ax <= 0.5f) {
    r = 1.5707964f - asin_small(ax);
    if (x < 0.0f) r = ORC_PI - r;
  } else {
    float s = sqrtf((1.0f - ax) * 0.5f);
    r = 2.0f * asin_small(s);
    if (x < 0.0f) r = ORC_PI - r;
  }
  return r;
}

/* ------------------------------------------------------- intersections */
/* Polygon::CheckIntersection (src/renderer/Polygon.h:63-106): single-sided
 * Moller-Trumbore, det > EPS, 0 <= u <= det, v >= 0, u+v <= det, t >= EPS. */
int orc_intersect_triangle(const float ray[8], const float tri[9], float *t_out, float *u_out, float *v_out) {
  v3 o = ld3(ray), d = ld3(ray + 4);
  v3 p0 = ld3(tri), e1 = ld3(tri + 3), e2 = ld3(tri + 6);
  v3 P = cross(d, e2);
  float det = dot(P, e1);
  if (!(det > ORC_EPS)) return 0;
  v3 T = sub(o, p0);
  float u = dot(P, T);
  if (!(u >= 0.0f && u <= det)) return 0;
  v3 Q = cross(T, e1);
  float v = dot(Q, d);
  if (!(v >= 0.0f && u + v <= det)) return 0;
  float t = dot(Q, e2) / det;
  if (!(t >= ORC_EPS)) return 0;
  *t_out = t;
  *u_out = u / det;
  *v_out = v / det;
  return 1;
}

/* Sphere::CheckIntersection (src/renderer/Sphere.h:21-48).  The reference
 * forms D = (v.(c-x))^2 - |v|^2(|c-x|^2 - r^2) in FP64; this binary32
 * restatement uses the algebraically equal D = r^2 - |(c-x) - b v|^2 (|v| = 1,
 * which the reference's t = b -/+ sqrt(D) already assumes) to avoid the
 * cancellation FP64 did not suffer.  Roots and selection as the reference:
 * t2 = b - sqrt(D) if t2 > EPS else t1 = b + sqrt(D) if t1 > EPS. */
int orc_intersect_sphere(const float ray[8], const float sph[4], float *t_out) {
  v3 x = ld3(ray), v = ld3(ray + 4), c = ld3(sph);
  float r = sph[3];
  v3 oc = sub(c, x);
  float b = dot(v, oc);
  v3 perp = sub(oc, scl(v, b));
  float D = r * r - dot(perp, perp);
  if (D < 0.0f) return 0;
  float sD = sqrtf(D);
  float t1 = sD + b;
  float t2 = -sD + b;
  if (t1 <= ORC_EPS && t2 <= ORC_EPS) return 0;
  *t_out = (t2 > ORC_EPS) ? t2 : t1;
  return 1;
}

/* ------------------------------------------------------------- scene handle */
typedef struct {
  float lo[3], hi[3];
  int32_t left, right;    /* inner: children; leaf: left = -1 */
  uint32_t first, count;  /* leaf range in h->order */
} obvh_node;

struct orc_handle {
  orc_scene s;
  size_t n_prim;
  /* own binary BVH */
  obvh_node *nodes;
  size_t n_nodes;
  uint32_t *order;
  /* external wide BVH */
  const unsigned char *wnodes;
  size_t n_wnodes;
  uint32_t *worder;
  size_t n_worder;
  /* environment sampling tables */
  uint64_t *env_q;        /* env_h x env_w quantized weights */
  uint64_t *env_row;      /* env_h x (env_w+1) prefix per row */
  uint64_t *env_marg;     /* env_h+1 prefix over rows */
};

static void prim_box(const orc_handle *h, size_t i, float lo[3], float hi[3]) {
  if (i < h->s.n_tri) {
    const float *t = h->s.tri + 9 * i;
    for (int a = 0; a < 3; a++) {
      float p0 = t[a], p1 = t[a] + t[3 + a], p2 = t[a] + t[6 + a];
      float mn = p0 < p1 ? p0 : p1, mx = p0 > p1 ? p0 : p1;
      mn = mn < p2 ? mn : p2; mx = mx > p2 ? mx : p2;
      lo[a] = mn; hi[a] = mx;
    }
  } else {
    const float *s = h->s.sph + 4 * (i - h->s.n_tri);
    for (int a = 0; a < 3; a++) { lo[a] = s[a] - s[3]; hi[a] = s[a] + s[3]; }
  }
}

/* Own binary BVH: binned SAH (16 bins), leaves <= 4 primitives, boxes padded
 * so the slab test never culls a primitive the binary32 tests accept. */
typedef struct { float lo[3], hi[3], c[3]; } pbox;

static float area(const float lo[3], const float hi[3]) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
  return dx * dy + dy * dz + dz * dx;
}

static size_t build_rec(orc_handle *h, pbox *pb, uint32_t *idx, size_t first, size_t count, size_t *cap, float pad) {
  size_t me = h->n_nodes++;
  if (me >= *cap) { *cap *= 2; h->nodes = (obvh_node *)realloc(h->nodes, *cap * sizeof(obvh_node)); }
  obvh_node nd;
  float clo[3] = {ORC_INF, ORC_INF, ORC_INF}, chi[3] = {-ORC_INF, -ORC_INF, -ORC_INF};
  for (int a = 0; a < 3; a++) { nd.lo[a] = ORC_INF; nd.hi[a] = -ORC_INF; }
  for (size_t i = first; i < first + count; i++) {
    pbox *b = &pb[idx[i]];
    for (int a = 0; a < 3; a++) {
      if (b->lo[a] < nd.lo[a]) nd.lo[a] = b->lo[a];
      if (b->hi[a] > nd.hi[a]) nd.hi[a] = b->hi[a];
      if (b->c[a] < clo[a]) clo[a] = b->c[a];
      if (b->c[a] > chi[a]) chi[a] = b->c[a];
    }
  }
  for (int a = 0; a < 3; a++) { nd.lo[a] -= pad; nd.hi[a] += pad; }
  int best_axis = -1, best_bin = -1;
  float best_cost = (float)count * area(nd.lo, nd.hi);
  if (count > 4) {
    for (int a = 0; a < 3; a++) {
      float ext = chi[a] - clo[a];
      if (!(ext > 0.0f)) continue;
      enum { NB = 16 };
      float blo[NB][3], bhi[NB][3];
      size_t bn[NB];
      for (int k = 0; k < NB; k++) { bn[k] = 0; for (int q = 0; q < 3; q++) { blo[k][q] = ORC_INF; bhi[k][q] = -ORC_INF; } }
      for (size_t i = first; i < first + count; i++) {
        pbox *b = &pb[idx[i]];
        int k = (int)((b->c[a] - clo[a]) / ext * NB);
        if (k >= NB) k = NB - 1;
        if (k < 0) k = 0;
        bn[k]++;
        for (int q = 0; q < 3; q++) { if (b->lo[q] < blo[k][q]) blo[k][q] = b->lo[q]; if (b->hi[q] > bhi[k][q]) bhi[k][q] = b->hi[q]; }
      }
      for (int split = 1; split < NB; split++) {
        float llo[3] = {ORC_INF, ORC_INF, ORC_INF}, lhi[3] = {-ORC_INF, -ORC_INF, -ORC_INF};
        float rlo[3] = {ORC_INF, ORC_INF, ORC_INF}, rhi[3] = {-ORC_INF, -ORC_INF, -ORC_INF};
        size_t nl = 0, nr = 0;
        for (int k = 0; k < split; k++) { nl += bn[k]; for (int q = 0; q < 3; q++) { if (blo[k][q] < llo[q]) llo[q] = blo[k][q]; if (bhi[k][q] > lhi[q]) lhi[q] = bhi[k][q]; } }
        for (int k = split; k < NB; k++) { nr += bn[k]; for (int q = 0; q < 3; q++) { if (blo[k][q] < rlo[q]) rlo[q] = blo[k][q]; if (bhi[k][q] > rhi[q]) rhi[q] = bhi[k][q]; } }
        if (nl == 0 || nr == 0) continue;
        float cost = 0.125f * area(nd.lo, nd.hi) + (float)nl * area(llo, lhi) + (float)nr * area(rlo, rhi);
        if (cost < best_cost || count > 8) {
          if (best_axis < 0 || cost < best_cost) { best_cost = cost; best_axis = a; best_bin = split; }
        }
      }
    }
  }
  if (best_axis < 0 && count > 8) {
    /* all centroids coincide along every axis: split the range in half */
    size_t half = count / 2;
    nd.left = (int32_t)build_rec(h, pb, idx, first, half, cap, pad);
    nd.right = (int32_t)build_rec(h, pb, idx, first + half, count - half, cap, pad);
    nd.first = nd.count = 0;
    h->nodes[me] = nd;
    return me;
  }
  if (best_axis < 0) {
    nd.left = nd.right = -1;
    nd.first = (uint32_t)first;
    nd.count = (uint32_t)count;
    h->nodes[me] = nd;
    return me;
  }
  float ext = chi[best_axis] - clo[best_axis];
  size_t i = first, j = first + count;
  while (i < j) {
    pbox *b = &pb[idx[i]];
    int k = (int)((b->c[best_axis] - clo[best_axis]) / ext * 16);
    if (k >= 16) k = 15;
    if (k < 0) k = 0;
    if (k < best_bin) i++;
    else { j--; uint32_t tmp = idx[i]; idx[i] = idx[j]; idx[j] = tmp; }
  }
  size_t nl = i - first;
  if (nl == 0 || nl == count) nl = count / 2;
  nd.left = (int32_t)build_rec(h, pb, idx, first, nl, cap, pad);
  nd.right = (int32_t)build_rec(h, pb, idx, first + nl, count - nl, cap, pad);
  nd.first = nd.count = 0;
  h->nodes[me] = nd;
  return me;
}

static void env_tables(orc_handle *h);

orc_handle *orc_prepare(const orc_scene *scene) {
  orc_handle *h = (orc_handle *)calloc(1, sizeof(orc_handle));
  h->s = *scene;
  h->n_prim = scene->n_tri + scene->n_sph;
  if (h->n_prim > 0) {
    pbox *pb = (pbox *)malloc(h->n_prim * sizeof(pbox));
    float ext = 1.0f;
    for (size_t i = 0; i < h->n_prim; i++) {
      prim_box(h, i, pb[i].lo, pb[i].hi);
      for (int a = 0; a < 3; a++) {
        pb[i].c[a] = 0.5f * (pb[i].lo[a] + pb[i].hi[a]);
        float m = fabsf(pb[i].lo[a]) > fabsf(pb[i].hi[a]) ? fabsf(pb[i].lo[a]) : fabsf(pb[i].hi[a]);
        if (m > ext) ext = m;
      }
    }
    h->order = (uint32_t *)malloc(h->n_prim * sizeof(uint32_t));
    for (size_t i = 0; i < h->n_prim; i++) h->order[i] = (uint32_t)i;
    size_t cap = 2 * h->n_prim + 1;
    h->nodes = (obvh_node *)malloc(cap * sizeof(obvh_node));
    build_rec(h, pb, h->order, 0, h->n_prim, &cap, ext * 1.0e-5f);
    free(pb);
  }
  if (scene->env && scene->env_w > 0 && scene->env_h > 0) env_tables(h);
  return h;
}

void orc_free(orc_handle *h) {
  if (!h) return;
  free(h->nodes); free(h->order); free(h->worder);
  free(h->env_q); free(h->env_row); free(h->env_marg);
  free(h);
}

int orc_set_wide_bvh(orc_handle *h, const void *nodes, size_t n_nodes, const uint32_t *prim_order, size_t n_order) {
  h->wnodes = (const unsigned char *)nodes;
  h->n_wnodes = n_nodes;
  free(h->worder);
  h->worder = (uint32_t *)malloc((n_order ? n_order : 1) * sizeof(uint32_t));
  memcpy(h->worder, prim_order, n_order * sizeof(uint32_t));
  h->n_worder = n_order;
  return 0;
}

/* ------------------------------------------------------------ traversal */
typedef struct { float t, u, v; int32_t prim; } hitrec;

int orc_intersect_plane(const float ray[8], float y, float *t) {
  if (fabsf(ray[5]) < ORC_EPS) return 0;
  float tt = (y - ray[1]) / ray[5];
  if (tt < ORC_EPS) return 0;
  *t = tt;
  return 1;
}

/* closest-hit update with the tie rule: smaller t wins, equal t -> smaller id */
static inline void consider(const orc_handle *h, const float ray[8], uint32_t pid, hitrec *best, uint64_t *nprim) {
  float t, u = 0.0f, v = 0.0f;
  int ok;
  if (nprim) (*nprim)++;
  if (pid < h->s.n_tri) ok = orc_intersect_triangle(ray, h->s.tri + 9 * (size_t)pid, &t, &u, &v);
  else ok = orc_intersect_sphere(ray, h->s.sph + 4 * (size_t)(pid - h->s.n_tri), &t);
  if (!ok) return;
  if (t < best->t || (t == best->t && best->prim >= 0 && (int32_t)pid < best->prim)) {
    best->t = t; best->u = u; best->v = v; best->prim = (int32_t)pid;
  }
}

/* Slab test restating BoundingBox::CheckIntersection4floatAABB
 * (src/renderer/BoundingBox.h:127-172) per box, with the far distance scaled
 * by (1 + 2^-20) so rounding never rejects a box the ray touches. */
static inline int slab(const float o[3], const float inv[3], const int sgn[3], const float lo[3], const float hi[3],
                       float tbest, float *tnear) {
  float tmin = 0.0f, tmax = tbest;
  for (int a = 0; a < 3; a++) {
    float n = ((sgn[a] ? hi[a] : lo[a]) - o[a]) * inv[a];
    float f = ((sgn[a] ? lo[a] : hi[a]) - o[a]) * inv[a];
    tmin = n > tmin ? n : tmin;
    tmax = f < tmax ? f : tmax;
  }
  *tnear = tmin;
  return tmin <= tmax * 1.00000095f;
}

static void ray_setup(const float ray[8], float o[3], float inv[3], int sgn[3]) {
  for (int a = 0; a < 3; a++) {
    float d = ray[4 + a];
    o[a] = ray[a];
    /* QBVH.cpp:56-60: zero direction component -> INF */
    inv[a] = (d == 0.0f) ? ORC_INF : 1.0f / d;
    sgn[a] = (d >= 0.0f) ? 0 : 1;
  }
}

static void trace_own(const orc_handle *h, const float ray[8], hitrec *best, uint64_t *cnt) {
  if (h->n_nodes == 0) return;
  float o[3], inv[3]; int sgn[3];
  ray_setup(ray, o, inv, sgn);
  int32_t stack[256]; float stt[256];
  int sp = 0;
  float tn;
  if (!slab(o, inv, sgn, h->nodes[0].lo, h->nodes[0].hi, best->t, &tn)) return;
  stack[sp] = 0; stt[sp] = tn; sp++;
  while (sp > 0) {
    sp--;
    int32_t ni = stack[sp];
    if (stt[sp] > best->t) continue;
    const obvh_node *nd = &h->nodes[ni];
    if (cnt) cnt[0]++;
    if (nd->left < 0) {
      for (uint32_t k = 0; k < nd->count; k++) consider(h, ray, h->order[nd->first + k], best, cnt ? &cnt[1] : NULL);
      continue;
    }
    float t0, t1;
    int h0 = slab(o, inv, sgn, h->nodes[nd->left].lo, h->nodes[nd->left].hi, best->t, &t0);
    int h1 = slab(o, inv, sgn, h->nodes[nd->right].lo, h->nodes[nd->right].hi, best->t, &t1);
    if (h0 && h1) {
      if (t0 <= t1) { stack[sp] = nd->right; stt[sp] = t1; sp++; stack[sp] = nd->left; stt[sp] = t0; sp++; }
      else { stack[sp] = nd->left; stt[sp] = t0; sp++; stack[sp] = nd->right; stt[sp] = t1; sp++; }
    } else if (h0) { stack[sp] = nd->left; stt[sp] = t0; sp++; }
    else if (h1) { stack[sp] = nd->right; stt[sp] = t1; sp++; }
  }
}

/* External 4-wide node, 128 bytes (layout shared with the product BVH file format). */
typedef struct {
  float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
  int32_t child[4];
  int32_t meta[4];
} wnode;

#define W_IS_LEAF(c) (((uint32_t)(c) & 0xC0000000u) == 0x80000000u)
#define W_LEAF_FIRST(c) ((uint32_t)(c) & 0x07FFFFFFu)
#define W_LEAF_COUNT(c) ((((uint32_t)(c) >> 27) & 7u) + 1u)

static void trace_wide(const orc_handle *h, const float ray[8], hitrec *best, uint64_t *cnt) {
  if (h->n_wnodes == 0) return;
  const wnode *nodes = (const wnode *)h->wnodes;
  float o[3], inv[3]; int sgn[3];
  ray_setup(ray, o, inv, sgn);
  int32_t stack[512]; float stt[512];
  int sp = 0;
  stack[sp] = 0; stt[sp] = 0.0f; sp++;
  while (sp > 0) {
    sp--;
    int32_t c = stack[sp];
    if (stt[sp] > best->t) continue;
    if (W_IS_LEAF(c)) {
      uint32_t f = W_LEAF_FIRST(c), n = W_LEAF_COUNT(c);
      for (uint32_t k = 0; k < n; k++) consider(h, ray, h->worder[f + k], best, cnt ? &cnt[1] : NULL);
      continue;
    }
    const wnode *nd = &nodes[c];
    if (cnt) cnt[0]++;
    float tn[4]; int hit[4];
    for (int k = 0; k < 4; k++) {
      hit[k] = 0;
      if (nd->child[k] == -1) continue;
      float lo[3] = {nd->lox[k], nd->loy[k], nd->loz[k]}, hi[3] = {nd->hix[k], nd->hiy[k], nd->hiz[k]};
      hit[k] = slab(o, inv, sgn, lo, hi, best->t, &tn[k]);
    }
    /* push far-to-near so the nearest child is popped first */
    int ordk[4], m = 0;
    for (int k = 0; k < 4; k++) if (hit[k]) ordk[m++] = k;
    for (int a = 1; a < m; a++) {
      int x = ordk[a], b = a - 1;
      while (b >= 0 && tn[ordk[b]] > tn[x]) { ordk[b + 1] = ordk[b]; b--; }
      ordk[b + 1] = x;
    }
    for (int a = m - 1; a >= 0; a--) { stack[sp] = nd->child[ordk[a]]; stt[sp] = tn[ordk[a]]; sp++; }
  }
}

static void closest(const orc_handle *h, const float ray[8], int accel, hitrec *best, uint64_t *cnt) {
  best->t = ray[3];
  best->prim = -1;
  best->u = best->v = 0.0f;
  if (accel == ORC_BRUTE) {
    for (size_t i = 0; i < h->n_prim; i++) consider(h, ray, (uint32_t)i, best, cnt ? &cnt[1] : NULL);
  } else if (accel == ORC_OWN_BVH) {
    trace_own(h, ray, best, cnt);
  } else {
    trace_wide(h, ray, best, cnt);
  }
  /* objects outside the BVH after it (Scene.cpp:72-80): infinite floors */
  for (size_t k = 0; k < h->s.n_pln; k++) {
    float t;
    if (!orc_intersect_plane(ray, h->s.pln[k], &t)) continue;
    int32_t id = (int32_t)(h->n_prim + k);
    if (t < best->t || (t == best->t && best->prim >= 0 && id < best->prim)) {
      best->t = t; best->u = best->v = 0.0f; best->prim = id;
    }
  }
}

typedef struct {
  const orc_handle *h; const float *rays; size_t n; int accel;
  int32_t *prim; float *t; atomic_size_t next; atomic_ullong c0, c1;
} trace_job;

static void *trace_worker(void *arg) {
  trace_job *J = (trace_job *)arg;
  uint64_t cnt[2] = {0, 0};
  for (;;) {
    size_t b = atomic_fetch_add(&J->next, 1024);
    if (b >= J->n) break;
    size_t e = b + 1024 < J->n ? b + 1024 : J->n;
    for (size_t i = b; i < e; i++) {
      hitrec hr;
      closest(J->h, J->rays + 8 * i, J->accel, &hr, cnt);
      J->prim[i] = hr.prim;
      J->t[i] = hr.prim >= 0 ? hr.t : ORC_INF;
    }
  }
  atomic_fetch_add(&J->c0, cnt[0]);
  atomic_fetch_add(&J->c1, cnt[1]);
  return NULL;
}

void orc_trace(const orc_handle *h, const float *rays, size_t n, int accel, int32_t *prim, float *t,
               uint64_t *counters, int nthreads) {
  trace_job J;
  J.h = h; J.rays = rays; J.n = n; J.accel = accel; J.prim = prim; J.t = t;
  atomic_init(&J.next, 0); atomic_init(&J.c0, 0); atomic_init(&J.c1, 0);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, trace_worker, &J);
  trace_worker(&J);
  for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
  if (counters) { counters[0] = atomic_load(&J.c0); counters[1] = atomic_load(&J.c1); }
}

/* any hit in [EPS, tmax): used by environment shadow rays */
static int occluded(const orc_handle *h, const float ray[8], int accel) {
  hitrec hr;
  closest(h, ray, accel, &hr, NULL);
  return hr.prim >= 0;
}

/* ------------------------------------------------------------ environment */
/* IBL::Sample(dir) (src/renderer/IBL.h:30-45) + PhiThetaToUV (:67-71):
 * theta = acos(y); phi = acos(x / sqrt(x^2+z^2)), phi = 2pi - phi if z < 0;
 * texel (int(u W) % W, int(v H) % H) with u = phi/2pi, v = theta/pi. */
void orc_env_texel(float dx, float dy, float dz, uint32_t w, uint32_t h, uint32_t *ix, uint32_t *iy) {
  float theta = orc_acos(dy);
  float rxz = sqrtf(dx * dx + dz * dz);
  float phi = (rxz > 0.0f) ? orc_acos(dx / rxz) : 0.0f;
  if (dz < 0.0f) phi = ORC_TWO_PI - phi;
  float u = phi / ORC_TWO_PI;
  float v = theta / ORC_PI;
  *ix = (uint32_t)((int)(u * (float)w) % (int)w);
  *iy = (uint32_t)((int)(v * (float)h) % (int)h);
}

/* Sphere::CheckIntersection's texture uv (src/renderer/Sphere.h:52-75):
 * u = (atan2(n.z, n.x) + pi) / 2pi with atan2 = +-acos(n.x / |n.xz|) by the
 * sign of n.z; v = 1 - (asin(n.y) + pi/2) / pi = acos(n.y) / pi. */
void orc_sphere_uv(float nx, float ny, float nz, float *u, float *v) {
  float rxz = sqrtf(nx * nx + nz * nz);
  float phi = (rxz > 0.0f) ? orc_acos(nx / rxz) : 0.0f;
  *u = (nz >= 0.0f ? ORC_PI + phi : ORC_PI - phi) / ORC_TWO_PI;
  *v = orc_acos(ny) / ORC_PI;
}

/* Image::GetPixelByUV (src/tools/Image.h:21-30): Utils::Clamp to [0, 1]
 * (NaN to 0 here), times 0.99999, texel (int(u w), int(v h)). */
void orc_texel_xy(float u, float v, uint32_t w, uint32_t h, uint32_t *x, uint32_t *y) {
  u = u > 0.0f ? (u < 1.0f ? u : 1.0f) : 0.0f;
  v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;
  uint32_t ix = (uint32_t)(int)((u * 0.99999f) * (float)w);
  uint32_t iy = (uint32_t)(int)((v * 0.99999f) * (float)h);
  *x = ix < w ? ix : w - 1;
  *y = iy < h ? iy : h - 1;
}

static inline v3 env_rgb(const orc_handle *h, v3 d) {
  uint32_t ix, iy;
  orc_env_texel(d.x, d.y, d.z, h->s.env_w, h->s.env_h, &ix, &iy);
  const float *p = h->s.env + 3 * ((size_t)iy * h->s.env_w + ix);
  return mk(p[0], p[1], p[2]);
}

/* Importance-sampling tables (no reference counterpart: the reference's
 * CreateImportanceSamplingMap returns immediately, src/renderer/IBL.h:104-105).
 * Weight per texel = luminance * sin(theta at the row centre), quantized to
 * 32-bit fixed point relative to the maximum so that prefix sums are exact
 * integers whatever order a parallel scan adds them in. */
static void env_tables(orc_handle *h) {
  uint32_t W = h->s.env_w, H = h->s.env_h;
  float *wf = (float *)malloc((size_t)W * H * sizeof(float));
  float wmax = 0.0f;
  for (uint32_t iy = 0; iy < H; iy++) {
    float s, c;
    orc_sincos_turns(((float)iy + 0.5f) / (2.0f * (float)H), &s, &c);
    for (uint32_t ix = 0; ix < W; ix++) {
      const float *p = h->s.env + 3 * ((size_t)iy * W + ix);
      float lum = (0.2126f * p[0] + 0.7152f * p[1]) + 0.0722f * p[2];
      float w = lum * s;
      if (!(w > 0.0f)) w = 0.0f;
      wf[(size_t)iy * W + ix] = w;
      if (w > wmax) wmax = w;
    }
  }
  h->env_q = (uint64_t *)malloc((size_t)W * H * sizeof(uint64_t));
  h->env_row = (uint64_t *)malloc((size_t)H * (W + 1) * sizeof(uint64_t));
  h->env_marg = (uint64_t *)malloc((size_t)(H + 1) * sizeof(uint64_t));
  double scale = wmax > 0.0f ? 4294967295.0 / (double)wmax : 0.0;
  uint64_t acc = 0;
  h->env_marg[0] = 0;
  for (uint32_t iy = 0; iy < H; iy++) {
    uint64_t r = 0;
    uint64_t *row = h->env_row + (size_t)iy * (W + 1);
    row[0] = 0;
    for (uint32_t ix = 0; ix < W; ix++) {
      uint64_t q = (uint64_t)((double)wf[(size_t)iy * W + ix] * scale);
      h->env_q[(size_t)iy * W + ix] = q;
      r += q;
      row[ix + 1] = r;
    }
    acc += r;
    h->env_marg[iy + 1] = acc;
  }
  free(wf);
}

int orc_env_tables_raw(const orc_handle *h, uint64_t *row, uint64_t *marg) {
  if (!h->env_q) return -1;
  memcpy(row, h->env_row, (size_t)h->s.env_h * (h->s.env_w + 1) * sizeof(uint64_t));
  memcpy(marg, h->env_marg, (size_t)(h->s.env_h + 1) * sizeof(uint64_t));
  return 0;
}

int orc_env_cdf(const orc_handle *h, float *marginal, float *conditional) {
  if (!h->env_q) return -1;
  uint32_t W = h->s.env_w, H = h->s.env_h;
  double tot = (double)h->env_marg[H];
  for (uint32_t i = 0; i <= H; i++) marginal[i] = tot > 0 ? (float)((double)h->env_marg[i] / tot) : 0.0f;
  for (uint32_t iy = 0; iy < H; iy++) {
    const uint64_t *row = h->env_row + (size_t)iy * (W + 1);
    double rt = (double)row[W];
    for (uint32_t ix = 0; ix <= W; ix++) conditional[(size_t)iy * (W + 1) + ix] = rt > 0 ? (float)((double)row[ix] / rt) : 0.0f;
  }
  return 0;
}

/* W*H / (2 pi^2 total) in binary64: the texel weight's factor in the pdf */
static double env_pdf_factor(uint32_t W, uint32_t H, uint64_t tot) {
  return ((double)W * (double)H) / (19.739208802178716 * (double)tot);
}

/* solid-angle pdf of the sampler at direction d inside texel i: the texel's
 * probability times W*H / (2 pi^2 sin(theta)), sin(theta) = |d.xz|, as
 * q * (W*H / (2 pi^2 total)) in binary64, rounded to binary32, divided by
 * sin(theta) in binary32. */
static float env_pdf_texel(const orc_handle *h, size_t i, float dx, float dz) {
  uint32_t W = h->s.env_w, H = h->s.env_h;
  uint64_t tot = h->env_marg[H];
  if (tot == 0) return 0.0f;
  uint64_t q = h->env_q[i];
  float st = sqrtf(dx * dx + dz * dz);
  if (q == 0 || !(st > 0.0f)) return 0.0f;
  return (float)((double)q * env_pdf_factor(W, H, tot)) / st;
}

/* the same for any direction: its texel by the latlong mapping */
float orc_env_pdf(const orc_handle *h, float dx, float dy, float dz) {
  if (!h->env_q) return 0.0f;
  uint32_t W = h->s.env_w, H = h->s.env_h;
  uint64_t tot = h->env_marg[H];
  if (tot == 0) return 0.0f;
  uint32_t ix, iy;
  orc_env_texel(dx, dy, dz, W, H, &ix, &iy);
  uint64_t q = h->env_q[(size_t)iy * W + ix];
  float st = sqrtf(dx * dx + dz * dz);
  if (q == 0 || !(st > 0.0f)) return 0.0f;
  return (float)((double)q * env_pdf_factor(W, H, tot)) / st;
}

/* largest i in [0, n) with a[i] <= key (a nondecreasing, a[0] <= key) */
static uint32_t search_le(const uint64_t *a, uint32_t n, uint64_t key) {
  uint32_t lo = 0, hi = n;  /* answer in [lo, hi) */
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] <= key) lo = mid; else hi = mid;
  }
  return lo;
}

/* Draw one direction from the tables with four uniforms in [0,1]: a texel
 * by the marginal and row CDFs (its index into *texel), then a point of it
 * uniform in (phi, theta). */
static int env_sample4(const orc_handle *h, float u1, float u2, float u3, float u4, v3 *dir, size_t *texel) {
  uint32_t W = h->s.env_w, H = h->s.env_h;
  uint64_t tot = h->env_marg[H];
  if (tot == 0) return 0;
  uint64_t t = (uint64_t)((double)u1 * (double)tot);
  if (t >= tot) t = tot - 1;
  uint32_t iy = search_le(h->env_marg, H, t);
  const uint64_t *row = h->env_row + (size_t)iy * (W + 1);
  uint64_t rt = row[W];
  uint64_t t2 = (uint64_t)((double)u2 * (double)rt);
  if (t2 >= rt) t2 = rt - 1;
  uint32_t ix = search_le(row, W, t2);
  float fx = ((float)ix + u3) / (float)W;
  float fy = ((float)iy + u4) / (2.0f * (float)H);
  if (fx >= 1.0f) fx = 0.99999994f;
  if (fy >= 0.5f) fy = 0.49999997f;
  float sp, cp, sth, cth;
  orc_sincos_turns(fx, &sp, &cp);
  orc_sincos_turns(fy, &sth, &cth);
  *dir = mk(sth * cp, cth, sth * sp);
  *texel = (size_t)iy * W + ix;
  return 1;
}

void orc_env_sample(const orc_handle *h, float u1, float u2, float dir[3], float *pdf, uint32_t *ix, uint32_t *iy) {
  /* u3 = u4 = 0.5: the texel centre */
  v3 d;
  size_t ti;
  if (!h->env_q || !env_sample4(h, u1, u2, 0.5f, 0.5f, &d, &ti)) { dir[0] = dir[1] = dir[2] = 0; *pdf = 0; return; }
  dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
  *pdf = orc_env_pdf(h, d.x, d.y, d.z);
  orc_env_texel(d.x, d.y, d.z, h->s.env_w, h->s.env_h, ix, iy);
}

/* ------------------------------------------------------------ camera */
/* Camera::SampleRayForPixel (src/renderer/Camera.h:65-93) and
 * CircleAperture::SampleOnePoint (src/renderer/Aperture.h:23-35). */
void orc_camera_ray(const orc_camera *cam, float x, float y, const float lens[2], float o[3], float d[3]) {
  v3 pos = ld3(cam->position), C = ld3(cam->screen_center), X = ld3(cam->screen_x), Y = ld3(cam->screen_y);
  float a = x / (float)cam->width - 0.5f;
  float b = y / (float)cam->height - 0.5f;
  v3 target = add(add(C, scl(X, a)), scl(Y, b));
  v3 dir = normalize(sub(target, pos));
  v3 org = pos;
  if (cam->aperture_radius > 0.0f && lens) {
    float st, ct;
    orc_sincos_turns(lens[0], &st, &ct);
    float r = cam->aperture_radius * sqrtf(lens[1]);
    float dx = r * ct, dy = r * st;
    float ft = cam->focal_distance / fabsf(dir.z);
    v3 focus = add(pos, scl(dir, ft));
    org = add(pos, mk(dx, dy, 0.0f));
    dir = normalize(sub(focus, org));
  }
  o[0] = org.x; o[1] = org.y; o[2] = org.z;
  d[0] = dir.x; d[1] = dir.y; d[2] = dir.z;
}

/* ------------------------------------------------------------ integrator */
typedef struct {
  v3 o, d, beta;
  uint32_t depth, branch;
} pending;

static inline int64_t fix(float v) {
  if (!(v > 0.0f)) return 0;
  if (v > ORC_FIX_CLAMP) v = ORC_FIX_CLAMP;
  return (int64_t)llrintf(v * ORC_FIX_SCALE);
}

typedef struct {
  const orc_handle *h; const orc_camera *cam; const orc_params *p; int accel;
  uint32_t row_begin, row_end;
  int64_t *sums;
  atomic_uint next_row;
  atomic_ullong segments, shadow;
} render_job;

static inline void add_contrib(int64_t *acc, v3 beta, v3 L, float s) {
  acc[0] += fix(beta.x * L.x * s);
  acc[1] += fix(beta.y * L.y * s);
  acc[2] += fix(beta.z * L.z * s);
}

/* One sample: the recursion of PathTracer::Radiance / Radiance_internal /
 * Radiance_{Lambert,Specular,Refraction} (src/renderer/PathTracer.cpp:180-591)
 * unrolled into a loop over path segments with a pending-branch stack for the
 * depth <= 2 glass split.  Throughput is carried forward: L += beta * Le. */
static void trace_sample(const render_job *J, uint32_t pixel, uint32_t subpx, uint32_t sample, v3 o, v3 d,
                         int64_t *acc, uint64_t *nseg, uint64_t *nshadow) {
  const orc_handle *h = J->h;
  const orc_params *p = J->p;
  const orc_scene *S = &h->s;
  int env_on = S->env != NULL && S->env_w > 0;
  int nee = env_on && p->env_is && h->env_q && h->env_marg[S->env_h] > 0;
  pending stk[8];
  int sp = 0;
  v3 beta = mk(1.0f, 1.0f, 1.0f);
  uint32_t depth = 0, branch = 0;
  uint32_t rng[4];
  orc_rng_init(orc_path_seed(p->seed, pixel, subpx, sample, 0), rng);
  int prev_mis = 0;
  float prev_pdf = 0.0f;
  for (;;) {
    float ray[8] = {o.x, o.y, o.z, ORC_INF, d.x, d.y, d.z, 0.0f};
    hitrec hr;
    (*nseg)++;
    closest(h, ray, J->accel, &hr, NULL);
    int cont = 0;
    if (hr.prim < 0) {
      /* PathTracer.cpp:357-363: miss -> IBL lookup or black background */
      if (env_on) {
        v3 L = env_rgb(h, d);
        float w = 1.0f;
        if (prev_mis) {
          float pe = orc_env_pdf(h, d.x, d.y, d.z);
          w = (prev_pdf * prev_pdf) / (prev_pdf * prev_pdf + pe * pe);
        }
        add_contrib(acc, beta, L, w);
      }
    } else {
      uint32_t pid = (uint32_t)hr.prim;
      v3 x = add(o, scl(d, hr.t));
      v3 n;
      if (pid < S->n_tri) {
        /* Polygon.h:178-179: interpolated normal, normalized */
        const float *tn = S->tri_n + 9 * (size_t)pid;
        n = normalize(add(add(scl(ld3(tn + 3), hr.u), scl(ld3(tn + 6), hr.v)), ld3(tn)));
      } else if (pid < S->n_tri + S->n_sph) {
        /* Sphere.h:47-48 */
        n = normalize(sub(x, ld3(S->sph + 4 * (size_t)(pid - S->n_tri))));
      } else {
        /* InfiniteFloor.h:34-36: +-y, facing the ray */
        n = mk(0.0f, d.y < 0.0f ? 1.0f : -1.0f, 0.0f);
      }
      /* PathTracer.cpp:189: orient against the ray */
      v3 nl = dot(n, d) < 0.0f ? n : scl(n, -1.0f);
      /* PathTracer.cpp:368-386: material roulette, one draw on every hit */
      uint32_t mfirst = S->prim_mat[pid];
      uint32_t mcount = S->prim_mat_count ? S->prim_mat_count[pid] : 1u;
      float value = orc_rng_float(rng);
      uint32_t sel = mcount - 1;
      for (uint32_t i = 0; i < mcount; i++) {
        float r = S->mat[mfirst + i].rate;
        if (value <= r) { sel = i; break; }
        value -= r;
      }
      const orc_material *m = &S->mat[mfirst + sel];
      /* PathTracer.cpp:327-346,388-389: hit colour = colour x texel */
      v3 col0 = ld3(m->color), em = ld3(m->emission);
      v3 col = col0;
      if (m->tex) {
        float tu = 0.0f, tv = 0.0f;
        if (pid < S->n_tri) {
          if (S->tri_uv) {
            const float *q = S->tri_uv + 6 * (size_t)pid;
            tu = (q[2] * hr.u + q[4] * hr.v) + q[0];
            tv = (q[3] * hr.u + q[5] * hr.v) + q[1];
          }
        } else if (pid < S->n_tri + S->n_sph) {
          orc_sphere_uv(n.x, n.y, n.z, &tu, &tv);
        }
        const orc_texture *tx = &S->tex[m->tex - 1];
        uint32_t ix, iy;
        orc_texel_xy(tu, tv, tx->w, tx->h, &ix, &iy);
        const float *px = tx->rgb + 3 * ((size_t)iy * tx->w + ix);
        col = mk(col0.x * px[0], col0.y * px[1], col0.z * px[2]);
      }
      /* PathTracer.cpp:398-416: Russian roulette on max(albedo) */
      float prr = col.x > (col.y > col.z ? col.y : col.z) ? col.x : (col.y > col.z ? col.y : col.z);
      if (p->max_depth == 0 && depth > 64) prr *= ldexpf(1.0f, -(int)(depth - 64));
      int killed = 0;
      if (depth > p->rr_depth) {
        if (orc_rng_float(rng) >= prr) killed = 1;
      } else {
        prr = 1.0f;
      }
      /* PathTracer.cpp:433-439 (and :406-410 on a roulette kill): += emission,
       * with next-event estimation only where the eye ray itself hits
       * (depth 0); its direct term is always 0 (:312-314 multiply by the
       * non-emissive surface's own emission) */
      if (!p->nee || depth == 0) add_contrib(acc, beta, em, 1.0f);
      /* unlimited depth (reference mode) is capped at 255 segments */
      int can_continue = !killed && depth + 1 < (p->max_depth == 0 ? 255u : p->max_depth);
      prev_mis = 0;
      if (!killed) {
        if (m->type == 0) {
          /* Radiance_Lambert (PathTracer.cpp:445-500): unnormalized basis */
          if (dot(em, em) == 0.0f) {
            v3 w = nl;
            v3 u = fabsf(nl.x) > ORC_EPS ? cross(mk(0.0f, 1.0f, 0.0f), w) : cross(mk(1.0f, 0.0f, 0.0f), w);
            v3 vv = cross(w, u);
            float u1 = orc_rng_float(rng), u2 = orc_rng_float(rng);
            float s1, c1;
            orc_sincos_turns(u1, &s1, &c1);
            float r2 = sqrtf(u2), r3 = sqrtf(1.0f - u2);
            v3 nd = normalize(add(add(scl(scl(u, r3), c1), scl(scl(vv, r3), s1)), scl(w, r2)));
            float k = 1.0f / prr;
            if (nee && can_continue) {
              float e1 = orc_rng_float(rng), e2 = orc_rng_float(rng), e3 = orc_rng_float(rng), e4 = orc_rng_float(rng);
              v3 ld;
              size_t ti;
              if (env_sample4(h, e1, e2, e3, e4, &ld, &ti)) {
                /* the sampled texel's pdf and radiance */
                float c = dot(nl, ld);
                float pe = env_pdf_texel(h, ti, ld.x, ld.z);
                if (c > 0.0f && pe > 0.0f) {
                  float pb = c * ORC_INV_PI;
                  float w = (pe * pe) / (pe * pe + pb * pb);
                  float s = ((pb / pe) * w) * k;
                  const float *lp = h->s.env + 3 * ti;
                  v3 L = mk(lp[0], lp[1], lp[2]);
                  v3 bc = mk(beta.x * col.x, beta.y * col.y, beta.z * col.z);
                  float sray[8] = {x.x, x.y, x.z, ORC_INF, ld.x, ld.y, ld.z, 0.0f};
                  (*nshadow)++;
                  if (!occluded(h, sray, J->accel)) add_contrib(acc, bc, L, s);
                }
              }
            }
            if (can_continue) {
              beta = mk(beta.x * (col.x * k), beta.y * (col.y * k), beta.z * (col.z * k));
              if (nee) { prev_mis = 1; prev_pdf = dot(nl, nd) * ORC_INV_PI; }
              o = x; d = nd; depth++; cont = 1;
            }
          }
        } else if (m->type == 1) {
          /* Radiance_Specular (PathTracer.cpp:503-520) */
          v3 r = normalize(sub(d, scl(scl(nl, 2.0f), dot(d, nl))));
          if (can_continue) {
            float k = 1.0f / prr;
            beta = mk(beta.x * (col.x * k), beta.y * (col.y * k), beta.z * (col.z * k));
            o = x; d = r; depth++; cont = 1;
          }
        } else {
          /* Radiance_Refraction (PathTracer.cpp:523-591) */
          int into = dot(n, nl) > 0.0f;
          v3 refl = normalize(sub(d, scl(scl(nl, 2.0f), dot(d, nl))));
          float nr = into ? 1.0f / m->ior : m->ior / 1.0f;
          float dd = dot(d, nl);
          float cos2t = 1.0f - nr * nr * (1.0f - dd * dd);
          if (cos2t < 0.0f) {
            /* total internal reflection weighs by the untextured colour (PathTracer.cpp:543-548) */
            if (can_continue) {
              float k = 1.0f / prr;
              beta = mk(beta.x * (col0.x * k), beta.y * (col0.y * k), beta.z * (col0.z * k));
              o = x; d = refl; depth++; cont = 1;
            }
          } else {
            v3 refr = normalize(sub(scl(d, nr), scl(scl(n, into ? 1.0f : -1.0f), dd * nr + sqrtf(cos2t))));
            float F0 = ((m->ior - 1.0f) * (m->ior - 1.0f)) / ((m->ior + 1.0f) * (m->ior + 1.0f));
            float c = 1.0f - (into ? -dd : -dot(refr, nl));
            float c5 = c * c * c * c * c;
            float Fr = F0 + (1.0f - F0) * c5;
            float Tr = (1.0f - Fr) * (nr * nr);
            if (depth > 2) {
              float prob = 0.1f + 0.8f * Fr;
              if (orc_rng_float(rng) < prob) {
                if (can_continue) {
                  float k = Fr / (prr * prob);
                  beta = mk(beta.x * (col.x * k), beta.y * (col.y * k), beta.z * (col.z * k));
                  o = x; d = refl; depth++; cont = 1;
                }
              } else {
                if (can_continue) {
                  float k = Tr / (prr * (1.0f - prob));
                  beta = mk(beta.x * (col.x * k), beta.y * (col.y * k), beta.z * (col.z * k));
                  o = x; d = refr; depth++; cont = 1;
                }
              }
            } else if (can_continue) {
              /* split: refracted branch pending, reflected branch continues */
              float kt = Tr / prr;
              pending *pe = &stk[sp++];
              pe->o = x; pe->d = refr;
              pe->beta = mk(beta.x * (col.x * kt), beta.y * (col.y * kt), beta.z * (col.z * kt));
              pe->depth = depth + 1;
              pe->branch = branch | (1u << depth);
              float kr = Fr / prr;
              beta = mk(beta.x * (col.x * kr), beta.y * (col.y * kr), beta.z * (col.z * kr));
              o = x; d = refl; depth++; cont = 1;
            }
          }
        }
      }
    }
    if (cont) continue;
    if (sp == 0) break;
    sp--;
    o = stk[sp].o; d = stk[sp].d; beta = stk[sp].beta; depth = stk[sp].depth; branch = stk[sp].branch;
    orc_rng_init(orc_path_seed(p->seed, pixel, subpx, sample, branch), rng);
    prev_mis = 0;
  }
}

static void *render_worker(void *arg) {
  render_job *J = (render_job *)arg;
  const orc_camera *cam = J->cam;
  const orc_params *p = J->p;
  uint32_t W = cam->width, H = cam->height, ss = p->supersamples;
  uint64_t nseg = 0, nsh = 0;
  for (;;) {
    uint32_t r = J->row_begin + atomic_fetch_add(&J->next_row, 1);
    if (r >= J->row_end) break;
    uint32_t y = H - 1 - r;  /* ScreenPixels.cpp:64: stored rows are flipped */
    for (uint32_t x = 0; x < W; x++) {
      uint32_t pixel = r * W + x;
      int64_t acc[3] = {0, 0, 0};
      for (uint32_t sub = 0; sub < ss * ss; sub++) {
        uint32_t sx = sub % ss, sy = sub / ss;
        /* PathTracer.cpp:158-161: sub-pixel centre (2s+1)/(2ss) */
        float rx = (2.0f * (float)sx + 1.0f) / (2.0f * (float)ss);
        float ry = (2.0f * (float)sy + 1.0f) / (2.0f * (float)ss);
        float lens[2] = {0.0f, 0.0f};
        if (cam->aperture_radius > 0.0f) {
          uint32_t lr[4];
          orc_rng_init(orc_path_seed(p->seed, pixel, sub, p->spp_begin, 0xFFFFFFFFu), lr);
          lens[0] = orc_rng_float(lr);
          lens[1] = orc_rng_float(lr);
        }
        float o[3], d[3];
        orc_camera_ray(cam, (float)x + rx, (float)y + ry, lens, o, d);
        for (uint32_t s = p->spp_begin + 1; s <= p->spp_end; s++)
          trace_sample(J, pixel, sub, s, mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), acc, &nseg, &nsh);
      }
      int64_t *dst = J->sums + 3 * ((size_t)(r - J->row_begin) * W + x);
      dst[0] = acc[0]; dst[1] = acc[1]; dst[2] = acc[2];
    }
  }
  atomic_fetch_add(&J->segments, nseg);
  atomic_fetch_add(&J->shadow, nsh);
  return NULL;
}

int orc_render_sums(const orc_handle *h, const orc_camera *cam, const orc_params *p, int accel, int nthreads,
                    uint32_t row_begin, uint32_t row_end, int64_t *out_sums, uint64_t *segments, uint64_t *shadow) {
  if (row_end > cam->height || row_begin > row_end || p->supersamples == 0) return -1;
  if (accel == ORC_WIDE_BVH && h->n_wnodes == 0) return -2;
  render_job J;
  J.h = h; J.cam = cam; J.p = p; J.accel = accel; J.row_begin = row_begin; J.row_end = row_end;
  J.sums = out_sums;
  atomic_init(&J.next_row, 0); atomic_init(&J.segments, 0); atomic_init(&J.shadow, 0);
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, render_worker, &J);
  render_worker(&J);
  for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
  if (segments) *segments = atomic_load(&J.segments);
  if (shadow) *shadow = atomic_load(&J.shadow);
  return 0;
}

int orc_render(const orc_handle *h, const orc_camera *cam, const orc_params *p, int accel, int nthreads,
               uint32_t row_begin, uint32_t row_end, float *out_rgb, uint64_t *segments, uint64_t *shadow) {
  size_t npx = (size_t)(row_end - row_begin) * cam->width;
  int64_t *sums = (int64_t *)malloc((npx ? npx : 1) * 3 * sizeof(int64_t));
  int rc = orc_render_sums(h, cam, p, accel, nthreads, row_begin, row_end, sums, segments, shadow);
  if (rc == 0) {
    double n = (double)p->supersamples * p->supersamples * (double)(p->spp_end - p->spp_begin);
    for (size_t i = 0; i < 3 * npx; i++) out_rgb[i] = (float)((double)sums[i] * (1.0 / 16777216.0) / n);
  }
  free(sums);
  return rc;
}
