/*
 * omr_oracle.h -- CPU restatement of omochi-renderer3's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * Arithmetic is IEEE single precision with a fixed operation order and no
 * contraction (built with -ffp-contract=off), so that the same inputs give the
 * same bits on the GPU.  Each function cites the reference file:line whose
 * behaviour it restates (paths relative to the reference checkout).
 *
 * Pinning status of each function is listed at the top of omr_oracle.c.
 */
#ifndef OMR_ORACLE_H
#define OMR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material record, 48 bytes.  Mirrors Material (src/renderer/Material.h:11-36)
 * plus the multi-material rate (src/renderer/SceneObject.h:17-30). */
typedef struct {
  uint32_t type;        /* 0 lambert, 1 specular, 2 refraction */
  float color[3];
  float emission[3];
  float ior;
  float rate;           /* normalized roulette rate */
  uint32_t tex;         /* 0 = none, k = texture k-1 (Material::texture_id) */
  uint32_t pad[2];
} orc_material;

/* Texture (Image, src/tools/Image.h:4-35): linear RGB rows from the top. */
typedef struct {
  uint32_t w, h;
  const float *rgb;     /* h x w x 3 */
} orc_texture;

typedef struct {
  size_t n_tri;
  const float *tri;      /* n_tri x 9: v0.xyz, e1.xyz, e2.xyz  (Polygon.h:103-105) */
  const float *tri_n;    /* n_tri x 9: n0.xyz, n1-n0, n2-n0     (Polygon.h:109-111) */
  size_t n_sph;
  const float *sph;      /* n_sph x 4: center.xyz, radius        (Sphere.h:13-19)  */
  const uint32_t *prim_mat;       /* (n_tri+n_sph+n_pln): first material index */
  const uint32_t *prim_mat_count; /* (n_tri+n_sph+n_pln): material count, or NULL = 1 */
  size_t n_mat;
  const orc_material *mat;
  uint32_t env_w, env_h;
  const float *env;      /* env_h rows x env_w x 3 (row 0 = +y), or NULL */
  size_t n_pln;
  const float *pln;      /* n_pln planes y = pln[i] (InfiniteFloor.h:12-41), tested after the BVH */
  size_t n_tex;
  const orc_texture *tex;
  const float *tri_uv;   /* n_tri x 6: uv0, uv1-uv0, uv2-uv0 (Polygon.h:22-24), or NULL = uv 0 */
} orc_scene;

typedef struct {
  uint32_t width, height;
  float position[3];
  float screen_center[3];
  float screen_x[3];
  float screen_y[3];
  float aperture_radius;
  float focal_distance;
} orc_camera;

typedef struct {
  uint32_t supersamples;
  uint32_t spp_begin;
  uint32_t spp_end;
  uint32_t max_depth;   /* 0 = unlimited */
  uint32_t rr_depth;
  uint32_t env_is;
  uint64_t seed;
  uint32_t nee;         /* the reference's next-event estimation mode (emission at depth 0 only) */
  uint32_t pad;
} orc_params;

typedef struct orc_handle orc_handle;

/* Acceleration modes for trace/render. */
enum { ORC_BRUTE = 0, ORC_OWN_BVH = 1, ORC_WIDE_BVH = 2 };

/* ---- RNG (src/tools/Random.h:9-29) ---- */
void orc_rng_init(uint32_t seed, uint32_t state[4]);
uint32_t orc_rng_next(uint32_t state[4]);
float orc_rng_float(uint32_t state[4]);
uint32_t orc_path_seed(uint64_t seed, uint32_t pixel, uint32_t sub, uint32_t sample, uint32_t branch);

/* ---- elementary math restated with fixed operation order ---- */
void orc_sincos_turns(float x, float *s, float *c);
float orc_acos(float x);
int orc_intersect_triangle(const float ray[8], const float tri[9], float *t, float *u, float *v);
int orc_intersect_sphere(const float ray[8], const float sph[4], float *t);
/* InfiniteFloor::CheckIntersection (src/renderer/InfiniteFloor.h:20-38):
 * |d.y| >= EPS, t = (y - o.y) / d.y >= EPS. */
int orc_intersect_plane(const float ray[8], float y, float *t);
void orc_env_texel(float dx, float dy, float dz, uint32_t w, uint32_t h, uint32_t *ix, uint32_t *iy);
/* Latlong uv of a sphere normal (Sphere.h:52-75): u = (atan2(n.z, n.x) + pi) / 2pi,
 * v = 1 - (asin(n.y) + pi/2) / pi. */
void orc_sphere_uv(float nx, float ny, float nz, float *u, float *v);
/* Image::GetPixelByUV (src/tools/Image.h:21-30) texel coordinates. */
void orc_texel_xy(float u, float v, uint32_t w, uint32_t h, uint32_t *x, uint32_t *y);
void orc_camera_ray(const orc_camera *cam, float x, float y, const float lens[2], float o[3], float d[3]);

/* ---- scene handle ---- */
orc_handle *orc_prepare(const orc_scene *scene);
void orc_free(orc_handle *h);
/* Attach an externally built 4-wide BVH: nodes are 128-byte records
 * {float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4]; int32 child[4]; int32 meta[4]};
 * prim_order maps a leaf slot to the original primitive index. */
int orc_set_wide_bvh(orc_handle *h, const void *nodes, size_t n_nodes, const uint32_t *prim_order, size_t n_order);

/* Closest hit for n rays (n x 8 floats: o.xyz, tmax, d.xyz, pad).  prim = -1 on miss.
 * counters (optional, 2 x uint64): inner nodes visited, primitives tested. */
void orc_trace(const orc_handle *h, const float *rays, size_t n, int accel,
               int32_t *prim, float *t, uint64_t *counters, int nthreads);

/* Render rows [row_begin, row_end) of the image (row index is top-down output
 * row); out_rgb receives (row_end-row_begin) x width x 3 floats of per-pixel
 * mean over samples (spp_begin, spp_end] x ss^2.  segments (optional) receives
 * the number of closest-hit queries; shadow (optional) the shadow rays. */
int orc_render(const orc_handle *h, const orc_camera *cam, const orc_params *p, int accel,
               int nthreads, uint32_t row_begin, uint32_t row_end, float *out_rgb,
               uint64_t *segments, uint64_t *shadow);

/* Same as orc_render but returns the raw fixed-point sums (x 2^-24) per pixel
 * channel, for bit-level comparisons. */
int orc_render_sums(const orc_handle *h, const orc_camera *cam, const orc_params *p, int accel,
                    int nthreads, uint32_t row_begin, uint32_t row_end, int64_t *out_sums,
                    uint64_t *segments, uint64_t *shadow);

/* Environment importance sampling tables (built on the CPU here). */
int orc_env_cdf(const orc_handle *h, float *marginal /* env_h+1 */, float *conditional /* env_h x (env_w+1) */);
void orc_env_sample(const orc_handle *h, float u1, float u2, float dir[3], float *pdf, uint32_t *ix, uint32_t *iy);
/* The integer tables themselves: row prefix sums h x (w+1), marginal h+1. */
int orc_env_tables_raw(const orc_handle *h, uint64_t *row, uint64_t *marg);
float orc_env_pdf(const orc_handle *h, float dx, float dy, float dz);

#ifdef __cplusplus
}
#endif
#endif
