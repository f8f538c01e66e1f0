/*
 * omr_capi.h -- C ABI of the MI355X path-tracing library (libomr.so).
 *
 * The reference has no FFI: its operator API is the C++ virtual interface of
 * Renderer / PathTracer / Scene (src/renderer/Renderer.h:10-23,
 * src/renderer/PathTracer.h:18-88, src/scenes/Scene.h:18-124).  The C++
 * drop-in (omochi-renderer3_amd/dropin/) keeps that interface and calls these
 * entry points; tests and benchmarks bind them with ctypes.
 *
 * Conventions: 0 = OK, negative = error code (no exceptions, no exit()).  The
 * caller owns host buffers; the library owns device memory and handles.  A
 * context is driven by one host thread, except omr_job_submit / omr_job_poll /
 * omr_job_wait / omr_job_cancel / omr_server_stats, which are thread-safe.
 */
#ifndef OMR_CAPI_H
#define OMR_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct omr_ctx omr_ctx;
typedef int32_t omr_scene_id;
typedef int64_t omr_job_id;

enum {
  OMR_OK = 0,
  OMR_ERR_INVALID = -1,   /* bad argument or scene data */
  OMR_ERR_NOMEM = -2,     /* device or host allocation failed */
  OMR_ERR_HIP = -3,       /* HIP runtime error */
  OMR_ERR_BUDGET = -4,    /* request does not fit the HBM budget */
  OMR_ERR_NOTFOUND = -5,  /* unknown scene or job id */
  OMR_ERR_STATE = -6,     /* wrong call order (e.g. server not started) */
  OMR_ERR_NODEVICE = -7,  /* no usable GPU */
  OMR_ERR_TIMEOUT = -8,   /* omr_job_wait timed out */
  OMR_ERR_CANCELLED = -9  /* the job was cancelled */
};

/* Material::REFLECTION_TYPE (src/renderer/Material.h:13-17) */
enum { OMR_MAT_LAMBERT = 0, OMR_MAT_SPECULAR = 1, OMR_MAT_REFRACTION = 2 };

/* Material (src/renderer/Material.h:11-36) + multi-material rate
 * (SceneObject::MultiMaterialElem, src/renderer/SceneObject.h:17-30).  48 bytes. */
typedef struct {
  uint32_t type;
  float color[3];
  float emission[3];
  float ior;         /* refraction_rate */
  float rate;        /* roulette rate among the primitive's materials (normalized) */
  uint32_t texture;  /* 0 = none; k = the scene's textures[k-1] (Material::texture_id,
                        src/renderer/Material.h:33): the hit colour is color x texel */
  uint32_t pad[2];
} omr_material;

/* Texture image (Image, src/tools/Image.h:4-35): linear RGB, rows from the
 * top of the image; sampled by uv with u, v clamped to [0, 1] and scaled by
 * 0.99999 (Image::GetPixelByUV, :21-30).  width, height <= 32768. */
typedef struct {
  uint32_t width, height;
  const float *rgb;                /* height x width x 3 */
} omr_texture;

/* Scene contents handed to omr_scene_upload; replaces the object lists of
 * Scene (src/scenes/Scene.h:112-116) built by AddObject / AddModel.
 * Primitive ids: triangles 0..n_triangles-1, then spheres, then planes. */
typedef struct {
  size_t n_triangles;
  const float *triangles;        /* n x 9: v0.xyz, e1.xyz = v1-v0, e2.xyz = v2-v0 (Polygon.h:103-105,112) */
  const float *tri_normals;      /* n x 9: n0.xyz, n1-n0, n2-n0 (Polygon.h:109-111); NULL = face normal */
  size_t n_spheres;
  const float *spheres;          /* n x 4: center.xyz, radius (Sphere.h:13-19) */
  const uint32_t *prim_material; /* per primitive: index of its first material */
  const uint32_t *prim_material_count; /* per primitive: number of materials, NULL = 1 */
  size_t n_materials;
  const omr_material *materials;
  uint32_t env_width, env_height;  /* latlong environment (IBL, src/renderer/IBL.h:14-45) */
  const float *env_rgb;            /* env_height rows x env_width x 3, row 0 = +y; NULL = black */
  size_t n_planes;                 /* infinite floors y = planes[i], two-sided, outside the BVH and
                                      tested after it (InfiniteFloor, src/renderer/InfiniteFloor.h:10-45;
                                      Scene::CheckIntersection's not-in-BVH list, src/scenes/Scene.cpp:72-80) */
  const float *planes;
  size_t n_textures;               /* textures referenced by omr_material.texture */
  const omr_texture *textures;
  const float *tri_uv;             /* n_triangles x 6: uv0.xy, uv1-uv0, uv2-uv0 (Polygon.h:22-24,86-96);
                                      NULL = uv 0.  Spheres use the latlong uv of the hit normal
                                      (Sphere.h:52-75), planes uv 0 */
} omr_scene_desc;

/* Camera (src/renderer/Camera.h:12-43) in its derived form; the drop-in
 * Camera class computes it from position / direction / up exactly as the
 * reference constructor does. */
typedef struct {
  uint32_t width, height;
  float position[3];
  float screen_center[3];
  float screen_x[3];        /* full screen width vector */
  float screen_y[3];        /* full screen height vector */
  float aperture_radius;    /* CircleAperture radius (Aperture.h:23-35); 0 = pinhole */
  float focal_distance;
} omr_camera;

/* Render parameters; the reference's compile-time MinDepth/MaxDepth
 * (src/renderer/PathTracer.cpp:177-178) become rr_depth / max_depth. */
typedef struct {
  uint32_t supersamples;   /* ss: ss x ss sub-pixel grid (PathTracer.cpp:156-161) */
  uint32_t spp_begin;      /* samples already accumulated ("previous_samples") */
  uint32_t spp_end;        /* samples after this call ("next_samples") */
  uint32_t max_depth;      /* segments per path branch; 0 = unlimited (reference mode) */
  uint32_t rr_depth;       /* Russian roulette when depth > rr_depth (reference: 5) */
  uint32_t env_is;         /* 1 = environment importance sampling with shadow rays */
  uint64_t seed;           /* per-path seeds = hash(seed, pixel, sub-pixel, sample, branch) */
  uint32_t nee;            /* 1 = the reference's next-event estimation mode
                              (PathTracer::EnableNextEventEstimation, PathTracer.cpp:406-439):
                              surface emission counts only where the camera ray hits, and the
                              direct-light term it adds is zero (:312-314).  Scenes whose mirror
                              or glass materials emit are refused (OMR_ERR_INVALID): the reference
                              adds that emission when their continuation ray hits (:513-516) */
  uint32_t pad;
} omr_render_params;

typedef struct {
  uint64_t segments;       /* closest-hit queries (Scene::CheckIntersection calls) */
  uint64_t shadow_rays;    /* environment shadow rays */
  uint64_t bounce_steps;   /* extend/shade steps executed */
  double gpu_ms;           /* device time of the render (events) */
  double wall_ms;          /* host wall time of the call */
  /* filled when profiling is on (omr_set_profiling): per-kernel device time
   * from events around every launch, and traversal work counted by extend */
  double extend_ms, shade_ms, shadow_ms;
  uint64_t extend_launches;
  uint64_t node_visits, prim_tests;  /* extend launches only */
  /* drain (omr_tuning.drain_below): paths handed to it, their closest-hit
   * queries (part of `segments`), and its device time when profiling */
  uint64_t drained_paths, drain_segments;
  double drain_ms;
  uint64_t drain_node_visits, drain_prim_tests;  /* drain launch, when profiling */
  uint64_t drain_launches;
  uint64_t shadow_launches;  /* shadow-ray launches, when profiling (scenes with an environment sampler) */
} omr_render_stats;

typedef struct {
  uint32_t n_nodes;      /* 4-wide nodes */
  uint32_t depth;
  uint32_t stack_bound;
  uint32_t n_prims;
  uint64_t device_bytes;
  double build_ms;
  uint32_t n_nodes8;     /* the same tree as 8-wide nodes (same leaves) */
  uint32_t depth8;
  uint32_t stack_bound8;
  uint32_t pad;
} omr_scene_info;

/* Launch tuning (chosen by omr_autotune or set by hand). */
typedef struct {
  uint32_t node_width;        /* 4 (128-byte nodes) or 8 (256-byte nodes): the BVH the traversal reads */
  uint32_t wg_size;           /* 64, 128 or 256 */
  uint32_t blocks_per_cu;     /* persistent grid = CUs x blocks_per_cu (shade runs 20 waves per CU, 16 for one job) */
  uint32_t pool_paths;        /* path slots in flight (max) */
  uint32_t steps_per_sync;    /* bounce steps per host check (graph length) */
  uint32_t use_graph;         /* capture bounce steps in a hipGraph */
  uint32_t refill_every;      /* queue refill cadence: 0 = finished paths regenerate inside
                                 shade; k = idle slots are refilled in pixel order every k steps */
  uint32_t fetch_budget;      /* traversal: 0 = one ray per lane per grid-stride turn; k = lanes
                                 that finished fetch new rays after every k node/leaf visits (the
                                 drain slices its traversals the same way, 8 visits when 0) */
  uint32_t fetch_min;         /* idle lanes (of 64) that trigger a fetch */
  uint32_t stack_short;       /* traversal stack entries per lane kept in LDS: 8 or 16 (deeper
                                 entries spill to a per-thread global area) */
  uint32_t drain_below;       /* once no samples are left to hand out and at most this many paths
                                 are queued, one drain launch runs them all to the end instead of
                                 further extend/shade steps (0 = never) */
  uint32_t record_mb;         /* device memory (MiB) a render may use for per-sample sums (24 bytes
                                 per sample): jobs whose records fit add their contributions there
                                 without atomics and a gather sums them per pixel at the end; the
                                 rest add to their pixel sums atomically (0 = atomics only).  The
                                 results are the same bits either way */
  uint32_t drain_rounds;      /* 1: the run is a sequence of rounds, each filling every idle path
                                 slot with a new sample and running all of them to their end in one
                                 drain launch (no path state through memory between segments);
                                 0: extend/shade bounce steps until the drain threshold */
  uint32_t round_steps;       /* drain rounds: extend/shade bounce steps (even, 0..16) before each
                                 round's drain -- the first bounces of fresh camera paths run as
                                 wide, coherent steps */
  uint32_t solo_round_min;    /* drain rounds: a job with at least this many samples left ends in
                                 a round of its own, run by the single-job kernels, instead of
                                 sharing one with the next job (0 = half the pool) */
  uint32_t lanes;             /* drain rounds: 1 or 2 path pools of pool_paths slots each; with 2,
                                 the consecutive rounds of a single job alternate between them on
                                 two streams, so one round's last, longest paths run beside the
                                 next round's first steps (twice the pool memory; batches keep
                                 one lane) */
} omr_tuning;

typedef struct {
  uint64_t pool_paths;        /* path slots (0 = the tuning's pool_paths; restored at omr_server_stop) */
  uint64_t hbm_budget_bytes;  /* device memory the server may hold (0 = 64 GiB): the path pool, its
                                 queues and traversal stacks, the resident scenes and the admitted
                                 requests' buffers.  omr_server_start fails with OMR_ERR_BUDGET when
                                 the pool and scenes alone exceed it; a request whose sums and image
                                 (12 bytes per pixel) cannot fit is rejected at submission; per-sample
                                 records (24 bytes per sample) are added only where they fit */
  uint32_t max_live_jobs;     /* 0 = 512 */
} omr_server_config;

typedef struct {
  uint64_t jobs_submitted, jobs_done, jobs_rejected, jobs_cancelled;
  uint64_t segments, shadow_rays;
  double busy_ms;             /* wall time with work on the device */
  double lat_p50_ms, lat_p99_ms, lat_max_ms;  /* submit -> result on host */
  uint64_t fixed_bytes;       /* budgeted device memory of the path pool, queues, stacks and scenes */
  uint64_t used_bytes;        /* budgeted device memory of the admitted requests */
} omr_server_stats;

/* ---- context ---- */
int omr_create(omr_ctx **ctx, int device);
void omr_destroy(omr_ctx *ctx);
const char *omr_last_error(omr_ctx *ctx);
int omr_device_info(omr_ctx *ctx, int *n_cu, uint64_t *hbm_bytes, char *name, size_t name_len);

/* ---- scenes: Scene::ConstructQBVH (src/scenes/Scene.cpp:36-49) becomes an
 * upload that builds the wide BVH on the host and keeps it resident ---- */
int omr_scene_upload(omr_ctx *ctx, const omr_scene_desc *desc, omr_scene_id *id);
int omr_scene_release(omr_ctx *ctx, omr_scene_id id);
int omr_scene_info_get(omr_ctx *ctx, omr_scene_id id, omr_scene_info *info);
/* Host copy of the built BVH (128-byte nodes; leaf slot -> primitive id). */
int omr_scene_bvh(omr_ctx *ctx, omr_scene_id id, const void **nodes, size_t *n_nodes, const uint32_t **prim_order,
                  size_t *n_order);
/* Host-only BVH build (no device needed): the same builder omr_scene_upload
 * runs.  The handle owns the arrays until omr_bvh_free. */
typedef struct omr_bvh omr_bvh;
int omr_bvh_build(const omr_scene_desc *desc, omr_bvh **out);
int omr_bvh_data(const omr_bvh *b, const void **nodes, size_t *n_nodes, const uint32_t **prim_order, size_t *n_order,
                 uint32_t *depth, uint32_t *stack_bound);
/* The 8-wide collapse of the same tree (256-byte nodes, the same leaf codes
 * and prim_order as the 4-wide nodes). */
int omr_bvh_data8(const omr_bvh *b, const void **nodes, size_t *n_nodes, uint32_t *depth, uint32_t *stack_bound);
void omr_bvh_free(omr_bvh *b);
/* Environment sampling tables as built on the device (uint64 prefix sums). */
int omr_scene_env_tables(omr_ctx *ctx, omr_scene_id id, uint64_t *row /* h x (w+1) */, uint64_t *marg /* h+1 */);

/* ---- closest hit: Scene::CheckIntersection (src/scenes/Scene.cpp:51-85) ----
 * rays: n x 8 floats (o.xyz, tmax, d.xyz, pad).  prim_id = -1 on a miss,
 * t = INF then.  _host: host buffers; _device: device pointers (no copies). */
int omr_trace_closest(omr_ctx *ctx, omr_scene_id scene, const float *rays, size_t n, int32_t *prim_id, float *t);
int omr_trace_closest_device(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n, int32_t *d_prim_id,
                             float *d_t);
/* omr_trace_closest_device without any device allocation (no hipMalloc,
 * hipFree or allocating library call).  No counterpart in the reference: it is
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc.  Its working
 * memory (fetch heads and traversal overflow stacks) is one fixed 32 MiB
 * device array per GPU (each process has its own), held until the process
 * ends and shared by every call on that GPU: calls take turns under a lock
 * per GPU and each returns only when its work is done, so they never overlap.
 * The array keeps the last call's contents; each call clears the
 * part that must start at zero.  The traversal grid shrinks to what the array
 * holds; a scene whose stack bound leaves no room for one workgroup is
 * refused with OMR_ERR_INVALID.  Same results as omr_trace_closest_device. */
int omr_trace_closest_device_noalloc(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n,
                                     int32_t *d_prim_id, float *d_t);
/* Node visits / primitive tests for a ray set (device pointers; stats pass). */
int omr_trace_stats_device(omr_ctx *ctx, omr_scene_id scene, const float *d_rays, size_t n, int32_t *d_prim_id,
                           float *d_t, uint64_t *node_visits, uint64_t *prim_tests);

/* ---- rendering: PathTracer::RenderScene for one sample batch
 * (src/renderer/PathTracer.cpp:55-175).  out_rgb: width x height x 3 floats,
 * top-down rows, mean radiance over (spp_begin, spp_end] x ss^2 samples. ---- */
int omr_render(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
               float *out_rgb);
/* Fixed-point per-pixel sums (x 2^-24), width x height x 3, for bit-level checks. */
int omr_render_sums(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
                    int64_t *out_sums);
int omr_last_render_stats(omr_ctx *ctx, omr_render_stats *stats);

/* Batched: n jobs packed into one persistent wavefront run (shared queues). */
int omr_render_batch(omr_ctx *ctx, size_t n, const omr_scene_id *scenes, const omr_camera *cams,
                     const omr_render_params *params, float *const *outs);

/* Profiling: launches bounce steps without a graph, times each kernel with
 * HIP events on the library's stream and counts traversal work. */
int omr_set_profiling(omr_ctx *ctx, int enable);

/* ---- tuning ---- */
int omr_get_tuning(omr_ctx *ctx, omr_tuning *t);
int omr_set_tuning(omr_ctx *ctx, const omr_tuning *t);

/* Autotuner: coordinate search over node width, workgroup size, grid
 * occupancy, pool size, queue refill cadence and steps per host check, each
 * candidate scored by device-timed Mrays/s of the given jobs (run batched).
 * Leaves the best tuning set on the context; stops after budget_ms. */
int omr_autotune(omr_ctx *ctx, size_t n, const omr_scene_id *scenes, const omr_camera *cams,
                 const omr_render_params *params, uint32_t budget_ms, omr_tuning *best);

/* ---- served mode: continuous batching with HBM-budget admission ---- */
int omr_server_start(omr_ctx *ctx, const omr_server_config *cfg);
int omr_server_stop(omr_ctx *ctx);
int omr_job_submit(omr_ctx *ctx, omr_scene_id scene, const omr_camera *cam, const omr_render_params *params,
                   float *out_rgb, omr_job_id *job);
/* state: 0 queued, 1 running, 2 done, 3 cancelled, 4 failed (device or HIP error; the
 * server stops serving), 5 rejected (does not fit the HBM budget) */
int omr_job_poll(omr_ctx *ctx, omr_job_id job, int *state);
/* OMR_OK when done; OMR_ERR_CANCELLED, OMR_ERR_TIMEOUT, or the failure's own code
 * (OMR_ERR_BUDGET for a rejection, OMR_ERR_HIP / OMR_ERR_NOMEM / OMR_ERR_STATE
 * for a failure). */
int omr_job_wait(omr_ctx *ctx, omr_job_id job, int64_t timeout_ms);
/* The error text of a failed or rejected job (empty otherwise). */
int omr_job_error(omr_ctx *ctx, omr_job_id job, char *msg, size_t len);
/* Submission-to-result latency of a done job (OMR_ERR_STATE otherwise). */
int omr_job_latency(omr_ctx *ctx, omr_job_id job, double *ms);
int omr_job_cancel(omr_ctx *ctx, omr_job_id job);
int omr_server_stats_get(omr_ctx *ctx, omr_server_stats *stats);

/* ---- served mode on a fake clock (host only, no device; no reference
 * counterpart: tests and capacity planning).  The server's own admission,
 * arena and latency code runs as a discrete-event simulation: requests arrive
 * at the given times, each is planned at submission (rejected when its image
 * exceeds the arena; per-sample records when they fit the record limit and,
 * with the image, the arena), admitted in submission order while the live
 * window has room and the arena a block (without records when only that
 * fits), and a modelled device finishes samples at `samples_per_ms` in
 * admission order; a request completes, and returns its block, when the end
 * of its sample range is finished. */
typedef struct omr_sim_request {
  double arrival_ms;           /* submission time; nondecreasing over the array */
  uint32_t width, height;      /* image */
  uint32_t supersamples, spp;  /* samples = width x height x supersamples^2 x spp */
  uint32_t cancel;             /* cancelled by the caller before its admission */
  uint32_t pad;
} omr_sim_request;

typedef struct omr_sim_config {
  uint64_t arena_bytes;         /* what the path pool and scenes leave of the HBM budget */
  uint64_t record_limit_bytes;  /* per-request limit on per-sample records */
  uint32_t max_live_jobs;       /* live window, 1..1024 */
  uint32_t pad;
  double samples_per_ms;        /* device model: samples finished per ms */
} omr_sim_config;

typedef struct omr_sim_outcome {
  int32_t state;                       /* 2 done, 3 cancelled, 5 rejected */
  uint32_t records;                    /* 1: per-sample records, 0: pixel atomics */
  uint64_t off, bytes;                 /* its arena block while live */
  double admit_ms, done_ms, latency_ms; /* -1 when it never ran */
} omr_sim_outcome;

typedef struct omr_sim_summary {
  uint64_t jobs_done, jobs_rejected, jobs_cancelled;
  uint64_t peak_used_bytes;  /* most arena bytes held by live requests at once */
  uint32_t peak_live_jobs, pad;
  double lat_p50_ms, lat_p99_ms, lat_max_ms;  /* over done requests, as omr_server_stats_get */
  double makespan_ms;                          /* time of the last event */
} omr_sim_summary;

int omr_serve_simulate(const omr_sim_request *requests, size_t n, const omr_sim_config *cfg,
                       omr_sim_outcome *outcomes, omr_sim_summary *summary);

/* ---- procedural benchmark scenes (no reference counterpart: the inputs the
 * measurements use).  kind: 0 Cornell box, 1 Cornell box + displaced
 * icosphere (level 6), 2 heightfield terrain + 200 spheres + HDR sky,
 * 3 displaced torus knot.  The library owns the arrays until
 * omr_procedural_free. ---- */
typedef struct {
  omr_scene_desc desc;
  omr_camera camera;
  void *internal;
} omr_procedural;
int omr_procedural_scene(int kind, uint64_t seed, uint32_t width, uint32_t height, omr_procedural *out);
void omr_procedural_free(omr_procedural *p);
/* The camera preset of a procedural kind at another image size (kind 4 is a
 * 2000 x 1000 terrain, 3,996,002 triangles, with its own sky). */
int omr_procedural_camera(int kind, uint32_t width, uint32_t height, omr_camera *out);

/* ---- file formats (host only: no context or GPU needed).  Errors:
 * omr_io_last_error() of the calling thread. ---- */
const char *omr_io_last_error(void);
/* Releases memory returned by the loaders below. */
void omr_free(void *p);
/* HDRImage::ReadFromRadianceFile (src/tools/HDRImage.cpp:12-27, 51-186):
 * Radiance RGBE, run-length encoded or flat, into *rgb = height rows x width
 * x 3 floats (the file's first scanline first).  Release with omr_free. */
int omr_hdr_read(const char *path, uint32_t *width, uint32_t *height, float **rgb);
/* HDRImage::WriteToRadianceFile (src/tools/HDRImage.cpp:29-49, 188-314):
 * RGBE with run-length encoded scanlines, byte for byte as the reference
 * writes them. */
int omr_hdr_write(const char *path, uint32_t width, uint32_t height, const float *rgb);
/* The image library call of ImageHandler::LoadFromFile
 * (src/tools/ImageHandler.cpp:54, stbi_load with native channels): a PNG
 * (8-bit samples, any colour type, interlaced or not) or baseline JPEG file
 * as 8-bit samples, rows from the top, *channels per pixel.  Release *px with
 * omr_free. */
int omr_image_read(const char *path, uint32_t *width, uint32_t *height, uint32_t *channels, uint8_t **px);
/* ImageHandler::LoadFromFile (src/tools/ImageHandler.cpp:20-90) as a texture:
 * linear RGB floats, rows from the top; .hdr files as Radiance RGBE, other
 * files b / 255 per sample, raised to 2.2 when inverse_gamma.  Release *rgb
 * with omr_free. */
int omr_texture_read(const char *path, int inverse_gamma, uint32_t *width, uint32_t *height, float **rgb);
/* A scene read from a file: desc points into memory the handle owns (until
 * omr_loaded_scene_free), textures included. */
typedef struct {
  omr_scene_desc desc;
  size_t n_objects;               /* objects of the scene file (a mesh counts its polygons) */
  void *internal;
} omr_loaded_scene;
/* SceneFromExternalFile (src/scenes/SceneFromExternalFile.cpp:46-531): the
 * reference's .scene text format, with meshes (Wavefront OBJ/MTL), spheres,
 * sphere lights, floors, infinite floors and a Radiance IBL. */
int omr_scene_file_load(const char *path, omr_loaded_scene *out);
/* Model::ReadFromObj + Model::Transform (src/renderer/Model.cpp:50-182,
 * 38-48): one OBJ mesh placed by translation, per-axis scale and rotation
 * (radians, applied x, then y, then z as in SceneFromExternalFile.cpp:523-526). */
int omr_obj_load(const char *path, const double position[3], const double scale[3], const double rotation[3],
                 int flip_v, omr_loaded_scene *out);
void omr_loaded_scene_free(omr_loaded_scene *s);

/* Reference Camera constructor (src/renderer/Camera.h:12-43), double precision:
 * fills the derived screen vectors of an omr_camera. */
void omr_camera_from_lookat(uint32_t width, uint32_t height, const double pos[3], const double dir[3],
                            const double up[3], double screen_height, double dist_to_screen, double focal_distance,
                            double aperture_radius, omr_camera *out);

#ifdef __cplusplus
}
#endif
#endif
