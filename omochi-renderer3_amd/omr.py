"""ctypes binding of libomr.so (include/omr_capi.h) for tests and benchmarks.

The GPU path is the only path: if lib/libomr.so is missing, or no GPU is
present, the calls raise instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libomr.so")

OMR_OK = 0
OMR_ERR_INVALID = -1
ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_BUDGET, ERR_NOTFOUND, ERR_STATE, ERR_NODEVICE, ERR_TIMEOUT, ERR_CANCELLED = \
    -1, -2, -3, -4, -5, -6, -7, -8, -9
ERRORS = {
    -1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "over HBM budget",
    -5: "not found", -6: "bad state", -7: "no device", -8: "timeout", -9: "cancelled",
}
LAMBERT, SPECULAR, REFRACTION = 0, 1, 2


class Material(C.Structure):
    _fields_ = [("type", C.c_uint32), ("color", C.c_float * 3), ("emission", C.c_float * 3),
                ("ior", C.c_float), ("rate", C.c_float), ("texture", C.c_uint32), ("pad", C.c_uint32 * 2)]


MATERIAL_DTYPE = np.dtype([("type", "<u4"), ("color", "<f4", 3), ("emission", "<f4", 3), ("ior", "<f4"),
                           ("rate", "<f4"), ("texture", "<u4"), ("pad", "<u4", 2)])
assert MATERIAL_DTYPE.itemsize == C.sizeof(Material) == 48


class SceneDesc(C.Structure):
    _fields_ = [("n_triangles", C.c_size_t), ("triangles", C.c_void_p), ("tri_normals", C.c_void_p),
                ("n_spheres", C.c_size_t), ("spheres", C.c_void_p), ("prim_material", C.c_void_p),
                ("prim_material_count", C.c_void_p), ("n_materials", C.c_size_t), ("materials", C.c_void_p),
                ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_rgb", C.c_void_p),
                ("n_planes", C.c_size_t), ("planes", C.c_void_p), ("n_textures", C.c_size_t),
                ("textures", C.c_void_p), ("tri_uv", C.c_void_p)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", C.c_void_p)]


class LoadedScene(C.Structure):
    _fields_ = [("desc", SceneDesc), ("n_objects", C.c_size_t), ("internal", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("position", C.c_float * 3),
                ("screen_center", C.c_float * 3), ("screen_x", C.c_float * 3), ("screen_y", C.c_float * 3),
                ("aperture_radius", C.c_float), ("focal_distance", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("supersamples", C.c_uint32), ("spp_begin", C.c_uint32), ("spp_end", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("env_is", C.c_uint32), ("seed", C.c_uint64),
                ("nee", C.c_uint32), ("pad", C.c_uint32)]


class RenderStats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("shadow_rays", C.c_uint64), ("bounce_steps", C.c_uint64),
                ("gpu_ms", C.c_double), ("wall_ms", C.c_double), ("extend_ms", C.c_double),
                ("shade_ms", C.c_double), ("shadow_ms", C.c_double), ("extend_launches", C.c_uint64),
                ("node_visits", C.c_uint64), ("prim_tests", C.c_uint64), ("drained_paths", C.c_uint64),
                ("drain_segments", C.c_uint64), ("drain_ms", C.c_double), ("drain_node_visits", C.c_uint64),
                ("drain_prim_tests", C.c_uint64), ("drain_launches", C.c_uint64),
                ("shadow_launches", C.c_uint64)]


class SceneInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("depth", C.c_uint32), ("stack_bound", C.c_uint32), ("n_prims", C.c_uint32),
                ("device_bytes", C.c_uint64), ("build_ms", C.c_double), ("n_nodes8", C.c_uint32),
                ("depth8", C.c_uint32), ("stack_bound8", C.c_uint32), ("pad", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("node_width", C.c_uint32), ("wg_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("pool_paths", C.c_uint32), ("steps_per_sync", C.c_uint32), ("use_graph", C.c_uint32),
                ("refill_every", C.c_uint32), ("fetch_budget", C.c_uint32), ("fetch_min", C.c_uint32),
                ("stack_short", C.c_uint32), ("drain_below", C.c_uint32), ("record_mb", C.c_uint32),
                ("drain_rounds", C.c_uint32), ("round_steps", C.c_uint32), ("solo_round_min", C.c_uint32),
                ("lanes", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("pool_paths", C.c_uint64), ("hbm_budget_bytes", C.c_uint64), ("max_live_jobs", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_uint64), ("jobs_done", C.c_uint64), ("jobs_rejected", C.c_uint64),
                ("jobs_cancelled", C.c_uint64), ("segments", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("busy_ms", C.c_double), ("lat_p50_ms", C.c_double), ("lat_p99_ms", C.c_double),
                ("lat_max_ms", C.c_double), ("fixed_bytes", C.c_uint64), ("used_bytes", C.c_uint64)]


class SimRequest(C.Structure):
    _fields_ = [("arrival_ms", C.c_double), ("width", C.c_uint32), ("height", C.c_uint32),
                ("supersamples", C.c_uint32), ("spp", C.c_uint32), ("cancel", C.c_uint32), ("pad", C.c_uint32)]


class SimConfig(C.Structure):
    _fields_ = [("arena_bytes", C.c_uint64), ("record_limit_bytes", C.c_uint64), ("max_live_jobs", C.c_uint32),
                ("pad", C.c_uint32), ("samples_per_ms", C.c_double)]


class SimOutcome(C.Structure):
    _fields_ = [("state", C.c_int32), ("records", C.c_uint32), ("off", C.c_uint64), ("bytes", C.c_uint64),
                ("admit_ms", C.c_double), ("done_ms", C.c_double), ("latency_ms", C.c_double)]


class SimSummary(C.Structure):
    _fields_ = [("jobs_done", C.c_uint64), ("jobs_rejected", C.c_uint64), ("jobs_cancelled", C.c_uint64),
                ("peak_used_bytes", C.c_uint64), ("peak_live_jobs", C.c_uint32), ("pad", C.c_uint32),
                ("lat_p50_ms", C.c_double), ("lat_p99_ms", C.c_double), ("lat_max_ms", C.c_double),
                ("makespan_ms", C.c_double)]


class Procedural(C.Structure):
    _fields_ = [("desc", SceneDesc), ("camera", Camera), ("internal", C.c_void_p)]


EXPORTS = [
    "omr_create", "omr_destroy", "omr_last_error", "omr_device_info", "omr_scene_upload", "omr_scene_release",
    "omr_scene_info_get", "omr_scene_bvh", "omr_scene_env_tables", "omr_trace_closest", "omr_trace_closest_device",
    "omr_trace_stats_device", "omr_trace_closest_device_noalloc", "omr_render", "omr_render_sums", "omr_last_render_stats", "omr_render_batch",
    "omr_get_tuning", "omr_set_tuning", "omr_server_start", "omr_server_stop", "omr_job_submit", "omr_job_poll",
    "omr_job_wait", "omr_job_error", "omr_job_latency", "omr_job_cancel", "omr_server_stats_get",
    "omr_serve_simulate", "omr_procedural_scene", "omr_procedural_free",
    "omr_camera_from_lookat", "omr_bvh_build", "omr_bvh_data", "omr_bvh_data8", "omr_bvh_free", "omr_set_profiling", "omr_autotune",
    "omr_procedural_camera", "omr_io_last_error", "omr_free", "omr_hdr_read", "omr_hdr_write", "omr_scene_file_load",
    "omr_obj_load", "omr_loaded_scene_free", "omr_image_read", "omr_texture_read",
]

_lib = None


def load():
    """Loads lib/libomr.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() or make -C {HERE}")
    lib = C.CDLL(LIB_PATH)
    vp, i32, sz, u64 = C.c_void_p, C.c_int32, C.c_size_t, C.c_uint64
    lib.omr_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.omr_destroy.argtypes = [vp]
    lib.omr_destroy.restype = None
    lib.omr_last_error.argtypes = [vp]
    lib.omr_last_error.restype = C.c_char_p
    lib.omr_device_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(u64), C.c_char_p, sz]
    lib.omr_scene_upload.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(i32)]
    lib.omr_scene_release.argtypes = [vp, i32]
    lib.omr_scene_info_get.argtypes = [vp, i32, C.POINTER(SceneInfo)]
    lib.omr_scene_bvh.argtypes = [vp, i32, C.POINTER(vp), C.POINTER(sz), C.POINTER(vp), C.POINTER(sz)]
    lib.omr_scene_env_tables.argtypes = [vp, i32, vp, vp]
    lib.omr_trace_closest.argtypes = [vp, i32, vp, sz, vp, vp]
    lib.omr_trace_closest_device.argtypes = [vp, i32, vp, sz, vp, vp]
    lib.omr_trace_closest_device_noalloc.argtypes = [vp, i32, vp, sz, vp, vp]
    lib.omr_trace_stats_device.argtypes = [vp, i32, vp, sz, vp, vp, C.POINTER(u64), C.POINTER(u64)]
    lib.omr_render.argtypes = [vp, i32, C.POINTER(Camera), C.POINTER(RenderParams), vp]
    lib.omr_render_sums.argtypes = [vp, i32, C.POINTER(Camera), C.POINTER(RenderParams), vp]
    lib.omr_last_render_stats.argtypes = [vp, C.POINTER(RenderStats)]
    lib.omr_render_batch.argtypes = [vp, sz, vp, vp, vp, vp]
    lib.omr_set_profiling.argtypes = [vp, C.c_int]
    lib.omr_autotune.argtypes = [vp, sz, vp, vp, vp, C.c_uint32, vp]
    lib.omr_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.omr_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.omr_server_start.argtypes = [vp, C.POINTER(ServerConfig)]
    lib.omr_server_stop.argtypes = [vp]
    lib.omr_job_submit.argtypes = [vp, i32, C.POINTER(Camera), C.POINTER(RenderParams), vp, C.POINTER(C.c_int64)]
    lib.omr_job_poll.argtypes = [vp, C.c_int64, C.POINTER(C.c_int)]
    lib.omr_job_wait.argtypes = [vp, C.c_int64, C.c_int64]
    lib.omr_job_error.argtypes = [vp, C.c_int64, C.c_char_p, sz]
    lib.omr_job_latency.argtypes = [vp, C.c_int64, C.POINTER(C.c_double)]
    lib.omr_job_cancel.argtypes = [vp, C.c_int64]
    lib.omr_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
    lib.omr_serve_simulate.argtypes = [C.POINTER(SimRequest), sz, C.POINTER(SimConfig), C.POINTER(SimOutcome),
                                       C.POINTER(SimSummary)]
    lib.omr_procedural_scene.argtypes = [C.c_int, u64, C.c_uint32, C.c_uint32, C.POINTER(Procedural)]
    lib.omr_procedural_free.argtypes = [C.POINTER(Procedural)]
    lib.omr_procedural_free.restype = None
    lib.omr_procedural_camera.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(Camera)]
    lib.omr_camera_from_lookat.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_double * 3),
                                           C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3), C.c_double,
                                           C.c_double, C.c_double, C.c_double, C.POINTER(Camera)]
    lib.omr_camera_from_lookat.restype = None
    lib.omr_bvh_build.argtypes = [C.POINTER(SceneDesc), C.POINTER(vp)]
    lib.omr_bvh_data.argtypes = [vp, C.POINTER(vp), C.POINTER(sz), C.POINTER(vp), C.POINTER(sz),
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.omr_bvh_data8.argtypes = [vp, C.POINTER(vp), C.POINTER(sz), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.omr_bvh_free.argtypes = [vp]
    lib.omr_bvh_free.restype = None
    lib.omr_io_last_error.argtypes = []
    lib.omr_io_last_error.restype = C.c_char_p
    lib.omr_free.argtypes = [vp]
    lib.omr_free.restype = None
    lib.omr_hdr_read.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(vp)]
    lib.omr_hdr_write.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, vp]
    lib.omr_scene_file_load.argtypes = [C.c_char_p, C.POINTER(LoadedScene)]
    lib.omr_obj_load.argtypes = [C.c_char_p, C.POINTER(C.c_double * 3), C.POINTER(C.c_double * 3),
                                 C.POINTER(C.c_double * 3), C.c_int, C.POINTER(LoadedScene)]
    lib.omr_loaded_scene_free.argtypes = [C.POINTER(LoadedScene)]
    lib.omr_loaded_scene_free.restype = None
    lib.omr_image_read.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                   C.POINTER(vp)]
    lib.omr_texture_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(vp)]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Scene:
    """A scene as numpy arrays (the omr_scene_desc layout)."""

    def __init__(self, tri, tri_n, sph, prim_mat, mats, prim_mat_count=None, env=None, camera=None, planes=None,
                 textures=None, tri_uv=None):
        self.planes = np.zeros(0, np.float32) if planes is None else np.ascontiguousarray(planes, dtype=np.float32)
        # textures: list of (height, width, 3) linear RGB arrays, rows from the top;
        # materials refer to them by texture = index + 1
        self.textures = [np.ascontiguousarray(t, dtype=np.float32) for t in (textures or [])]
        self.tri_uv = None if tri_uv is None else np.ascontiguousarray(tri_uv, dtype=np.float32).reshape(-1, 6)
        self.tri = np.ascontiguousarray(tri, dtype=np.float32).reshape(-1, 9)
        self.tri_n = None if tri_n is None else np.ascontiguousarray(tri_n, dtype=np.float32).reshape(-1, 9)
        self.sph = np.ascontiguousarray(sph, dtype=np.float32).reshape(-1, 4)
        self.prim_mat = np.ascontiguousarray(prim_mat, dtype=np.uint32)
        self.prim_mat_count = None if prim_mat_count is None else np.ascontiguousarray(prim_mat_count, dtype=np.uint32)
        self.mats = np.ascontiguousarray(mats, dtype=MATERIAL_DTYPE)
        self.env = None if env is None else np.ascontiguousarray(env, dtype=np.float32)
        self.camera = camera

    @property
    def n_prims(self):
        return self.tri.shape[0] + self.sph.shape[0] + self.planes.shape[0]

    def desc(self):
        d = SceneDesc()
        d.n_triangles = self.tri.shape[0]
        d.triangles = _ptr(self.tri) if self.tri.shape[0] else None
        d.tri_normals = _ptr(self.tri_n) if self.tri_n is not None and self.tri.shape[0] else None
        d.n_spheres = self.sph.shape[0]
        d.spheres = _ptr(self.sph) if self.sph.shape[0] else None
        d.prim_material = _ptr(self.prim_mat)
        d.prim_material_count = _ptr(self.prim_mat_count)
        d.n_materials = self.mats.shape[0]
        d.materials = _ptr(self.mats)
        if self.env is not None:
            d.env_height, d.env_width = self.env.shape[0], self.env.shape[1]
            d.env_rgb = _ptr(self.env)
        d.n_planes = self.planes.shape[0]
        d.planes = _ptr(self.planes) if self.planes.shape[0] else None
        if self.textures:
            self._tex = (Texture * len(self.textures))()
            for k, t in enumerate(self.textures):
                self._tex[k].height, self._tex[k].width = t.shape[0], t.shape[1]
                self._tex[k].rgb = _ptr(t)
            d.n_textures = len(self.textures)
            d.textures = C.cast(self._tex, C.c_void_p)
        d.tri_uv = _ptr(self.tri_uv) if self.tri_uv is not None and self.tri.shape[0] else None
        return d


def _scene_from_desc(d, camera=None):
    """numpy copy of an omr_scene_desc the library owns."""
    nt, ns, npl = d.n_triangles, d.n_spheres, d.n_planes

    def arr(ptr, n, dt):
        if n == 0 or not ptr:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,)).copy()

    tri = arr(d.triangles, nt * 9, np.float32).reshape(-1, 9)
    trn = arr(d.tri_normals, nt * 9, np.float32).reshape(-1, 9)
    sph = arr(d.spheres, ns * 4, np.float32).reshape(-1, 4)
    pm = arr(d.prim_material, nt + ns + npl, np.uint32)
    pc = arr(d.prim_material_count, nt + ns + npl, np.uint32) if d.prim_material_count else None
    mats = np.frombuffer(C.string_at(d.materials, d.n_materials * 48), dtype=MATERIAL_DTYPE).copy()
    env = None
    if d.env_rgb:
        env = arr(d.env_rgb, d.env_width * d.env_height * 3, np.float32).reshape(d.env_height, d.env_width, 3)
    planes = arr(d.planes, npl, np.float32)
    textures = []
    if d.n_textures and d.textures:
        tx = C.cast(d.textures, C.POINTER(Texture))
        for k in range(d.n_textures):
            w, h = tx[k].width, tx[k].height
            textures.append(arr(tx[k].rgb, w * h * 3, np.float32).reshape(h, w, 3))
    uv = arr(d.tri_uv, nt * 6, np.float32).reshape(-1, 6) if d.tri_uv else None
    return Scene(tri, trn, sph, pm, mats, pc, env, camera, planes, textures, uv)


def _io_error(lib, what):
    return OmrError(OMR_ERR_INVALID, f"{what}: {lib.omr_io_last_error().decode(errors='replace')}")


def hdr_read(path):
    """Radiance RGBE image -> float32 array (height, width, 3), first scanline first."""
    lib = load()
    w, h, p = C.c_uint32(), C.c_uint32(), C.c_void_p()
    if lib.omr_hdr_read(os.fsencode(path), C.byref(w), C.byref(h), C.byref(p)) != OMR_OK:
        raise _io_error(lib, "omr_hdr_read")
    try:
        n = w.value * h.value * 3
        a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n,)).copy()
        return a.reshape(h.value, w.value, 3)
    finally:
        lib.omr_free(p)


def hdr_write(path, img):
    """Writes a float (height, width, 3) image as run-length encoded RGBE."""
    lib = load()
    a = np.ascontiguousarray(img, dtype=np.float32)
    if lib.omr_hdr_write(os.fsencode(path), a.shape[1], a.shape[0], _ptr(a)) != OMR_OK:
        raise _io_error(lib, "omr_hdr_write")


def image_read(path):
    """PNG / baseline JPEG file -> uint8 array (height, width, channels), rows from the top."""
    lib = load()
    w, h, c, p = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_void_p()
    if lib.omr_image_read(os.fsencode(path), C.byref(w), C.byref(h), C.byref(c), C.byref(p)) != OMR_OK:
        raise _io_error(lib, "omr_image_read")
    try:
        n = w.value * h.value * c.value
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(n,)).copy().reshape(h.value, w.value,
                                                                                                  c.value)
    finally:
        lib.omr_free(p)


def texture_read(path, inverse_gamma=True):
    """An image file as the renderer's texture: float32 (height, width, 3), linear."""
    lib = load()
    w, h, p = C.c_uint32(), C.c_uint32(), C.c_void_p()
    if lib.omr_texture_read(os.fsencode(path), int(inverse_gamma), C.byref(w), C.byref(h), C.byref(p)) != OMR_OK:
        raise _io_error(lib, "omr_texture_read")
    try:
        n = w.value * h.value * 3
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n,)).copy().reshape(h.value, w.value, 3)
    finally:
        lib.omr_free(p)


def serve_simulate(requests, arena_bytes, samples_per_ms, record_limit_bytes=None, max_live_jobs=512):
    """Served mode's admission, arena and latency code on a fake clock (host
    only).  requests: dicts with arrival_ms, w, h, ss, spp and optional
    cancel.  Returns (outcomes as a numpy structured array, summary dict)."""
    lib = load()
    n = len(requests)
    arr = (SimRequest * max(1, n))()
    for i, r in enumerate(requests):
        arr[i] = SimRequest(float(r["arrival_ms"]), r["w"], r["h"], r["ss"], r["spp"], int(r.get("cancel", 0)), 0)
    cfg = SimConfig(int(arena_bytes), int(arena_bytes // 4 if record_limit_bytes is None else record_limit_bytes),
                    int(max_live_jobs), 0, float(samples_per_ms))
    out = (SimOutcome * max(1, n))()
    sm = SimSummary()
    rc = lib.omr_serve_simulate(arr, n, C.byref(cfg), out, C.byref(sm))
    if rc != OMR_OK:
        raise OmrError(rc, "omr_serve_simulate failed")
    dt = np.dtype([("state", np.int32), ("records", np.uint32), ("off", np.uint64), ("bytes", np.uint64),
                   ("admit_ms", np.float64), ("done_ms", np.float64), ("latency_ms", np.float64)])
    res = np.frombuffer(C.string_at(out, C.sizeof(SimOutcome) * n), dtype=dt).copy() if n else np.zeros(0, dt)
    return res, {k: getattr(sm, k) for k, _ in SimSummary._fields_ if k != "pad"}


def load_scene_file(path):
    """A .scene file (the reference's text format) as a Scene."""
    lib = load()
    ls = LoadedScene()
    if lib.omr_scene_file_load(os.fsencode(path), C.byref(ls)) != OMR_OK:
        raise _io_error(lib, "omr_scene_file_load")
    try:
        return _scene_from_desc(ls.desc)
    finally:
        lib.omr_loaded_scene_free(C.byref(ls))


def load_obj(path, position=(0, 0, 0), scale=(1, 1, 1), rotation=(0, 0, 0), flip_v=False):
    """One Wavefront OBJ mesh (with its MTL materials), placed, as a Scene."""
    lib = load()
    ls = LoadedScene()
    v3 = C.c_double * 3
    if lib.omr_obj_load(os.fsencode(path), v3(*position), v3(*scale), v3(*rotation), int(flip_v), C.byref(ls)) != OMR_OK:
        raise _io_error(lib, "omr_obj_load")
    try:
        return _scene_from_desc(ls.desc)
    finally:
        lib.omr_loaded_scene_free(C.byref(ls))


def procedural(kind, seed=0x6F6D7233, width=256, height=256):
    """Builds a procedural benchmark scene; returns a Scene (numpy copies)."""
    lib = load()
    p = Procedural()
    rc = lib.omr_procedural_scene(kind, seed, width, height, C.byref(p))
    if rc != OMR_OK:
        raise RuntimeError(f"omr_procedural_scene failed ({rc})")
    try:
        cam = Camera()
        C.memmove(C.byref(cam), C.byref(p.camera), C.sizeof(Camera))
        return _scene_from_desc(p.desc, cam)
    finally:
        lib.omr_procedural_free(C.byref(p))


def build_bvh(scene, width=4):
    """Host-only build of the wide BVH: (node bytes, leaf order, depth, stack
    bound) of its 4-wide (128-byte) or 8-wide (256-byte) nodes."""
    if width not in (4, 8):
        raise ValueError("width must be 4 or 8")
    lib = load()
    h = C.c_void_p()
    d = scene.desc()
    rc = lib.omr_bvh_build(C.byref(d), C.byref(h))
    if rc != OMR_OK:
        raise OmrError(rc, "omr_bvh_build failed")
    try:
        nodes, nn, order, no = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
        depth, sb = C.c_uint32(), C.c_uint32()
        lib.omr_bvh_data(h, C.byref(nodes), C.byref(nn), C.byref(order), C.byref(no), C.byref(depth), C.byref(sb))
        if width == 8:
            lib.omr_bvh_data8(h, C.byref(nodes), C.byref(nn), C.byref(depth), C.byref(sb))
        nb = np.frombuffer(C.string_at(nodes, nn.value * 32 * width), dtype=np.uint8).copy()
        ob = np.frombuffer(C.string_at(order, no.value * 4), dtype=np.uint32).copy()
        return nb, ob, depth.value, sb.value
    finally:
        lib.omr_bvh_free(h)


def procedural_camera(kind, width, height):
    cam = Camera()
    rc = load().omr_procedural_camera(kind, width, height, C.byref(cam))
    if rc != OMR_OK:
        raise OmrError(rc, "omr_procedural_camera failed")
    return cam


def camera_lookat(width, height, pos, direction, up=(0.0, 1.0, 0.0), screen_height=30.0, dist=40.0, focal=200.0,
                  aperture=0.0):
    lib = load()
    cam = Camera()
    P, D, U = (C.c_double * 3)(*pos), (C.c_double * 3)(*direction), (C.c_double * 3)(*up)
    lib.omr_camera_from_lookat(width, height, C.byref(P), C.byref(D), C.byref(U), screen_height, dist, focal,
                               aperture, C.byref(cam))
    return cam


def params(ss=1, spp_begin=0, spp_end=1, max_depth=8, rr_depth=3, env_is=0, seed=0x6F6D7233, nee=0):
    return RenderParams(ss, spp_begin, spp_end, max_depth, rr_depth, env_is, seed, nee, 0)


class OmrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Context:
    def __init__(self, device=0):
        self.lib = load()
        self.h = C.c_void_p()
        rc = self.lib.omr_create(C.byref(self.h), device)
        if rc != OMR_OK:
            raise OmrError(rc, "omr_create failed (no GPU?)")
        self._keep = {}

    def close(self):
        if self.h:
            self.lib.omr_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OMR_OK:
            raise OmrError(rc, self.lib.omr_last_error(self.h).decode())

    def device_info(self):
        n_cu, hbm, name = C.c_int(), C.c_uint64(), C.create_string_buffer(128)
        self._check(self.lib.omr_device_info(self.h, C.byref(n_cu), C.byref(hbm), name, 128))
        return {"n_cu": n_cu.value, "hbm_bytes": hbm.value, "name": name.value.decode()}

    def upload(self, scene):
        sid = C.c_int32()
        d = scene.desc()
        self._check(self.lib.omr_scene_upload(self.h, C.byref(d), C.byref(sid)))
        return sid.value

    def release(self, sid):
        self._check(self.lib.omr_scene_release(self.h, sid))

    def scene_info(self, sid):
        info = SceneInfo()
        self._check(self.lib.omr_scene_info_get(self.h, sid, C.byref(info)))
        return {k: getattr(info, k) for k, _ in SceneInfo._fields_}

    def scene_bvh(self, sid):
        nodes, nn, order, no = C.c_void_p(), C.c_size_t(), C.c_void_p(), C.c_size_t()
        self._check(self.lib.omr_scene_bvh(self.h, sid, C.byref(nodes), C.byref(nn), C.byref(order), C.byref(no)))
        nb = np.frombuffer(C.string_at(nodes, nn.value * 128), dtype=np.uint8).copy()
        ob = np.frombuffer(C.string_at(order, no.value * 4), dtype=np.uint32).copy()
        return nb, ob

    def env_tables(self, sid, w, h):
        row = np.zeros(h * (w + 1), dtype=np.uint64)
        marg = np.zeros(h + 1, dtype=np.uint64)
        self._check(self.lib.omr_scene_env_tables(self.h, sid, _ptr(row), _ptr(marg)))
        return row.reshape(h, w + 1), marg

    def trace(self, sid, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        self._check(self.lib.omr_trace_closest(self.h, sid, _ptr(rays), n, _ptr(prim), _ptr(t)))
        return prim, t

    def trace_device(self, sid, d_rays, n, d_prim, d_t, stats=False):
        if stats:
            nv, nt = C.c_uint64(), C.c_uint64()
            self._check(self.lib.omr_trace_stats_device(self.h, sid, d_rays, n, d_prim, d_t, C.byref(nv),
                                                        C.byref(nt)))
            return nv.value, nt.value
        self._check(self.lib.omr_trace_closest_device(self.h, sid, d_rays, n, d_prim, d_t))
        return None

    def trace_device_noalloc(self, sid, d_rays, n, d_prim, d_t):
        """Closest hit on device buffers through the allocation-free entry point."""
        self._check(self.lib.omr_trace_closest_device_noalloc(self.h, sid, d_rays, n, d_prim, d_t))

    def render(self, sid, cam, prm):
        out = np.empty((cam.height, cam.width, 3), dtype=np.float32)
        self._check(self.lib.omr_render(self.h, sid, C.byref(cam), C.byref(prm), _ptr(out)))
        return out

    def render_sums(self, sid, cam, prm):
        out = np.empty((cam.height, cam.width, 3), dtype=np.int64)
        self._check(self.lib.omr_render_sums(self.h, sid, C.byref(cam), C.byref(prm), _ptr(out)))
        return out

    def render_batch(self, sids, cams, prms):
        n = len(sids)
        outs = [np.empty((c.height, c.width, 3), dtype=np.float32) for c in cams]
        s_arr = (C.c_int32 * n)(*sids)
        c_arr = (Camera * n)(*cams)
        p_arr = (RenderParams * n)(*prms)
        o_arr = (C.c_void_p * n)(*[o.ctypes.data for o in outs])
        self._check(self.lib.omr_render_batch(self.h, n, s_arr, c_arr, p_arr, o_arr))
        return outs

    def last_stats(self):
        st = RenderStats()
        self._check(self.lib.omr_last_render_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in RenderStats._fields_}

    def set_profiling(self, on=True):
        self._check(self.lib.omr_set_profiling(self.h, 1 if on else 0))

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.omr_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def autotune(self, sids, cams, prms, budget_ms=20000):
        n = len(sids)
        t = Tuning()
        self._check(self.lib.omr_autotune(self.h, n, (C.c_int32 * n)(*sids), (Camera * n)(*cams),
                                          (RenderParams * n)(*prms), budget_ms, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        t = Tuning(**cur)
        self._check(self.lib.omr_set_tuning(self.h, C.byref(t)))

    # ---- served mode ----
    def server_start(self, pool_paths=0, hbm_budget_bytes=0, max_live_jobs=0):
        cfg = ServerConfig(pool_paths, hbm_budget_bytes, max_live_jobs)
        self._check(self.lib.omr_server_start(self.h, C.byref(cfg)))

    def server_stop(self):
        self._check(self.lib.omr_server_stop(self.h))

    def submit(self, sid, cam, prm, out):
        job = C.c_int64()
        self._keep[id(out)] = out
        rc = self.lib.omr_job_submit(self.h, sid, C.byref(cam), C.byref(prm), _ptr(out), C.byref(job))
        self._check(rc)
        return job.value

    def poll(self, job):
        st = C.c_int()
        self._check(self.lib.omr_job_poll(self.h, job, C.byref(st)))
        return st.value

    def wait(self, job, timeout_ms=-1):
        return self.lib.omr_job_wait(self.h, job, timeout_ms)

    def job_error(self, job):
        buf = C.create_string_buffer(512)
        self._check(self.lib.omr_job_error(self.h, job, buf, 512))
        return buf.value.decode()

    def job_latency(self, job):
        ms = C.c_double()
        self._check(self.lib.omr_job_latency(self.h, job, C.byref(ms)))
        return ms.value

    def cancel(self, job):
        return self.lib.omr_job_cancel(self.h, job)

    def server_stats(self):
        st = ServerStats()
        self._check(self.lib.omr_server_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
