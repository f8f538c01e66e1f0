"""ctypes binding of the CPU restatement (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker / CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")
BRUTE, OWN_BVH, WIDE_BVH = 0, 1, 2


class OScene(C.Structure):
    _fields_ = [("n_tri", C.c_size_t), ("tri", C.c_void_p), ("tri_n", C.c_void_p), ("n_sph", C.c_size_t),
                ("sph", C.c_void_p), ("prim_mat", C.c_void_p), ("prim_mat_count", C.c_void_p), ("n_mat", C.c_size_t),
                ("mat", C.c_void_p), ("env_w", C.c_uint32), ("env_h", C.c_uint32), ("env", C.c_void_p),
                ("n_pln", C.c_size_t), ("pln", C.c_void_p), ("n_tex", C.c_size_t), ("tex", C.c_void_p),
                ("tri_uv", C.c_void_p)]


class OTexture(C.Structure):
    _fields_ = [("w", C.c_uint32), ("h", C.c_uint32), ("rgb", C.c_void_p)]


class OCamera(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("position", C.c_float * 3),
                ("screen_center", C.c_float * 3), ("screen_x", C.c_float * 3), ("screen_y", C.c_float * 3),
                ("aperture_radius", C.c_float), ("focal_distance", C.c_float)]


class OParams(C.Structure):
    _fields_ = [("supersamples", C.c_uint32), ("spp_begin", C.c_uint32), ("spp_end", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("env_is", C.c_uint32), ("seed", C.c_uint64),
                ("nee", C.c_uint32), ("pad", C.c_uint32)]


_lib = None


def load(build_if_missing=True):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH) and build_if_missing:
        subprocess.run(["make", "-C", HERE, "all"], check=True, capture_output=True)
    lib = C.CDLL(LIB_PATH)
    vp, sz, u32 = C.c_void_p, C.c_size_t, C.c_uint32
    lib.orc_rng_init.argtypes = [u32, vp]
    lib.orc_rng_init.restype = None
    lib.orc_rng_next.argtypes = [vp]
    lib.orc_rng_next.restype = u32
    lib.orc_rng_float.argtypes = [vp]
    lib.orc_rng_float.restype = C.c_float
    lib.orc_path_seed.argtypes = [C.c_uint64, u32, u32, u32, u32]
    lib.orc_path_seed.restype = u32
    lib.orc_sincos_turns.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_sincos_turns.restype = None
    lib.orc_acos.argtypes = [C.c_float]
    lib.orc_acos.restype = C.c_float
    lib.orc_intersect_triangle.argtypes = [vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_intersect_sphere.argtypes = [vp, vp, C.POINTER(C.c_float)]
    lib.orc_intersect_plane.argtypes = [vp, C.c_float, C.POINTER(C.c_float)]
    lib.orc_env_texel.argtypes = [C.c_float, C.c_float, C.c_float, u32, u32, C.POINTER(u32), C.POINTER(u32)]
    lib.orc_env_texel.restype = None
    lib.orc_sphere_uv.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.orc_sphere_uv.restype = None
    lib.orc_texel_xy.argtypes = [C.c_float, C.c_float, u32, u32, C.POINTER(u32), C.POINTER(u32)]
    lib.orc_texel_xy.restype = None
    lib.orc_camera_ray.argtypes = [C.POINTER(OCamera), C.c_float, C.c_float, vp, vp, vp]
    lib.orc_camera_ray.restype = None
    lib.orc_prepare.argtypes = [C.POINTER(OScene)]
    lib.orc_prepare.restype = vp
    lib.orc_free.argtypes = [vp]
    lib.orc_free.restype = None
    lib.orc_set_wide_bvh.argtypes = [vp, vp, sz, vp, sz]
    lib.orc_trace.argtypes = [vp, vp, sz, C.c_int, vp, vp, vp, C.c_int]
    lib.orc_trace.restype = None
    lib.orc_render.argtypes = [vp, C.POINTER(OCamera), C.POINTER(OParams), C.c_int, C.c_int, u32, u32, vp, vp, vp]
    lib.orc_render_sums.argtypes = [vp, C.POINTER(OCamera), C.POINTER(OParams), C.c_int, C.c_int, u32, u32, vp, vp,
                                    vp]
    lib.orc_env_cdf.argtypes = [vp, vp, vp]
    lib.orc_env_tables_raw.argtypes = [vp, vp, vp]
    lib.orc_env_sample.argtypes = [vp, C.c_float, C.c_float, vp, C.POINTER(C.c_float), C.POINTER(u32),
                                   C.POINTER(u32)]
    lib.orc_env_sample.restype = None
    lib.orc_env_pdf.argtypes = [vp, C.c_float, C.c_float, C.c_float]
    lib.orc_env_pdf.restype = C.c_float
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def to_ocamera(cam):
    oc = OCamera()
    C.memmove(C.byref(oc), C.byref(cam), C.sizeof(OCamera))
    return oc


class Oracle:
    """Holds a prepared scene (the oracle's own BVH + optional wide BVH)."""

    def __init__(self, scene, wide_bvh=None):
        self.lib = load()
        self.scene = scene
        s = OScene()
        s.n_tri = scene.tri.shape[0]
        s.tri = _p(scene.tri) if s.n_tri else None
        if scene.tri_n is None and s.n_tri:
            raise ValueError("the oracle needs explicit triangle normals")
        s.tri_n = _p(scene.tri_n) if s.n_tri else None
        s.n_sph = scene.sph.shape[0]
        s.sph = _p(scene.sph) if s.n_sph else None
        s.prim_mat = _p(scene.prim_mat)
        s.prim_mat_count = _p(scene.prim_mat_count)
        s.n_mat = scene.mats.shape[0]
        s.mat = _p(scene.mats)
        if scene.env is not None:
            s.env_h, s.env_w = scene.env.shape[0], scene.env.shape[1]
            s.env = _p(scene.env)
        planes = getattr(scene, "planes", None)
        if planes is not None and planes.shape[0]:
            s.n_pln = planes.shape[0]
            s.pln = _p(planes)
        textures = getattr(scene, "textures", None) or []
        if textures:
            self._tex = (OTexture * len(textures))()
            for k, t in enumerate(textures):
                self._tex[k].h, self._tex[k].w = t.shape[0], t.shape[1]
                self._tex[k].rgb = _p(t)
            s.n_tex = len(textures)
            s.tex = C.cast(self._tex, C.c_void_p)
        tri_uv = getattr(scene, "tri_uv", None)
        if tri_uv is not None and s.n_tri:
            s.tri_uv = _p(tri_uv)
        self._s = s
        self.h = self.lib.orc_prepare(C.byref(s))
        self._wide = None
        if wide_bvh is not None:
            self.set_wide(*wide_bvh)

    def set_wide(self, nodes, order):
        self._wide = (np.ascontiguousarray(nodes), np.ascontiguousarray(order, dtype=np.uint32))
        self.lib.orc_set_wide_bvh(self.h, _p(self._wide[0]), self._wide[0].size // 128, _p(self._wide[1]),
                                  self._wide[1].size)

    def close(self):
        if self.h:
            self.lib.orc_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace(self, rays, accel=OWN_BVH, threads=None, counters=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.empty(n, dtype=np.int32)
        t = np.empty(n, dtype=np.float32)
        cnt = np.zeros(2, dtype=np.uint64)
        self.lib.orc_trace(self.h, _p(rays), n, accel, _p(prim), _p(t), _p(cnt), threads or os.cpu_count())
        if counters:
            return prim, t, cnt
        return prim, t

    def render_sums(self, cam, prm, accel=OWN_BVH, threads=None, rows=None):
        oc = to_ocamera(cam)
        op = OParams()
        C.memmove(C.byref(op), C.byref(prm), C.sizeof(OParams))
        r0, r1 = rows if rows else (0, cam.height)
        out = np.empty((r1 - r0, cam.width, 3), dtype=np.int64)
        seg, sh = C.c_uint64(), C.c_uint64()
        rc = self.lib.orc_render_sums(self.h, C.byref(oc), C.byref(op), accel, threads or os.cpu_count(), r0, r1,
                                      _p(out), C.byref(seg), C.byref(sh))
        if rc != 0:
            raise RuntimeError(f"orc_render_sums failed ({rc})")
        return out, seg.value, sh.value

    def render(self, cam, prm, accel=OWN_BVH, threads=None, rows=None):
        sums, seg, sh = self.render_sums(cam, prm, accel, threads, rows)
        n = float(prm.supersamples * prm.supersamples * (prm.spp_end - prm.spp_begin))
        return (sums.astype(np.float64) * (1.0 / 16777216.0) / n).astype(np.float32), seg, sh

    def env_tables(self):
        h, w = self.scene.env.shape[0], self.scene.env.shape[1]
        row = np.zeros(h * (w + 1), dtype=np.uint64)
        marg = np.zeros(h + 1, dtype=np.uint64)
        if self.lib.orc_env_tables_raw(self.h, _p(row), _p(marg)) != 0:
            raise RuntimeError("scene has no environment")
        return row.reshape(h, w + 1), marg

    def env_pdf(self, d):
        return self.lib.orc_env_pdf(self.h, float(d[0]), float(d[1]), float(d[2]))


def rng_floats(seed, n):
    lib = load()
    st = (C.c_uint32 * 4)()
    lib.orc_rng_init(seed, st)
    return np.array([lib.orc_rng_float(st) for _ in range(n)], dtype=np.float32)


def rng_next(seed, n):
    lib = load()
    st = (C.c_uint32 * 4)()
    lib.orc_rng_init(seed, st)
    return np.array([lib.orc_rng_next(st) for _ in range(n)], dtype=np.uint32)
