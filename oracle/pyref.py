"""ctypes binding of oracle/_ref/libref.so: the reference's own code compiled
from its sources (see oracle/Makefile and oracle/ref_harness.cpp).

Only the fixture script under tests/golden/ uses this, in a container that has
the reference checkout; nothing on the GPU host loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref.so")


def load():
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "ref"], check=True, capture_output=True)
    lib = C.CDLL(LIB_PATH)
    vp, sz = C.c_void_p, C.c_size_t
    lib.ref_random.argtypes = [C.c_uint32, sz, vp, vp]
    lib.ref_random.restype = None
    lib.ref_scene_new.argtypes = [vp, vp, sz, vp, sz, vp, vp, vp, C.c_int, C.c_char_p]
    lib.ref_scene_new.restype = vp
    lib.ref_scene_free.argtypes = [vp]
    lib.ref_scene_free.restype = None
    lib.ref_trace.argtypes = [vp, vp, sz, vp, vp]
    lib.ref_trace.restype = None
    lib.ref_env_texel.argtypes = [C.c_char_p, sz, vp, vp, vp, vp]
    lib.ref_render.argtypes = [vp, C.c_uint32, C.c_uint32, vp, vp, vp, C.c_double, C.c_double, C.c_uint32,
                               C.c_uint32, vp, vp]
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def random(seed, n):
    lib = load()
    nx = np.empty(n, dtype=np.uint32)
    nd = np.empty(n, dtype=np.float64)
    lib.ref_random(seed, n, _p(nx), _p(nd))
    return nx, nd


class RefScene:
    def __init__(self, scene, bvh=False, hdr=None):
        self.lib = load()
        self.scene = scene
        self.h = self.lib.ref_scene_new(_p(scene.tri), _p(scene.tri_n), scene.tri.shape[0], _p(scene.sph),
                                        scene.sph.shape[0], _p(scene.prim_mat), _p(scene.prim_mat_count),
                                        _p(scene.mats), 1 if bvh else 0, (hdr or "").encode())

    def close(self):
        if self.h:
            self.lib.ref_scene_free(self.h)
            self.h = None

    def trace(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        prim = np.empty(rays.shape[0], dtype=np.int32)
        t = np.empty(rays.shape[0], dtype=np.float64)
        self.lib.ref_trace(self.h, _p(rays), rays.shape[0], _p(prim), _p(t))
        return prim, t

    def render(self, w, h, pos, direction, up, screen_h, dist, ss, batches, threads=1):
        return _render(self.lib, self.lib.ref_render, self.h, w, h, pos, direction, up, screen_h, dist, ss, batches,
                       threads)


def _render(lib, fn, handle, w, h, pos, direction, up, screen_h, dist, ss, batches, threads):
    """PathTracer::RenderScene through a harness entry: (mean, var) of the
    batch means; lib.ref_last_seconds() is its wall time afterwards."""
    lib.ref_set_threads.argtypes = [C.c_size_t]
    lib.ref_set_threads.restype = C.c_int
    lib.ref_last_seconds.argtypes = []
    lib.ref_last_seconds.restype = C.c_double
    if lib.ref_set_threads(threads) != 0:
        raise ValueError("the reference's row bands support 1 or 2 threads")
    mean = np.empty((h, w, 3), dtype=np.float64)
    var = np.empty((h, w, 3), dtype=np.float64)
    P, D, U = (np.asarray(v, dtype=np.float64) for v in (pos, direction, up))
    try:
        fn(handle, w, h, _p(P), _p(D), _p(U), screen_h, dist, ss, batches, _p(mean), _p(var))
    finally:
        lib.ref_set_threads(1)
    return mean, var


def last_seconds():
    """Wall time of the last reference render."""
    lib = load()
    lib.ref_last_seconds.restype = C.c_double
    return float(lib.ref_last_seconds())


def env_texels(hdr, dirs):
    lib = load()
    dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
    n = dirs.shape[0]
    ix = np.empty(n, dtype=np.int32)
    iy = np.empty(n, dtype=np.int32)
    wh = np.empty(2, dtype=np.int32)
    rc = lib.ref_env_texel(hdr.encode(), n, _p(dirs), _p(ix), _p(iy), _p(wh))
    if rc != 0:
        raise RuntimeError("HDR read failed")
    return ix, iy, int(wh[0]), int(wh[1])


def _load_io(lib):
    vp, sz = C.c_void_p, C.c_size_t
    lib.ref_scene_add_planes.argtypes = [vp, vp, sz, vp, vp, vp, sz]
    lib.ref_scene_add_planes.restype = None
    lib.ref_file_scene_new.argtypes = [C.c_char_p]
    lib.ref_file_scene_new.restype = vp
    lib.ref_file_scene_free.argtypes = [vp]
    lib.ref_file_scene_free.restype = None
    lib.ref_file_scene_objects.argtypes = [vp]
    lib.ref_file_scene_objects.restype = sz
    lib.ref_file_trace.argtypes = [vp, vp, sz, vp, vp, vp, vp, vp, vp]
    lib.ref_file_trace.restype = None
    lib.ref_hdr_read.argtypes = [C.c_char_p, vp, vp]
    lib.ref_hdr_read.restype = C.c_int
    lib.ref_hdr_write.argtypes = [C.c_char_p, C.c_int32, C.c_int32, vp]
    lib.ref_hdr_write.restype = C.c_int
    lib.ref_image_load.argtypes = [C.c_char_p, vp, vp]
    lib.ref_image_load.restype = C.c_int
    lib.ref_texture_load.argtypes = [C.c_char_p, C.c_int, vp, vp]
    lib.ref_texture_load.restype = C.c_int
    lib.ref_sphere_uv.argtypes = [vp, sz, vp]
    lib.ref_sphere_uv.restype = None
    lib.ref_texel_xy.argtypes = [C.c_int32, C.c_int32, vp, sz, vp]
    lib.ref_texel_xy.restype = None
    return lib


def add_planes(ref_scene, scene):
    """Appends scene.planes as InfiniteFloor objects (after every other
    object, outside the BVH) to a RefScene built from the same scene."""
    lib = _load_io(ref_scene.lib)
    y = np.ascontiguousarray(scene.planes, dtype=np.float32)
    first = scene.tri.shape[0] + scene.sph.shape[0]
    lib.ref_scene_add_planes(ref_scene.h, _p(y), y.shape[0], _p(scene.prim_mat), _p(scene.prim_mat_count),
                             _p(scene.mats), first)


class RefFileScene:
    """SceneFromExternalFile on a .scene file (relative names inside it
    resolve against the current directory)."""

    def __init__(self, path):
        self.lib = _load_io(load())
        self.h = self.lib.ref_file_scene_new(path.encode())
        if not self.h:
            raise RuntimeError(f"reference could not load {path}")

    def objects(self):
        return int(self.lib.ref_file_scene_objects(self.h))

    def close(self):
        if self.h:
            self.lib.ref_file_scene_free(self.h)
            self.h = None

    def render(self, w, h, pos, direction, up, screen_h, dist, ss, batches, threads=1):
        fn = self.lib.ref_file_render
        fn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                       C.c_double, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
        fn.restype = C.c_int
        return _render(self.lib, fn, self.h, w, h, pos, direction, up, screen_h, dist, ss, batches, threads)

    def trace(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        nrm = np.zeros((n, 3), np.float64)
        mat = np.zeros((n, 4, 9), np.float32)
        nmat = np.zeros(n, np.int32)
        obj = np.empty(n, np.int64)
        uv = np.zeros((n, 2), np.float64)
        self.lib.ref_file_trace(self.h, _p(rays), n, _p(t), _p(nrm), _p(mat), _p(nmat), _p(obj), _p(uv))
        return dict(t=t, normal=nrm, mat=mat, nmat=nmat, obj=obj, uv=uv)


def hdr_read(path):
    lib = _load_io(load())
    wh = np.zeros(2, np.int32)
    if lib.ref_hdr_read(path.encode(), _p(wh), None) != 0:
        raise RuntimeError(f"reference could not read {path}")
    out = np.empty((wh[1], wh[0], 3), np.float32)
    lib.ref_hdr_read(path.encode(), _p(wh), _p(out))
    return out


def hdr_write(path, rgb):
    lib = _load_io(load())
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[0], rgb.shape[1]
    if lib.ref_hdr_write(path.encode(), w, h, _p(rgb)) != 0:
        raise RuntimeError(f"reference could not write {path}")


def image_load(path):
    """stbi_load as the reference calls it: uint8 (h, w, channels)."""
    lib = _load_io(load())
    whc = np.zeros(3, np.int32)
    if lib.ref_image_load(path.encode(), _p(whc), None) != 0:
        raise RuntimeError(f"reference could not read {path}")
    out = np.empty((whc[1], whc[0], whc[2]), np.uint8)
    lib.ref_image_load(path.encode(), _p(whc), _p(out))
    return out


def texture_load(path, inverse_gamma=True):
    """ImageHandler::LoadFromFile's pixels as float32 (h, w, 3)."""
    lib = _load_io(load())
    wh = np.zeros(2, np.int32)
    if lib.ref_texture_load(path.encode(), int(inverse_gamma), _p(wh), None) != 0:
        raise RuntimeError(f"reference could not read {path}")
    out = np.empty((wh[1], wh[0], 3), np.float32)
    lib.ref_texture_load(path.encode(), int(inverse_gamma), _p(wh), _p(out))
    return out


def sphere_uv(normals):
    lib = _load_io(load())
    nrm = np.ascontiguousarray(normals, np.float64).reshape(-1, 3)
    uv = np.empty((nrm.shape[0], 2), np.float64)
    lib.ref_sphere_uv(_p(nrm), nrm.shape[0], _p(uv))
    return uv


def texel_xy(w, h, uv):
    lib = _load_io(load())
    uv = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    xy = np.empty((uv.shape[0], 2), np.int32)
    lib.ref_texel_xy(w, h, _p(uv), uv.shape[0], _p(xy))
    return xy


class RefCornell:
    """The reference's own CornellBoxScene (src/scenes/CornellBoxScene.cpp:10-36)."""

    def __init__(self):
        self.lib = load()
        vp, sz, u32 = C.c_void_p, C.c_size_t, C.c_uint32
        self.lib.ref_cornell_new.argtypes = []
        self.lib.ref_cornell_new.restype = vp
        self.lib.ref_cornell_free.argtypes = [vp]
        self.lib.ref_cornell_free.restype = None
        self.lib.ref_cornell_objects.argtypes = [vp]
        self.lib.ref_cornell_objects.restype = sz
        self.lib.ref_cornell_trace.argtypes = [vp, vp, sz, vp, vp, vp, vp, vp, vp]
        self.lib.ref_cornell_trace.restype = None
        self.lib.ref_cornell_render.argtypes = [vp, u32, u32, vp, vp, vp, C.c_double, C.c_double, u32, u32, vp, vp,
                                                C.c_int]
        self.h = self.lib.ref_cornell_new()

    def objects(self):
        return int(self.lib.ref_cornell_objects(self.h))

    def close(self):
        if self.h:
            self.lib.ref_cornell_free(self.h)
            self.h = None

    def trace(self, rays):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        nrm = np.zeros((n, 3), np.float64)
        mat = np.zeros((n, 4, 9), np.float32)
        nmat = np.zeros(n, np.int32)
        obj = np.empty(n, np.int64)
        uv = np.zeros((n, 2), np.float64)
        self.lib.ref_cornell_trace(self.h, _p(rays), n, _p(t), _p(nrm), _p(mat), _p(nmat), _p(obj), _p(uv))
        return dict(t=t, normal=nrm, mat=mat, nmat=nmat, obj=obj, uv=uv)

    def render(self, w, h, pos, direction, up, screen_h, dist, ss, batches, nee=False):
        mean = np.empty((h, w, 3), dtype=np.float64)
        var = np.empty((h, w, 3), dtype=np.float64)
        P, D, U = (np.asarray(v, dtype=np.float64) for v in (pos, direction, up))
        self.lib.ref_cornell_render(self.h, w, h, _p(P), _p(D), _p(U), screen_h, dist, ss, batches, _p(mean), _p(var),
                                    1 if nee else 0)
        return mean, var
