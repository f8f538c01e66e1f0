"""Seeded ray sets and small utilities shared by the tests."""
import numpy as np

INF = float.fromhex("0x1.fffffep+126")


def make_rays(o, d, tmax=INF):
    n = o.shape[0]
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = tmax
    r[:, 4:7] = d
    return r


def unit_dirs(rng, n):
    z = 1.0 - 2.0 * rng.random(n)
    phi = 2.0 * np.pi * rng.random(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1).astype(np.float32)


def scene_bounds(scene):
    pts = [scene.tri[:, 0:3], scene.tri[:, 0:3] + scene.tri[:, 3:6], scene.tri[:, 0:3] + scene.tri[:, 6:9]]
    if scene.sph.shape[0]:
        pts.append(scene.sph[:, 0:3] - scene.sph[:, 3:4])
        pts.append(scene.sph[:, 0:3] + scene.sph[:, 3:4])
    p = np.concatenate(pts)
    return p.min(0), p.max(0)


def hit_rays(scene, n, seed, box_scale=1.5, focus=None):
    """Half the rays start uniformly in box_scale x the bounds (or `focus` =
    (lo, hi)) with uniform directions; half are aimed at points on random
    triangles from outside."""
    rng = np.random.default_rng(seed)
    lo, hi = focus if focus is not None else scene_bounds(scene)
    c, e = (lo + hi) / 2, (hi - lo) / 2 * box_scale
    n1 = n // 2
    o1 = (c + e * (2 * rng.random((n1, 3)) - 1)).astype(np.float32)
    d1 = unit_dirs(rng, n1)
    n2 = n - n1
    if scene.tri.shape[0]:
        ti = rng.integers(0, scene.tri.shape[0], n2)
        a, b = rng.random(n2), rng.random(n2)
        flip = a + b > 1
        a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
        tgt = scene.tri[ti, 0:3] + scene.tri[ti, 3:6] * a[:, None] + scene.tri[ti, 6:9] * b[:, None]
        o2 = (c + e * 1.2 * unit_dirs(rng, n2)).astype(np.float32)
        d2 = tgt - o2
        d2 = (d2 / np.linalg.norm(d2, axis=1, keepdims=True)).astype(np.float32)
    else:
        o2 = (c + e * (2 * rng.random((n2, 3)) - 1)).astype(np.float32)
        d2 = unit_dirs(rng, n2)
    return make_rays(np.concatenate([o1, o2]), np.concatenate([d1, d2]))


class DeviceArray:
    """A device copy of a numpy array through the HIP runtime libomr.so links
    (libamdhip64.so.7).  Tests use it for the entry points that take device
    pointers; torch bundles a HIP runtime of its own, which cannot share the
    device with the one already loaded."""

    _hip = None

    @classmethod
    def hip(cls):
        import ctypes as C

        if cls._hip is None:
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, a):
        import ctypes as C

        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        h = self.hip()
        if h.hipMalloc(C.byref(self.ptr), max(1, self.a.nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        if self.a.nbytes and h.hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    @property
    def address(self):
        return self.ptr.value

    def get(self):
        out = np.empty_like(self.a)
        self.hip().hipDeviceSynchronize()
        if out.nbytes and self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy to the host failed")
        return out

    def free(self):
        if self.ptr.value:
            self.hip().hipFree(self.ptr)
            self.ptr = type(self.ptr)()
