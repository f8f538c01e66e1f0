"""Closed-form checks of the CPU restatement's building blocks."""
import ctypes as C

import numpy as np
import pytest

import helpers


def _lib():
    import pyoracle

    return pyoracle.load()


def test_sincos_turns_closed_form():
    lib = _lib()
    xs = np.concatenate([np.linspace(0, 1, 4097), np.random.default_rng(1).random(2000)]).astype(np.float32)
    s, c = C.c_float(), C.c_float()
    err = 0.0
    for x in xs:
        lib.orc_sincos_turns(float(x), C.byref(s), C.byref(c))
        a = 2 * np.pi * float(x)
        err = max(err, abs(s.value - np.sin(a)), abs(c.value - np.cos(a)))
        assert abs(s.value * s.value + c.value * c.value - 1.0) < 1e-6
    assert err < 3e-7


def test_acos_closed_form():
    lib = _lib()
    xs = np.concatenate([np.linspace(-1, 1, 8193), [-1.0, 1.0, 0.5, -0.5, 0.0]]).astype(np.float32)
    err = max(abs(lib.orc_acos(float(x)) - np.arccos(np.float64(x))) for x in xs)
    assert err < 5e-7
    assert lib.orc_acos(2.0) == pytest.approx(0.0) and lib.orc_acos(-2.0) == pytest.approx(np.pi)


def _tri(ray, tri):
    lib = _lib()
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    r = np.asarray(ray, np.float32)
    tr = np.asarray(tri, np.float32)
    ok = lib.orc_intersect_triangle(r.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p), C.byref(t),
                                    C.byref(u), C.byref(v))
    return ok, t.value, u.value, v.value


def test_triangle_single_sided_closed_form():
    """Polygon::CheckIntersection: det > EPS (front face only), barycentric
    bounds, t >= EPS (src/renderer/Polygon.h:63-106)."""
    inf = helpers.INF
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]          # normal e1 x e2 = +z
    ok, t, u, v = _tri([0.25, 0.25, 5, inf, 0, 0, -1, 0], tri)   # from +z side
    assert ok and t == pytest.approx(5.0) and u == pytest.approx(0.25) and v == pytest.approx(0.25)
    assert not _tri([0.25, 0.25, -5, inf, 0, 0, 1, 0], tri)[0]    # back face passes through
    assert not _tri([0.9, 0.9, 5, inf, 0, 0, -1, 0], tri)[0]      # u + v > 1
    assert not _tri([0.25, 0.25, 0.0001, inf, 0, 0, -1, 0], tri)[0]  # t < EPS
    assert _tri([0.25, 0.25, 0.0006, inf, 0, 0, -1, 0], tri)[0]
    small = [0, 0, 0, 0.01, 0, 0, 0, 0.01, 0]  # det = 1e-4 <= EPS: culled as in the reference
    assert not _tri([0.002, 0.002, 5, inf, 0, 0, -1, 0], small)[0]


def test_sphere_closed_form():
    lib = _lib()
    sph = np.array([0, 0, 0, 2], np.float32)

    def hit(o, d):
        t = C.c_float()
        r = np.array(list(o) + [helpers.INF] + list(d) + [0], np.float32)
        ok = lib.orc_intersect_sphere(r.ctypes.data_as(C.c_void_p), sph.ctypes.data_as(C.c_void_p), C.byref(t))
        return ok, t.value

    assert hit((0, 0, 10), (0, 0, -1)) == (1, pytest.approx(8.0))        # near root
    assert hit((0, 0, 0), (1, 0, 0)) == (1, pytest.approx(2.0))          # inside: far root
    assert hit((0, 0, 10), (0, 0, 1))[0] == 0                            # behind
    assert hit((0, 3, 10), (0, 0, -1))[0] == 0                           # misses
    ok, t = hit((0, 1.0, 10), (0, 0, -1))
    assert ok and t == pytest.approx(10 - np.sqrt(3.0), abs=1e-5)


@pytest.mark.parametrize("kind", [0, 1])
def test_traversals_agree_with_brute_force(kind):
    """Own binary BVH and the product's 4-wide BVH (traversed by the oracle)
    return exactly the brute-force closest hit."""
    import omr
    import pyoracle

    scene = omr.procedural(kind, width=32, height=32)
    nodes, order, _, _ = omr.build_bvh(scene)
    o = pyoracle.Oracle(scene, (nodes, order))
    rays = helpers.hit_rays(scene, 3000 if kind else 20000, seed=5)
    pb, tb = o.trace(rays, pyoracle.BRUTE)
    for accel in (pyoracle.OWN_BVH, pyoracle.WIDE_BVH):
        p, t = o.trace(rays, accel)
        assert np.array_equal(p, pb)
        assert np.array_equal(t.view(np.uint32), tb.view(np.uint32))


def _furnace_scene(albedo, env_value, floor):
    import omr

    mats = np.zeros(1, dtype=omr.MATERIAL_DTYPE)
    mats["type"] = omr.LAMBERT
    mats["color"] = albedo
    mats["rate"] = 1.0
    env = np.full((32, 64, 3), env_value, np.float32)
    if floor:
        s = 1000.0
        tri = np.array([[-s, 0, -s, 0, 0, 2 * s, 2 * s, 0, 0], [s, 0, -s, -2 * s, 0, 2 * s, 0, 0, 2 * s]], np.float32)
        trn = np.zeros((2, 9), np.float32)
        trn[:, 1] = 1.0
        return omr.Scene(tri, trn, np.zeros((0, 4)), np.zeros(2, np.uint32), mats, env=env)
    return omr.Scene(np.zeros((0, 9)), None, np.array([[0, 0, 0, 1.0]]), np.zeros(1, np.uint32), mats, env=env)


def test_white_furnace_convex_sphere_exact():
    """A convex Lambert sphere of albedo a under a constant environment L:
    every bounce escapes, so each path returns exactly a*L (closed form)."""
    import omr
    import pyoracle

    a, L = 0.6, 2.0
    scene = _furnace_scene((a, a, a), L, floor=False)
    cam = omr.camera_lookat(16, 16, (0, 0, 6), (0, 0, -1), screen_height=2.0, dist=4.0)
    prm = omr.params(ss=1, spp_end=4, max_depth=8, rr_depth=8)
    img, seg, _ = pyoracle.Oracle(scene).render(cam, prm, pyoracle.BRUTE)
    on = np.abs(img - a * L) < 1e-5
    off = np.abs(img - L) < 1e-6
    assert (on | off).all()
    assert on[8, 8].all() and off[0, 0].all()


def test_env_importance_sampling_unbiased_on_a_floor():
    """Environment importance sampling with MIS against the lookup-only
    estimator on an axis-aligned Lambert floor (where the reference's
    unnormalized basis is exact): both converge to a*L."""
    import omr
    import pyoracle

    a, L = 0.5, 1.5
    scene = _furnace_scene((a, a, a), L, floor=True)
    cam = omr.camera_lookat(8, 8, (0, 10, 0), (0.0001, -1, 0), screen_height=2.0, dist=4.0)
    o = pyoracle.Oracle(scene)
    for env_is in (0, 1):
        prm = omr.params(ss=1, spp_end=256, max_depth=4, rr_depth=8, env_is=env_is)
        img, seg, sh = o.render(cam, prm, pyoracle.BRUTE)
        assert abs(img.mean() - a * L) < 0.01 * a * L
        assert (sh > 0) == bool(env_is)


def test_env_cdf_chi_square():
    """Texel frequencies of the sampler follow the tabulated probabilities."""
    import omr
    import pyoracle

    rng = np.random.default_rng(3)
    env = (rng.random((16, 32, 3)) ** 4 * 10).astype(np.float32)
    mats = np.zeros(1, dtype=omr.MATERIAL_DTYPE)
    mats["rate"] = 1
    scene = omr.Scene(np.zeros((0, 9)), None, np.array([[0, 0, 0, 1.0]]), np.zeros(1, np.uint32), mats, env=env)
    o = pyoracle.Oracle(scene)
    lib = o.lib
    marg = np.zeros(17, np.float32)
    cond = np.zeros(16 * 33, np.float32)
    lib.orc_env_cdf(o.h, marg.ctypes.data_as(C.c_void_p), cond.ctypes.data_as(C.c_void_p))
    cond = cond.reshape(16, 33)
    p = np.diff(marg)[:, None] * np.diff(cond, axis=1)
    n = 200000
    counts = np.zeros((16, 32))
    d = np.zeros(3, np.float32)
    pdf, ix, iy = C.c_float(), C.c_uint32(), C.c_uint32()
    us = rng.random((n, 2)).astype(np.float32)
    for u1, u2 in us:
        lib.orc_env_sample(o.h, float(u1), float(u2), d.ctypes.data_as(C.c_void_p), C.byref(pdf), C.byref(ix),
                           C.byref(iy))
        counts[iy.value, ix.value] += 1
    exp = p * n
    m = exp > 5
    chi2 = (((counts - exp) ** 2)[m] / exp[m]).sum()
    dof = m.sum() - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof)
    # the solid-angle pdf integrates to 1 over the sphere
    dirs = helpers.unit_dirs(rng, 200000)
    pdfs = np.array([o.env_pdf(v) for v in dirs[:20000]])
    assert abs(pdfs.mean() * 4 * np.pi - 1.0) < 0.05


def _seed_fn():
    import pyoracle

    lib = pyoracle.load()
    return lambda *a: int(lib.orc_path_seed(*a))


def test_path_seed_depends_on_every_argument():
    """Per-path seeding hash(seed, pixel, sub-pixel, sample, branch): changing
    any one argument (the 64-bit seed's low or high word included) changes
    the stream, and swapping two arguments does too -- a hash that dropped or
    transposed an argument would reuse streams across samples."""
    seed_of = _seed_fn()
    rng = np.random.default_rng(5)
    for _ in range(2000):
        s = int(rng.integers(0, 2 ** 63))
        p, u, k = (int(x) for x in rng.integers(0, 1 << 20, 3))
        b = int(rng.integers(0, 8))
        h = seed_of(s, p, u, k, b)
        for alt in ((s ^ 1, p, u, k, b), (s ^ (1 << 40), p, u, k, b), (s, p + 1, u, k, b), (s, p, u + 1, k, b),
                    (s, p, u, k + 1, b), (s, p, u, k, b ^ 1), (s, u, p, k, b), (s, p, k, u, b), (s, k, u, p, b)):
            if alt[1:] != (p, u, k, b):
                assert seed_of(*alt) != h, (s, p, u, k, b, alt)
    # the lens stream (branch ~0) differs from every path stream of the sample
    assert seed_of(7, 3, 2, 1, 0xFFFFFFFF) not in {seed_of(7, 3, 2, 1, b) for b in range(64)}


def test_path_seed_streams_are_distinct_and_decorrelated():
    """Over a grid of 32 pixels x 16 sub-pixels x 4 samples x 4 branches no
    two paths share a seed; one input bit flips half the output bits on
    average; the first draws of neighbouring pixels' streams are uniform on
    [0, 1] and uncorrelated."""
    import itertools

    import pyoracle

    seed_of = _seed_fn()
    grid = [seed_of(0x6F6D7233, p, s, k, b)
            for p, s, k, b in itertools.product(range(32), range(16), range(1, 5), (0, 1, 2, 5))]
    assert len(set(grid)) == len(grid)
    rng = np.random.default_rng(9)
    flips = []
    for _ in range(400):
        a = [int(rng.integers(0, 2 ** 63))] + [int(x) for x in rng.integers(0, 1 << 24, 4)]
        h = seed_of(*a)
        i = int(rng.integers(0, 5))
        bit = int(rng.integers(0, 24))
        b = list(a)
        b[i] ^= 1 << bit
        flips.append(bin(h ^ seed_of(*b)).count("1"))
    assert 14.5 < np.mean(flips) < 17.5
    first = np.array([pyoracle.rng_floats(seed_of(11, p, 0, 1, 0), 1)[0] for p in range(8001)], np.float64)
    counts, _ = np.histogram(first, bins=20, range=(0.0, 1.0))
    chi2 = float(((counts - 400.05) ** 2 / 400.05).sum())
    assert chi2 < 43.8  # 19 degrees of freedom, p = 0.001
    assert abs(np.corrcoef(first[:-1], first[1:])[0, 1]) < 0.05
