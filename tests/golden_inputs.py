"""Seeded inputs shared by the fixture script and the tests that read the
fixtures (so the fixtures only need to hold outputs)."""
import numpy as np

import helpers

RNG_SEEDS = [1, 2, 12345, 0xDEADBEEF, 4294967295]
RNG_N = 64

# name -> (procedural kind, ray count, seed, reference uses its binary BVH)
HIT_SETS = {
    "c0": (0, 20000, 101, False),
    "c1": (1, 8000, 202, True),
}

REF_RENDER = dict(w=24, h=24, pos=(50.0, 50.0, 260.0), dir=(0.0, 0.0, -1.0), screen_h=30.0, dist=40.0, ss=2,
                  batches=256)


def hit_rays(scene, n, seed):
    return helpers.hit_rays(scene, n, seed)


def env_dirs():
    rng = np.random.default_rng(77)
    d = helpers.unit_dirs(rng, 4000).astype(np.float64)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1], [0.6, 0.8, 0], [0.6, -0.8, 0], [0, 0.6, 0.8],
                     [0, -0.6, -0.8], [0.70710678, 0, 0.70710678], [-0.70710678, 0, -0.70710678]], dtype=np.float64)
    return np.concatenate([axes, d])


# RGBE images written by the reference's HDRImage writer: name -> (w, h, seed).
# Widths 8..32767 take the run-length encoded scanline path, others the flat one.
HDR_SETS = {"rle": (48, 7, 5), "flat": (5, 3, 6)}
SKY = (64, 32, 7)

# Planes (infinite floors) added to the procedural Cornell box for hit vectors.
PLANES_Y = (12.5, 71.0)
PLANE_HITS = (0, 20000, 303)

LOADER_RAYS = (30000, 404)

# The reference's own CornellBoxScene (src/scenes/CornellBoxScene.cpp:10-36),
# restated as tests/golden/scenes/cornell_ref.scene: a ray set inside the box
# and a render with the camera of the reference's settings.txt for it.
CORNELL_RAYS = (20000, 505)
CORNELL_BOX = ((0.0, 0.0, 0.0), (100.0, 81.6, 250.0))
CORNELL_RENDER = dict(w=64, h=36, pos=(50.0, 55.0, 230.0), dir=(0.00001, -0.15, -1.0), screen_h=30.0, dist=40.0,
                      ss=2, batches=256)

# A Lambert sphere lit only by the reference's Barce_Rooftop_C_Env.hdr
# (tests/golden/barce.hdr: the same texels, written by the RGBE writer).
IBL_SPHERE = dict(center=(0.0, 0.0, 0.0), radius=20.0, albedo=(0.8, 0.8, 0.8))
IBL_RENDER = dict(w=32, h=32, pos=(0.0, 0.0, 90.0), dir=(0.0, 0.0, -1.0), screen_h=30.0, dist=40.0, ss=2,
                  batches=256)


# tests/golden/scenes/textured.scene (textured spheres, a textured OBJ mesh
# and floor, a sphere light, an infinite floor, sky.hdr) rendered from above.
TEXTURED_RENDER = dict(w=48, h=32, pos=(0.0, 30.0, 70.0), dir=(0.0, -25.0, -70.0), screen_h=30.0, dist=40.0,
                       ss=2, batches=256)


def cornell_rays(scene):
    n, seed = CORNELL_RAYS
    lo, hi = (np.asarray(v, np.float32) for v in CORNELL_BOX)
    return helpers.hit_rays(scene, n, seed, box_scale=1.0, focus=(lo, hi))


def ibl_sphere_scene(env):
    """The Lambert sphere as an omr.Scene (no triangles) with `env` as sky."""
    import omr

    mats = np.zeros(1, dtype=omr.MATERIAL_DTYPE)
    mats["type"] = 0
    mats["color"] = IBL_SPHERE["albedo"]
    mats["rate"] = 1.0
    sph = np.array([list(IBL_SPHERE["center"]) + [IBL_SPHERE["radius"]]], np.float32)
    return omr.Scene(np.zeros((0, 9), np.float32), np.zeros((0, 9), np.float32), sph, np.zeros(1, np.uint32), mats,
                     env=env)


def hdr_image(w, h, seed):
    """Non-negative float image with constant runs (RLE runs of every
    length), zeros, values below RGBE's smallest exponent and large values."""
    rng = np.random.default_rng(seed)
    img = (rng.random((h, w, 3)) ** 3 * 4.0).astype(np.float32)
    for y in range(h):
        a = int(rng.integers(0, max(1, w - 1)))
        L = int(rng.integers(1, max(2, w // 2)))
        img[y, a:a + L] = img[y, a]
    img[0, : min(w, 3)] = 0.0
    img[-1, 0] = (1e-39, 2e-38, 1e-30)
    img[-1, -1] = (65504.0, 1.5e5, 3.0)
    if w > 10:
        img[1, 2:9, 1] = 0.25          # a run in one channel only
    return img


def sky_image():
    """Latlong sky: a vertical gradient, a bright sun and a dark ground."""
    w, h, seed = SKY
    rng = np.random.default_rng(seed)
    y = (np.arange(h) + 0.5) / h
    x = (np.arange(w) + 0.5) / w
    g = np.clip(1.2 - 1.1 * y, 0.05, None)[:, None] * np.ones(w)[None, :]
    img = np.stack([0.5 * g, 0.7 * g, 1.1 * g], axis=2)
    sx, sy = 0.3, 0.22
    d2 = ((x[None, :] - sx) * 2) ** 2 + (y[:, None] - sy) ** 2
    img += (60.0 * np.exp(-d2 / 0.002))[:, :, None] * np.array([1.0, 0.9, 0.7])
    img[h // 2 + 2:] = 0.08
    img *= 1.0 + 0.05 * rng.random(img.shape)
    return img.astype(np.float32)


def plane_scene():
    """Procedural Cornell box plus two infinite floors (glass, blue lambert)."""
    import omr

    s = omr.procedural(PLANE_HITS[0], width=64, height=64)
    mats = np.concatenate([s.mats, s.mats[[5, 1]]])
    m0 = len(s.mats)
    pm = np.concatenate([s.prim_mat, np.array([m0, m0 + 1], np.uint32)])
    return omr.Scene(s.tri, s.tri_n, s.sph, pm, mats, planes=np.array(PLANES_Y, np.float32))


def loader_rays(which="loader"):
    """Rays through a test .scene: a third from a box around the objects in
    uniform directions (many reach the infinite floor from above or below),
    a third aimed from outside at points inside the objects' box and a third
    at points in the box of the placed mesh."""
    n, seed = LOADER_RAYS if which == "loader" else TEXTURED_RAYS
    rng = np.random.default_rng(seed)
    if which == "loader":
        start = ([-60.0, -20.0, -60.0], [60.0, 50.0, 40.0])
        boxes = (([-25.0, -2.0, -40.0], [45.0, 42.0, 25.0]), ([4.0, 3.0, -32.0], [26.0, 26.0, -8.0]))
    else:
        start = ([-50.0, -10.0, -50.0], [50.0, 50.0, 50.0])
        boxes = (([-22.0, -1.0, -20.0], [22.0, 18.0, 25.0]), ([-20.0, 0.0, -22.0], [0.0, 17.0, -8.0]))
    lo, hi = np.array(start[0]), np.array(start[1])
    n1 = n // 3
    o = [lo + (hi - lo) * rng.random((n1, 3))]
    d = [helpers.unit_dirs(rng, n1)]
    for (tlo, thi), m in ((boxes[0], n1), (boxes[1], n - 2 * n1)):
        tlo, thi = np.array(tlo), np.array(thi)
        tgt = tlo + (thi - tlo) * rng.random((m, 3))
        oo = (tlo + thi) / 2 + 90.0 * helpers.unit_dirs(rng, m)
        dd = tgt - oo
        o.append(oo)
        d.append(dd / np.linalg.norm(dd, axis=1, keepdims=True))
    return helpers.make_rays(np.concatenate(o).astype(np.float32), np.concatenate(d).astype(np.float32))


TEXTURED_RAYS = (30000, 505)

# Texture fixtures in tests/golden/tex/: name -> (format, width, height, seed, options)
TEX_SIZE = (37, 21)


def texture_image(w, h, seed):
    """Smooth colour ramps, a hard edge and noise: rows from the top, uint8 RGB."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    r = 255.0 * x / max(1, w - 1)
    g = 255.0 * y / max(1, h - 1)
    b = np.where((x // 6 + y // 5) % 2 == 0, 230.0, 25.0)
    img = np.stack([r, g, b], axis=2) + rng.normal(0, 6, (h, w, 3))
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


PNG_SETS = {
    "png_rgb": dict(ctype=2, interlace=False, seed=11),
    "png_rgba_adam7": dict(ctype=6, interlace=True, seed=12),
    "png_pal": dict(ctype=3, interlace=False, seed=13),
    "png_pal_trns": dict(ctype=3, interlace=True, seed=14, trns=True),
    "png_grey": dict(ctype=0, interlace=False, seed=15),
    "png_grey_alpha": dict(ctype=4, interlace=False, seed=16),
}
JPEG_SETS = {
    "jpg_444": dict(sampling=(1, 1), quality=90, restart=0, separate_scans=False, grey=False, seed=21),
    "jpg_422": dict(sampling=(2, 1), quality=80, restart=0, separate_scans=False, grey=False, seed=22),
    "jpg_420_rst": dict(sampling=(2, 2), quality=75, restart=2, separate_scans=False, grey=False, seed=23),
    "jpg_420_scans": dict(sampling=(2, 2), quality=85, restart=0, separate_scans=True, grey=False, seed=24),
    "jpg_grey": dict(sampling=(1, 1), quality=85, restart=3, separate_scans=False, grey=True, seed=25),
}


def png_samples(name):
    """(samples (h, w, spp), palette, trns, rgba reference) for a PNG set."""
    o = PNG_SETS[name]
    w, h = TEX_SIZE
    rgb = texture_image(w, h, o["seed"])
    alpha = ((np.arange(w)[None, :] * 7 + np.arange(h)[:, None] * 3) % 256).astype(np.uint8)[..., None]
    if o["ctype"] == 2:
        return rgb, None, None
    if o["ctype"] == 6:
        return np.concatenate([rgb, alpha], axis=2), None, None
    if o["ctype"] == 0:
        return rgb[..., :1], None, None
    if o["ctype"] == 4:
        return np.concatenate([rgb[..., :1], alpha], axis=2), None, None
    # palette: 40 colours quantized from the image
    pal = texture_image(8, 5, o["seed"] + 100).reshape(-1, 3)
    d = ((rgb[:, :, None, :].astype(int) - pal[None, None, :, :].astype(int)) ** 2).sum(-1)
    idx = d.argmin(-1).astype(np.uint8)[..., None]
    trns = (np.arange(len(pal)) * 6 % 256).astype(np.uint8)[:30] if o.get("trns") else None
    return idx, pal, trns


def sphere_normals():
    """Unit normals for the sphere uv vectors (axes included)."""
    rng = np.random.default_rng(606)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 0, 1], [0, 0, -1], [0, 1, 0], [0, -1, 0]], float)
    return np.concatenate([helpers.unit_dirs(rng, 4000).astype(np.float64), axes])


def uv_queries():
    """(u, v) pairs for the texel vectors, some outside [0, 1]."""
    rng = np.random.default_rng(607)
    edge = np.array([[0, 0], [1, 1], [0.5, 0.5], [-1, 2], [0.999999, 0.25]], float)
    return np.concatenate([rng.random((4000, 2)) * 1.2 - 0.1, edge])


TEXEL_SIZES = ((37, 21), (64, 32), (1, 1))
