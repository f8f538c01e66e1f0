"""The CPU restatement (oracle/) against the reference's own outputs.

The fixtures in tests/golden/ were written by tests/golden/make_golden.py
from the reference sources compiled in place (oracle/Makefile ref target).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import golden_inputs as gi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_rng_matches_reference_random():
    """Random (src/tools/Random.h:9-29): seeding recurrence + xorshift128."""
    import pyoracle

    with open(os.path.join(GOLDEN, "rng_kat.json")) as f:
        kat = json.load(f)
    for seed in gi.RNG_SEEDS:
        ref = kat[str(seed)]
        assert list(pyoracle.rng_next(seed, gi.RNG_N)) == ref["next"]
        f32 = pyoracle.rng_floats(seed, gi.RNG_N).astype(np.float64)
        # binary32 restatement of next()/UINT_MAX: within half an ulp of 1.0
        assert np.max(np.abs(f32 - np.array(ref["next_double"]))) <= 2.0 ** -24


@pytest.mark.parametrize("name", sorted(gi.HIT_SETS))
def test_closest_hit_matches_reference(name):
    """Scene::CheckIntersection over Polygon/Sphere (FP64) vs the binary32
    restatement: identical primitive ids, t within 1e-5 relative."""
    import omr
    import pyoracle

    kind, n, seed, _ = gi.HIT_SETS[name]
    scene = omr.procedural(kind, width=64, height=64)
    rays = gi.hit_rays(scene, n, seed)
    g = np.load(os.path.join(GOLDEN, f"hits_{name}.npz"))
    assert str(g["rays_sha256"]) == _sha(rays), "ray generator drifted from the fixture"
    o = pyoracle.Oracle(scene)
    prim, t = o.trace(rays, pyoracle.OWN_BVH)
    assert np.array_equal(prim, g["prim"])
    hit = prim >= 0
    assert hit.mean() > 0.5
    rel = np.abs(t[hit].astype(np.float64) - g["t"][hit]) / g["t"][hit]
    assert rel.max() < 1e-5


def test_env_texel_matches_reference_ibl():
    """IBL::Sample texel selection (src/renderer/IBL.h:30-45) on the
    reference's own 360x180 HDR: the binary32 restatement picks the same
    texel except within float rounding of a texel border."""
    import pyoracle

    g = np.load(os.path.join(GOLDEN, "env_texels.npz"))
    dirs = gi.env_dirs()
    assert str(g["dirs_sha256"]) == _sha(dirs)
    lib = pyoracle.load()
    import ctypes as C

    w, h = int(g["w"]), int(g["h"])
    ix = np.empty(len(dirs), np.int64)
    iy = np.empty(len(dirs), np.int64)
    for i, d in enumerate(dirs.astype(np.float32)):
        a, b = C.c_uint32(), C.c_uint32()
        lib.orc_env_texel(float(d[0]), float(d[1]), float(d[2]), w, h, C.byref(a), C.byref(b))
        ix[i], iy[i] = a.value, b.value
    same = (ix == g["ix"]) & (iy == g["iy"])
    assert same.mean() > 0.998
    # every disagreement is a neighbouring texel at a border
    dx = np.minimum(np.abs(ix - g["ix"]), w - np.abs(ix - g["ix"]))
    assert np.all(dx[~same] <= 1) and np.all(np.abs(iy - g["iy"])[~same] <= 1)
    # the first eight directions are away from texel borders; the two
    # diagonals at phi = pi/4, 5pi/4 sit exactly on one (u*W = 45.0, 225.0)
    assert same[:8].all(), "directions away from texel borders must match exactly"


def test_integrator_matches_reference_pathtracer_statistically():
    """PathTracer::RenderScene (reference mode: RR after depth 5, no depth
    cap) vs the restatement on the same Cornell box and camera.  The RNG
    streams differ (per-row vs per-path), so the check is statistical:
    per-pixel z-scores of the means, and the image mean."""
    import omr
    import pyoracle

    g = np.load(os.path.join(GOLDEN, "ref_render_c0.npz"))
    c = gi.REF_RENDER
    scene = omr.procedural(0, width=64, height=64)
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    nb = 256
    prm = omr.params(ss=c["ss"], spp_end=nb, max_depth=0, rr_depth=5, seed=99)
    o = pyoracle.Oracle(scene)
    img, _, _ = o.render(cam, prm, pyoracle.BRUTE)
    ref_mean, ref_var = g["mean"].astype(np.float64), g["var"].astype(np.float64)
    # both estimators have the same per-batch variance; ref_var is that of a
    # batch mean of ss^2 samples
    se = np.sqrt(ref_var / c["batches"] + ref_var / nb) + 1e-6
    z = (img.astype(np.float64) - ref_mean) / se
    assert np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4).mean() < 0.01
    rel = abs(img.mean() - ref_mean.mean()) / ref_mean.mean()
    assert rel < 0.01


def _load_cornell_ref():
    import omr

    here = os.path.join(GOLDEN, "scenes")
    cwd = os.getcwd()
    os.chdir(here)
    try:
        return omr.load_scene_file("cornell_ref.scene")
    finally:
        os.chdir(cwd)


def test_cornell_ref_scene_matches_reference_cornellboxscene():
    """cornell_ref.scene, read by the scene-file loader, is the reference's
    own CornellBoxScene (src/scenes/CornellBoxScene.cpp:10-36): on rays in
    the box, the restatement's closest hits have the reference's object
    (same order: floor triangles, then spheres), distance (1e-5 relative) and
    materials with their normalized roulette rates."""
    import pyoracle

    scene = _load_cornell_ref()
    g = np.load(os.path.join(GOLDEN, "hits_cornell_ref.npz"))
    rays = gi.cornell_rays(scene)
    assert str(g["rays_sha256"]) == _sha(rays), "ray generator drifted from the fixture"
    assert int(g["n_objects"]) == scene.tri.shape[0] + scene.sph.shape[0]
    prim, t = pyoracle.Oracle(scene).trace(rays, pyoracle.OWN_BVH)
    assert np.array_equal(prim, g["obj"].astype(np.int32))
    hit = prim >= 0
    assert hit.mean() > 0.9
    # binary32 against FP64: where the hit is well conditioned (the ray starts
    # at least one unit away and meets the surface at more than 3 degrees from
    # grazing) t agrees within 1e-5 relative; everywhere the hit points agree
    # along the surface normal within 1e-4 (the box is 100 units wide)
    dt = np.abs(t[hit].astype(np.float64) - g["t"][hit])
    cos = np.abs(np.sum(g["normal"][hit].astype(np.float64) * rays[hit, 4:7], axis=1))
    good = (g["t"][hit] > 1.0) & (cos > 0.05)
    assert good.mean() > 0.9
    assert (dt[good] / g["t"][hit][good]).max() < 1e-5
    assert (dt * cos).max() < 1e-4
    # materials of the hit object: type, colour, emission, ior, rate
    first = scene.prim_mat[prim[hit]]
    count = scene.prim_mat_count[prim[hit]]
    assert np.array_equal(count, g["nmat"][hit].astype(np.uint32))
    for k in range(2):
        sel = count > k
        m = scene.mats[first[sel] + k]
        ref = g["mat"][hit][sel, k]
        mine = np.column_stack([m["type"].astype(np.float32), m["color"], m["emission"], m["ior"], m["rate"]])
        assert np.allclose(mine, ref, rtol=1e-6, atol=1e-7)
    assert (count == 2).sum() > 100  # the two-material sphere is hit


def _zscores(img, g, nb):
    # (a constant pixel's variance can come out a rounding error below zero)
    ref_mean, ref_var = g["mean"].astype(np.float64), np.maximum(g["var"].astype(np.float64), 0.0)
    se = np.sqrt(ref_var / int(g["batches"]) + ref_var / nb) + 1e-6
    return (img.astype(np.float64) - ref_mean) / se, ref_mean


def test_integrator_matches_reference_on_its_cornellboxscene():
    """The reference's own CornellBoxScene rendered by its PathTracer
    (1 thread, reference mode: RR after depth 5) against the restatement:
    the curved Lambert sphere with a 10 % mirror layer (material roulette
    and the unnormalized Lambert basis on a curved surface), the mirror and
    glass spheres and the sphere light, per-pixel z-scores and image mean."""
    import omr
    import pyoracle

    scene = _load_cornell_ref()
    g = np.load(os.path.join(GOLDEN, "ref_render_cornell.npz"))
    c = gi.CORNELL_RENDER
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    nb = 256
    prm = omr.params(ss=c["ss"], spp_end=nb, max_depth=0, rr_depth=5, seed=31)
    img, _, _ = pyoracle.Oracle(scene).render(cam, prm, pyoracle.OWN_BVH)
    z, ref_mean = _zscores(img, g, nb)
    assert np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4).mean() < 0.01
    assert abs(img.mean() - ref_mean.mean()) / ref_mean.mean() < 0.01


def test_integrator_matches_reference_lambert_sphere_under_its_hdr():
    """A Lambert sphere lit only by the reference's Barce_Rooftop_C_Env.hdr,
    rendered by the reference's PathTracer and by the restatement: every
    sphere pixel is albedo x the environment seen through one Lambert bounce
    with the reference's unnormalized basis (PathTracer.cpp:463-469) on a
    curved surface, so a bounce direction drawn differently moves the
    per-pixel means; background pixels are the IBL lookup itself."""
    import omr
    import pyoracle

    env = omr.hdr_read(os.path.join(GOLDEN, "barce.hdr"))
    scene = gi.ibl_sphere_scene(env)
    g = np.load(os.path.join(GOLDEN, "ref_render_ibl_sphere.npz"))
    c = gi.IBL_RENDER
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    nb = 512
    prm = omr.params(ss=c["ss"], spp_end=nb, max_depth=0, rr_depth=5, seed=17)
    img, _, _ = pyoracle.Oracle(scene).render(cam, prm, pyoracle.BRUTE)
    z, ref_mean = _zscores(img, g, nb)
    # thresholds: the restatement gives mean |z| ~0.24, no |z| > 4 and the
    # image mean within 1e-4; the same restatement with a normalized Lambert
    # basis gives 0.61, 4 % above 4 and 1.3e-3, and fails all three
    assert np.abs(z).mean() < 0.4
    assert (np.abs(z) > 4).mean() < 0.005
    assert abs(img.mean() - ref_mean.mean()) / ref_mean.mean() < 5e-4


def test_integrator_matches_reference_on_textured_scene():
    """textured.scene rendered by the reference's PathTracer (its loader, its
    image library's texture decode, its IBL) and by the restatement on the
    scene the project's loader and decoders build: the textured branch of the
    integrator (hit colour = material colour x texel, PathTracer.cpp:327-346,
    388-389; Russian roulette on the textured colour, the refractive
    textured material's total internal reflection on the untextured one),
    sphere uv and mesh uv, the textured floor, an emissive sphere, the sky
    seen on a miss; per-pixel z-scores and image mean."""
    import omr
    import pyoracle

    here = os.path.join(GOLDEN, "scenes")
    cwd = os.getcwd()
    os.chdir(here)
    try:
        scene = omr.load_scene_file("textured.scene")
    finally:
        os.chdir(cwd)
    g = np.load(os.path.join(GOLDEN, "ref_render_textured.npz"))
    c = gi.TEXTURED_RENDER
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    nb = 256
    prm = omr.params(ss=c["ss"], spp_end=nb, max_depth=0, rr_depth=5, seed=23)
    img, _, _ = pyoracle.Oracle(scene).render(cam, prm, pyoracle.OWN_BVH)
    # the standard error gets a floor of 1 % of the pixel value: the project's
    # JPEG decoder is within 3 levels of the reference's image library, which
    # shows as a systematic difference of up to ~1 % on pixels that are
    # deterministic in both renders (a textured mirror reflecting the sky has
    # zero variance)
    ref_mean, ref_var = g["mean"].astype(np.float64), np.maximum(g["var"].astype(np.float64), 0.0)
    se = np.sqrt(ref_var / int(g["batches"]) + ref_var / nb) + 0.01 * np.abs(ref_mean) + 1e-6
    z = (img.astype(np.float64) - ref_mean) / se
    # the restatement: mean |z| ~0.76, 0.5 % above 4, image mean within 0.5 %
    assert np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4).mean() < 0.01
    assert abs(img.mean() - ref_mean.mean()) / ref_mean.mean() < 0.01


def test_next_event_estimation_mode_matches_reference():
    """PathTracer with EnableNextEventEstimation(true) on the reference's
    CornellBoxScene: emission counts only where the camera ray hits
    (PathTracer.cpp:406-439) and the direct-light term is zero
    (:312-314), so the image holds the directly seen light, the light's
    mirror images and nothing lit by it -- the restatement's nee mode against
    the reference render, z-scores and image mean (0.30 against 0.70 with
    it off)."""
    import omr
    import pyoracle

    scene = _load_cornell_ref()
    g = np.load(os.path.join(GOLDEN, "ref_render_cornell_nee.npz"))
    c = gi.CORNELL_RENDER
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    nb = 256
    prm = omr.params(ss=c["ss"], spp_end=nb, max_depth=0, rr_depth=5, seed=37, nee=1)
    img, _, _ = pyoracle.Oracle(scene).render(cam, prm, pyoracle.OWN_BVH)
    z, ref_mean = _zscores(img, g, nb)
    assert np.abs(z).mean() < 1.0
    assert (np.abs(z) > 4).mean() < 0.01
    assert abs(img.mean() - ref_mean.mean()) / ref_mean.mean() < 0.01
    off = np.load(os.path.join(GOLDEN, "ref_render_cornell.npz"))["mean"].mean()
    assert ref_mean.mean() < 0.6 * off
