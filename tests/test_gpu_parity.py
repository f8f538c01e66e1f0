"""GPU parity: the HIP path against the CPU restatement (bit-exact)."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _oracle(scene, ctx, sid):
    import pyoracle

    return pyoracle.Oracle(scene, ctx.scene_bvh(sid))


@pytest.mark.parametrize("kind", [0, 1])
def test_trace_bit_exact(ctx, kind):
    import omr
    import pyoracle

    scene = omr.procedural(kind, width=64, height=64)
    sid = ctx.upload(scene)
    rays = helpers.hit_rays(scene, 200_000, seed=11 + kind)
    print(f"trace kind={kind} n={rays.shape[0]}", flush=True)
    gp, gt = ctx.trace(sid, rays)
    o = _oracle(scene, ctx, sid)
    cp, ct = o.trace(rays, pyoracle.WIDE_BVH)
    assert np.array_equal(gp, cp), f"{(gp != cp).sum()} hit-id mismatches"
    hit = cp >= 0
    assert np.array_equal(gt[hit].view(np.uint32), ct[hit].view(np.uint32))
    assert hit.mean() > 0.3
    ctx.release(sid)


def test_trace_far_origins_and_axis_directions(ctx):
    """The slab test's edge cases: origins 100 times the scene's extent away
    (beyond what the box padding covers for the fused slab distances, so the
    far-origin slack is used; much farther, binary32 hit distances tie
    within their rounding and no box test reproduces the brute force's
    order, the reference's included), directions with zero,
    denormal and tiny components (the +-2^100 inverse), and origins on the
    bounding planes; ids and distances equal the CPU restatement's brute
    force (every primitive tested, no boxes)."""
    import omr
    import pyoracle

    scene = omr.procedural(1, width=64, height=64)
    sid = ctx.upload(scene)
    rng = np.random.default_rng(23)
    lo, hi = helpers.scene_bounds(scene)
    c, e = (lo + hi) / 2, (hi - lo) / 2
    n = 5_000
    ti = rng.integers(0, scene.tri.shape[0], n)
    a, b = rng.random(n), rng.random(n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    tgt = scene.tri[ti, 0:3] + scene.tri[ti, 3:6] * a[:, None] + scene.tri[ti, 6:9] * b[:, None]
    parts_o, parts_d = [], []
    for far in (30.0, 100.0):
        o = (c + e.max() * far * helpers.unit_dirs(rng, n)).astype(np.float32)
        d = tgt - o
        parts_o.append(o)
        parts_d.append((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    # aimed rays with one component forced to 0, -0, a denormal or 2^-70
    # (the origin's coordinate on that axis is the target's)
    o = (c + e * 1.2 * helpers.unit_dirs(rng, n)).astype(np.float32)
    d = (tgt - o).astype(np.float32)
    ax = rng.integers(0, 3, n)
    small = np.array([0.0, -0.0, 1e-40, -1e-40, 2.0 ** -70, -(2.0 ** -70)], np.float32)
    d[np.arange(n), ax] = small[rng.integers(0, small.size, n)]
    o[np.arange(n), ax] = tgt[np.arange(n), ax]
    parts_o.append(o)
    parts_d.append((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32))
    # axis-aligned rays from inside the bounds and from the bounding planes
    o = (c + e * (2 * rng.random((n, 3)) - 1)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    d[np.arange(n), rng.integers(0, 3, n)] = rng.choice(np.array([1.0, -1.0], np.float32), n)
    plane = rng.integers(0, 3, n)
    on = rng.random(n) < 0.5
    o[on, plane[on]] = np.where(rng.random(on.sum()) < 0.5, lo[plane[on]], hi[plane[on]])
    parts_o.append(o)
    parts_d.append(d)
    rays = helpers.make_rays(np.concatenate(parts_o), np.concatenate(parts_d))
    print(f"trace edge cases n={rays.shape[0]}", flush=True)
    gp, gt = ctx.trace(sid, rays)
    cp, ct = _oracle(scene, ctx, sid).trace(rays, pyoracle.BRUTE)
    bad = np.nonzero(gp != cp)[0]
    assert bad.size == 0, f"{bad.size} hit-id mismatches, by part {np.bincount(bad // n, minlength=4)}"
    hit = cp >= 0
    assert np.array_equal(gt[hit].view(np.uint32), ct[hit].view(np.uint32))
    assert hit[: 2 * n].mean() > 0.5 and hit[2 * n: 3 * n].mean() > 0.3
    ctx.release(sid)


def test_env_tables_bit_exact(ctx):
    """The device-scanned importance-sampling tables equal the sequential
    CPU tables exactly (integer prefix sums)."""
    import omr
    import pyoracle

    scene = omr.procedural(3, width=32, height=32)
    sid = ctx.upload(scene)
    h, w = scene.env.shape[0], scene.env.shape[1]
    grow, gmarg = ctx.env_tables(sid, w, h)
    crow, cmarg = pyoracle.Oracle(scene).env_tables()
    assert np.array_equal(grow, crow) and np.array_equal(gmarg, cmarg)
    assert gmarg[-1] > 0
    ctx.release(sid)


@pytest.mark.parametrize("env_is,tune", [(0, {}), (1, {}), (1, dict(stack_short=8, fetch_budget=8)),
                                         (1, dict(drain_below=1 << 20)), (1, dict(record_mb=0)),
                                         (1, dict(record_mb=0, drain_below=1 << 20)), (1, dict(node_width=8)),
                                         (1, dict(pool_paths=512, lanes=2))])
def test_render_env_bit_exact(ctx, env_is, tune):
    """Torus knot under an HDR sky: lookup on escape, and with importance
    sampling + shadow rays + MIS (also with the 8-entry LDS stack and
    dynamic fetch in the shadow kernel, and with pixel atomics instead of
    per-sample records)."""
    import omr
    import pyoracle

    saved = ctx.get_tuning()
    try:
        ctx.set_tuning(**tune)
        scene = omr.procedural(3, width=40, height=30)
        sid = ctx.upload(scene)
        prm = omr.params(ss=1, spp_end=2, max_depth=6, rr_depth=3, env_is=env_is, seed=77)
        print(f"render knot env_is={env_is} tuning {tune}", flush=True)
        g = ctx.render_sums(sid, scene.camera, prm)
        st = ctx.last_stats()
        c, seg, sh = _oracle(scene, ctx, sid).render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
        assert not (g != c).any(axis=2).any(), f"{(g != c).any(axis=2).sum()} pixels differ"
        assert st["segments"] == seg and st["shadow_rays"] == sh
        assert (sh > 0) == bool(env_is)
        if tune.get("drain_below"):
            assert st["drained_paths"] > 0
        ctx.release(sid)
    finally:
        ctx.set_tuning(**saved)


def test_render_aperture_and_multimaterial_bit_exact(ctx):
    """Thin-lens camera (CircleAperture) and a two-material sphere (material
    roulette over normalized rates)."""
    import omr
    import pyoracle

    base = omr.procedural(0, width=32, height=32)
    mats = np.concatenate([base.mats, base.mats[[2, 4]]])
    mats["rate"][-2:] = [0.7, 0.3]
    pm = base.prim_mat.copy()
    cnt = np.ones_like(pm)
    pm[-2], cnt[-2] = len(base.mats), 2      # mirror sphere becomes grey-lambert 70% / mirror 30%
    scene = omr.Scene(base.tri, base.tri_n, base.sph, pm, mats, prim_mat_count=cnt)
    cam = omr.camera_lookat(40, 32, (50, 50, 260), (0, 0, -1), focal=200.0, aperture=2.5)
    sid = ctx.upload(scene)
    prm = omr.params(ss=2, spp_end=2, max_depth=8, rr_depth=3, seed=5)
    g = ctx.render_sums(sid, cam, prm)
    c, seg, _ = _oracle(scene, ctx, sid).render_sums(cam, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any()
    ctx.release(sid)


@pytest.mark.parametrize("kind,ss,spp,depth,rr", [(0, 2, 2, 8, 3), (1, 2, 1, 16, 3), (0, 1, 3, 0, 5)])
def test_render_bit_exact(ctx, kind, ss, spp, depth, rr):
    import omr
    import pyoracle

    scene = omr.procedural(kind, width=48, height=40)
    sid = ctx.upload(scene)
    prm = omr.params(ss=ss, spp_end=spp, max_depth=depth, rr_depth=rr, seed=1234 + kind)
    print(f"render kind={kind} ss={ss} spp={spp} depth={depth}", flush=True)
    g = ctx.render_sums(sid, scene.camera, prm)
    st = ctx.last_stats()
    o = _oracle(scene, ctx, sid)
    c, seg, _ = o.render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
    diff = (g != c).any(axis=2)
    assert not diff.any(), f"{diff.sum()} pixels differ"
    assert st["segments"] == seg
    ctx.release(sid)


def test_tuning_rejects_invalid_values(ctx):
    """omr_set_tuning refuses values outside what the kernels are built for
    (node width, lanes, odd round steps, group size, stack depth) with
    OMR_ERR_INVALID and leaves the tuning as it was."""
    import omr

    saved = ctx.get_tuning()
    for bad in (dict(node_width=5), dict(lanes=3), dict(lanes=0), dict(round_steps=3), dict(wg_size=96),
                dict(stack_short=12), dict(pool_paths=(1 << 26) + 1)):
        with pytest.raises(omr.OmrError) as ei:
            ctx.set_tuning(**bad)
        assert ei.value.code == omr.ERR_INVALID, bad
        assert ctx.get_tuning() == saved, bad


@pytest.mark.parametrize("tune", [dict(fetch_budget=8, fetch_min=16), dict(drain_rounds=0),
                                  dict(drain_rounds=0, refill_every=2),
                                  dict(drain_rounds=0, fetch_budget=4, fetch_min=32, refill_every=1, use_graph=0,
                                       wg_size=256),
                                  dict(drain_rounds=0, wg_size=64, blocks_per_cu=2, pool_paths=4096, steps_per_sync=4),
                                  dict(wg_size=64, blocks_per_cu=2, pool_paths=4096),
                                  dict(pool_paths=4096, lanes=1), dict(pool_paths=4096, lanes=2),
                                  dict(pool_paths=1000, wg_size=128, blocks_per_cu=1),
                                  dict(stack_short=8), dict(stack_short=8, fetch_budget=8, wg_size=64,
                                                            blocks_per_cu=16),
                                  dict(drain_rounds=0, drain_below=1 << 20),
                                  dict(drain_rounds=0, drain_below=2000, pool_paths=8192),
                                  dict(record_mb=0), dict(drain_rounds=0, record_mb=0, refill_every=2),
                                  dict(node_width=8),
                                  dict(node_width=8, drain_rounds=0, stack_short=8, wg_size=128),
                                  dict(node_width=8, drain_rounds=0, fetch_budget=0, use_graph=0, wg_size=256,
                                       drain_below=1 << 20)])
def test_tuning_modes_bit_exact(ctx, tune):
    """Every launch configuration the autotuner can pick gives the same bits:
    drain rounds (one or many per render, small pools, on one lane or
    alternating between two pools on two streams) and bounce steps with
    dynamic-fetch traversal, bulk refill cadence, graph off, small pools, the
    drain launch that finishes the last paths in one go; the 8-entry LDS
    stack (more entries spill to global memory); 8-wide nodes (the oracle
    traverses the 4-wide tree: the tie rule makes the hit independent of the
    tree)."""
    import omr
    import pyoracle

    saved = ctx.get_tuning()
    try:
        ctx.set_tuning(**tune)
        scene = omr.procedural(1, width=40, height=32)
        sid = ctx.upload(scene)
        o = _oracle(scene, ctx, sid)
        rays = helpers.hit_rays(scene, 50_000, seed=21)
        gp, gt = ctx.trace(sid, rays)
        cp, ct = o.trace(rays, pyoracle.WIDE_BVH)
        assert np.array_equal(gp, cp)
        prm = omr.params(ss=2, spp_end=2, max_depth=12, rr_depth=3, seed=31)
        print("tuning", tune, flush=True)
        g = ctx.render_sums(sid, scene.camera, prm)
        c, seg, _ = o.render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
        assert not (g != c).any(axis=2).any()
        st = ctx.last_stats()
        assert st["segments"] == seg
        if tune.get("drain_below"):
            assert st["drained_paths"] > 0 and 0 < st["drain_segments"] <= seg
        if tune.get("drain_rounds", 1) and tune.get("pool_paths"):
            assert st["drain_launches"] == -(-40 * 32 * 4 * 2 // tune["pool_paths"])  # one per round
        ctx.release(sid)
    finally:
        ctx.set_tuning(**saved)


def test_trace_planes_bit_exact(ctx):
    """Infinite floors (tested after the BVH, ids after the spheres)."""
    import golden_inputs as gi
    import pyoracle

    scene = gi.plane_scene()
    sid = ctx.upload(scene)
    rays = helpers.hit_rays(scene, 100_000, seed=41)
    print(f"trace planes n={rays.shape[0]}", flush=True)
    gp, gt = ctx.trace(sid, rays)
    cp, ct = _oracle(scene, ctx, sid).trace(rays, pyoracle.WIDE_BVH)
    assert np.array_equal(gp, cp)
    hit = cp >= 0
    assert np.array_equal(gt[hit].view(np.uint32), ct[hit].view(np.uint32))
    assert (cp >= scene.tri.shape[0] + scene.sph.shape[0]).sum() > 10_000
    ctx.release(sid)


@pytest.mark.parametrize("name,env_is", [("loader", 0), ("loader", 1), ("textured", 0), ("textured", 1),
                                         ("cornell_ref", 0)])
def test_render_loaded_scene_bit_exact(ctx, name, env_is):
    """Scenes read from the .scene format render to the same bits as the
    restatement: loader.scene (OBJ mesh with MTL materials, multi-material
    sphere and floor, emissive sphere, glass infinite floor, RGBE sky) and
    textured.scene (PNG / JPEG textures on spheres, floor and mesh, a
    textured refractive material)."""
    import os

    import omr
    import pyoracle

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
    cwd = os.getcwd()
    os.chdir(here)
    try:
        scene = omr.load_scene_file(f"{name}.scene")
    finally:
        os.chdir(cwd)
    if name == "textured":
        assert len(scene.textures) == 6
        cam = omr.camera_lookat(48, 32, (0.0, 25.0, 80.0), (0.0, -0.3, -1.0), screen_height=30.0, dist=40.0)
    elif name == "cornell_ref":  # the reference's CornellBoxScene, its settings.txt camera
        cam = omr.camera_lookat(64, 36, (50.0, 55.0, 230.0), (0.00001, -0.15, -1.0), screen_height=30.0, dist=40.0)
    else:
        cam = omr.camera_lookat(48, 32, (0.0, 30.0, 110.0), (0.0, -0.25, -1.0), screen_height=30.0, dist=40.0)
    sid = ctx.upload(scene)
    prm = omr.params(ss=2, spp_end=2, max_depth=8, rr_depth=3, env_is=env_is, seed=2024)
    print(f"render {name}.scene env_is={env_is}", flush=True)
    g = ctx.render_sums(sid, cam, prm)
    st = ctx.last_stats()
    c, seg, sh = _oracle(scene, ctx, sid).render_sums(cam, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any(), f"{(g != c).any(axis=2).sum()} pixels differ"
    assert st["segments"] == seg and st["shadow_rays"] == sh
    assert g.any()
    ctx.release(sid)


@pytest.mark.parametrize("env_is", [0, 1])
def test_render_lambert_sphere_under_reference_hdr_bit_exact(ctx, env_is):
    """The Lambert sphere under the reference's Barce HDR (the scene whose
    restatement render is checked against the reference's PathTracer),
    reference mode (RR after depth 5, no depth cap), with and without
    environment importance sampling."""
    import os

    import golden_inputs as gi
    import omr
    import pyoracle

    env = omr.hdr_read(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "barce.hdr"))
    scene = gi.ibl_sphere_scene(env)
    c = gi.IBL_RENDER
    cam = omr.camera_lookat(c["w"], c["h"], c["pos"], c["dir"], screen_height=c["screen_h"], dist=c["dist"])
    sid = ctx.upload(scene)
    prm = omr.params(ss=2, spp_end=4, max_depth=0, rr_depth=5, env_is=env_is, seed=2025)
    print(f"render ibl sphere env_is={env_is}", flush=True)
    g = ctx.render_sums(sid, cam, prm)
    st = ctx.last_stats()
    c_, seg, sh = _oracle(scene, ctx, sid).render_sums(cam, prm, pyoracle.WIDE_BVH)
    assert not (g != c_).any(axis=2).any(), f"{(g != c_).any(axis=2).sum()} pixels differ"
    assert st["segments"] == seg and st["shadow_rays"] == sh
    ctx.release(sid)


def test_render_next_event_estimation_mode_bit_exact(ctx):
    """The reference's next-event estimation mode (emission counted where the
    camera ray hits only) on its CornellBoxScene: GPU equals the restatement;
    a scene whose mirror material emits is refused in that mode."""
    import os

    import omr
    import pyoracle

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")
    cwd = os.getcwd()
    os.chdir(here)
    try:
        scene = omr.load_scene_file("cornell_ref.scene")
    finally:
        os.chdir(cwd)
    cam = omr.camera_lookat(64, 36, (50.0, 55.0, 230.0), (0.00001, -0.15, -1.0), screen_height=30.0, dist=40.0)
    sid = ctx.upload(scene)
    prm = omr.params(ss=2, spp_end=2, max_depth=0, rr_depth=5, seed=4, nee=1)
    print("render cornell_ref nee", flush=True)
    g = ctx.render_sums(sid, cam, prm)
    c, seg, _ = _oracle(scene, ctx, sid).render_sums(cam, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any()
    assert ctx.last_stats()["segments"] == seg
    ctx.release(sid)
    mats = scene.mats.copy()
    mats["emission"][mats["type"] == 1] = (1.0, 1.0, 1.0)
    bad = omr.Scene(scene.tri, scene.tri_n, scene.sph, scene.prim_mat, mats, prim_mat_count=scene.prim_mat_count)
    sid = ctx.upload(bad)
    with pytest.raises(omr.OmrError):
        ctx.render_sums(sid, cam, prm)
    ctx.release(sid)
