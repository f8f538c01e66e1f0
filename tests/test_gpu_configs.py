"""Every benchmark configuration on the GPU, checked against the CPU
restatement (bit-exact) or against the same jobs run alone.

Configurations (the workloads bench.py measures):
  c0  Cornell box, 256x256, 16 spp, depth 8, RR after 3 (full size)
  c1  Cornell box + displaced level-6 icosphere: 1,000,000 seeded rays for the
      hit-id oracle (half inside 1.5x the mesh bounds with uniform directions,
      half aimed at mesh points from outside), and the render at its own spp
      and depth on a reduced image
  c2  terrain (997,002 triangles) + 200 spheres under a 2048x1024 sky with a
      sun lobe: importance-sampling tables, 1M rays, the render with
      importance sampling on a reduced image
  c3  the 64 uneven batched jobs over four scenes, each equal to the job alone
  c4  a served stream over the eight resident scenes, each result equal to
      the request alone
"""
import importlib.util
import os
import time

import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_workloads", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _oracle(scene, ctx, sid):
    import pyoracle

    return pyoracle.Oracle(scene, ctx.scene_bvh(sid))


def _same_hits(gp, gt, cp, ct):
    assert np.array_equal(gp, cp), f"{int((gp != cp).sum())} hit-id mismatches of {gp.size}"
    hit = cp >= 0
    assert np.array_equal(gt[hit].view(np.uint32), ct[hit].view(np.uint32)), "hit distances differ"


def test_c0_full_render_bit_exact(ctx):
    """Config 0 at its full size: 256x256, 4x4 sub-pixels, depth 8."""
    import omr
    import pyoracle

    scene = omr.procedural(0, width=256, height=256)
    sid = ctx.upload(scene)
    prm = omr.params(ss=4, spp_end=1, max_depth=8, rr_depth=3, seed=0x6F6D7233)
    print("c0 256x256 16 spp", flush=True)
    g = ctx.render_sums(sid, scene.camera, prm)
    st = ctx.last_stats()
    c, seg, _ = _oracle(scene, ctx, sid).render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any(), f"{int((g != c).any(axis=2).sum())} pixels differ"
    assert st["segments"] == seg
    ctx.release(sid)


@pytest.mark.parametrize("node_width", [4, 8])
def test_c1_hit_ids_1m_rays(ctx, node_width):
    """Config 1's hit-id oracle: 1,000,000 seeded rays through the mesh's BVH
    (4-wide and 8-wide nodes), ids and distances identical to the restatement
    traversing the 4-wide BVH and its own binary BVH (different trees: the
    tie rule makes the result independent of the tree)."""
    import omr
    import pyoracle

    scene = omr.procedural(1, width=64, height=64)
    mesh = scene.tri[12:]
    pts = np.concatenate([mesh[:, 0:3], mesh[:, 0:3] + mesh[:, 3:6], mesh[:, 0:3] + mesh[:, 6:9]])
    rays = helpers.hit_rays(scene, 1_000_000, seed=0x6F6D7233, focus=(pts.min(0), pts.max(0)))
    sid = ctx.upload(scene)
    print(f"c1 trace n={rays.shape[0]} node_width={node_width}", flush=True)
    saved = ctx.get_tuning()
    try:
        ctx.set_tuning(node_width=node_width)
        gp, gt = ctx.trace(sid, rays)
    finally:
        ctx.set_tuning(**saved)
    o = _oracle(scene, ctx, sid)
    cp, ct = o.trace(rays, pyoracle.WIDE_BVH)
    _same_hits(gp, gt, cp, ct)
    bp, bt = o.trace(rays, pyoracle.OWN_BVH)
    _same_hits(gp, gt, bp, bt)
    assert (gp >= 12).mean() > 0.4  # most rays hit the mesh
    ctx.release(sid)


def test_c1_render_bit_exact(ctx):
    """Config 1's render settings (8x8 sub-pixels = 64 spp, depth 16, half
    the mesh Lambert and half glass) on a 48x48 image."""
    import omr
    import pyoracle

    scene = omr.procedural(1, width=48, height=48)
    sid = ctx.upload(scene)
    prm = omr.params(ss=8, spp_end=1, max_depth=16, rr_depth=3, seed=0x6F6D7233)
    print("c1 48x48 64 spp", flush=True)
    g = ctx.render_sums(sid, scene.camera, prm)
    st = ctx.last_stats()
    c, seg, _ = _oracle(scene, ctx, sid).render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any(), f"{int((g != c).any(axis=2).sum())} pixels differ"
    assert st["segments"] == seg
    ctx.release(sid)


@pytest.fixture(scope="module")
def c2(ctx):
    import omr

    scene = omr.procedural(2, width=160, height=90)
    sid = ctx.upload(scene)
    yield scene, sid
    ctx.release(sid)


def test_c2_env_tables_2048x1024(ctx, c2):
    """Importance-sampling tables of the 2048x1024 sun-lobe sky, scanned on
    the device: identical to the sequential CPU prefix sums."""
    import pyoracle

    scene, sid = c2
    h, w = scene.env.shape[0], scene.env.shape[1]
    assert (w, h) == (2048, 1024)
    grow, gmarg = ctx.env_tables(sid, w, h)
    crow, cmarg = pyoracle.Oracle(scene).env_tables()
    assert np.array_equal(grow, crow) and np.array_equal(gmarg, cmarg)
    # the sun lobe dominates: a few rows hold most of the weight
    per_row = np.diff(gmarg.astype(np.float64))
    assert per_row.max() > 20 * np.median(per_row)


def test_c2_hit_ids_1m_rays(ctx, c2):
    """1,000,000 rays over the terrain and spheres (1M-triangle BVH)."""
    import pyoracle

    scene, sid = c2
    lo, hi = helpers.scene_bounds(scene)
    rays = helpers.hit_rays(scene, 1_000_000, seed=2, focus=(lo, hi), box_scale=1.0)
    print(f"c2 trace n={rays.shape[0]}", flush=True)
    gp, gt = ctx.trace(sid, rays)
    cp, ct = _oracle(scene, ctx, sid).trace(rays, pyoracle.WIDE_BVH)
    _same_hits(gp, gt, cp, ct)
    assert (gp >= 0).mean() > 0.3
    assert (gp >= scene.tri.shape[0]).sum() > 100  # sphere hits


@pytest.mark.parametrize("tune", [{}, dict(record_mb=0), dict(node_width=8)])
def test_c2_render_env_is_bit_exact(ctx, c2, tune):
    """Config 2's path (terrain + spheres, depth 12, environment importance
    sampling with shadow rays and MIS, bucket-directory CDF inversion on the
    2048x1024 tables) on a 160x90 image, 2x2 sub-pixels x 2 samples; with
    per-sample records, with pixel atomics, and with 8-wide nodes."""
    import omr
    import pyoracle

    scene, sid = c2
    saved = ctx.get_tuning()
    try:
        ctx.set_tuning(**tune)
        prm = omr.params(ss=2, spp_end=2, max_depth=12, rr_depth=3, env_is=1, seed=0x6F6D7233)
        print(f"c2 160x90 env_is {tune}", flush=True)
        g = ctx.render_sums(sid, scene.camera, prm)
        st = ctx.last_stats()
    finally:
        ctx.set_tuning(**saved)
    c, seg, sh = _oracle(scene, ctx, sid).render_sums(scene.camera, prm, pyoracle.WIDE_BVH)
    assert not (g != c).any(axis=2).any(), f"{int((g != c).any(axis=2).sum())} pixels differ"
    assert st["segments"] == seg and st["shadow_rays"] == sh and sh > 0


def test_c3_batched_jobs_equal_alone(ctx):
    """Config 3's 64 uneven jobs over four resident scenes (pixel count
    log-uniform in 64^2..1024^2, spp 1..256, depth 4/8/16) in one persistent
    run: every image bit-identical to the job rendered alone."""
    import omr

    bench = _bench()
    jobs = bench.batched_jobs()
    assert len(jobs) == 64 and {j["kind"] for j in jobs} == {0, 1, 2, 3}
    scenes = {k: omr.procedural(k, width=64, height=64) for k in (0, 1, 2, 3)}
    sids = {k: ctx.upload(s) for k, s in scenes.items()}
    cams = [omr.procedural_camera(j["kind"], j["w"], j["h"]) for j in jobs]
    prms = [bench._prm(omr, j) for j in jobs]
    ids = [sids[j["kind"]] for j in jobs]
    print("c3 batch of 64", flush=True)
    batch = ctx.render_batch(ids, cams, prms)
    st = ctx.last_stats()
    assert st["segments"] > 0
    for i in range(len(jobs)):
        alone = ctx.render(ids[i], cams[i], prms[i])
        assert np.array_equal(alone, batch[i]), f"job {i} ({jobs[i]}) differs"
    for s in sids.values():
        ctx.release(s)


def test_c4_served_stream_equals_alone():
    """Config 4's served form over the eight resident scenes (5.75M
    triangles, two sky maps): 48 requests drawn by the benchmark's seeded
    generator (pixel count log-uniform in 64^2..1920x1080, spp in
    {1,4,16,64,256}, random seeds), submitted in seeded bursts where each
    burst waits until the previous one's first request is admitted, so
    later requests join work in flight.  Every result equals the request
    rendered alone on the plain path."""
    import omr

    bench = _bench()
    ctx = omr.Context(0)
    kinds = bench.SERVED_SCENES
    scenes = [omr.procedural(k, seed=0x6F6D7233 + s, width=64, height=64) for k, s in kinds]
    sids = [ctx.upload(s) for s in scenes]
    assert sum(s.tri.shape[0] for s in scenes) > 5_700_000
    reqs = bench.served_requests(48, kinds, seed=7)

    def job_args(r):
        return sids[r["scene"]], omr.procedural_camera(kinds[r["scene"]][0], r["w"], r["h"]), bench._prm(omr, r)

    alone = [ctx.render(*job_args(r)) for r in reqs]
    rng = np.random.default_rng(5)
    bursts = []
    k = 0
    while k < len(reqs):
        n = int(rng.integers(1, 8))
        bursts.append(list(range(k, min(len(reqs), k + n))))
        k += n
    ctx.server_start()
    outs = [np.zeros_like(a) for a in alone]
    ids = [None] * len(reqs)
    try:
        for b in bursts:
            for i in b:
                ids[i] = ctx.submit(*job_args(reqs[i]), outs[i])
            t0 = time.time()
            while ctx.poll(ids[b[0]]) == 0 and time.time() - t0 < 60:
                time.sleep(0.0005)
        print(f"c4 submitted {len(reqs)} in {len(bursts)} bursts", flush=True)
        for i in ids:
            assert ctx.wait(i, 120000) == 0, ctx.job_error(i)
        st = ctx.server_stats()
    finally:
        ctx.server_stop()
    assert st["jobs_done"] == len(reqs)
    for i, (a, o) in enumerate(zip(alone, outs)):
        assert np.array_equal(a, o), f"request {i} ({reqs[i]}) differs"
    ctx.close()
