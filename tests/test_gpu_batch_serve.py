"""Batched and served modes: every job's image is bit-identical whether it runs
alone or packed with other jobs into the shared queues."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs(omr):
    s0 = omr.procedural(0, width=64, height=48)
    s1 = omr.procedural(1, width=80, height=64)
    spec = [  # (scene, w, h, ss, spp_begin, spp_end, depth, rr, seed)
        (0, 64, 48, 2, 0, 2, 8, 3, 1),
        (1, 80, 64, 1, 0, 3, 16, 3, 2),
        (0, 17, 9, 3, 0, 1, 4, 3, 3),
        (1, 33, 21, 2, 2, 5, 16, 3, 4),
        (0, 64, 48, 1, 0, 1, 0, 5, 5),
        (1, 5, 3, 4, 0, 7, 8, 3, 6),
    ]
    return (s0, s1), spec


def _cam(omr, scene, w, h):
    import ctypes as C

    c = omr.Camera()
    C.memmove(C.byref(c), C.byref(scene.camera), C.sizeof(omr.Camera))
    c.width, c.height = w, h
    return c


@pytest.mark.parametrize("record_mb", [None, 0, 1])
def test_batch_equals_alone(ctx, record_mb):
    """Also across the two ways of summing: per-sample records (default),
    pixel atomics (record_mb=0), and a budget that gives records to the first
    jobs of the batch only (record_mb=1) while each job alone has them."""
    import omr

    saved = ctx.get_tuning()
    if record_mb is not None:
        ctx.set_tuning(record_mb=record_mb)
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    cams, prms, ids = [], [], []
    for si, w, h, ss, b, e, d, rr, seed in spec:
        cams.append(_cam(omr, scenes[si], w, h))
        prms.append(omr.params(ss=ss, spp_begin=b, spp_end=e, max_depth=d, rr_depth=rr, seed=seed))
        ids.append(sids[si])
    print("batch of", len(spec), flush=True)
    batch = ctx.render_batch(ids, cams, prms)
    try:
        for i in range(len(spec)):
            alone = ctx.render(ids[i], cams[i], prms[i])
            assert np.array_equal(alone, batch[i]), f"job {i} differs"
            assert alone.mean() > 0
    finally:
        ctx.set_tuning(**saved)
    for s in sids:
        ctx.release(s)


@pytest.mark.parametrize("tune", [dict(pool_paths=4096, solo_round_min=1024), dict(pool_paths=4096, solo_round_min=0),
                                  dict(pool_paths=8192, solo_round_min=1 << 30),
                                  dict(pool_paths=4096, solo_round_min=1024, lanes=1)])
def test_batch_placed_rounds_equal_alone(ctx, tune):
    """Drain rounds placed by the host over a small pool: rounds inside one
    job run the single-job kernels, rounds across jobs the per-lane ones,
    and a job with enough samples left ends in a round of its own (with one
    lane or two: batches keep one); every image equals the job alone."""
    import omr

    saved = ctx.get_tuning()
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    cams, prms, ids = [], [], []
    for si, w, h, ss, b, e, d, rr, seed in spec:
        cams.append(_cam(omr, scenes[si], w, h))
        prms.append(omr.params(ss=ss, spp_begin=b, spp_end=e, max_depth=d, rr_depth=rr, seed=seed))
        ids.append(sids[si])
    alone = [ctx.render(ids[i], cams[i], prms[i]) for i in range(len(spec))]
    print("placed rounds", tune, flush=True)
    try:
        ctx.set_tuning(**tune)
        batch = ctx.render_batch(ids, cams, prms)
        st = ctx.last_stats()
    finally:
        ctx.set_tuning(**saved)
    for i in range(len(spec)):
        assert np.array_equal(alone[i], batch[i]), f"job {i} differs"
    assert st["drain_launches"] >= -(-sum(
        w * h * ss * ss * (e - b) for _, w, h, ss, b, e, _, _, _ in spec) // tune["pool_paths"])
    for s in sids:
        ctx.release(s)


def test_batch_rounds_with_and_without_sky_equal_alone(ctx):
    """Placed rounds over jobs of a scene with an HDR sky (importance
    sampling on) and of scenes without one: rounds whose jobs have no sky run
    the kernels built without the environment code, the others the full
    ones; every image equals the job alone."""
    import omr

    saved = ctx.get_tuning()
    scenes = (omr.procedural(1, width=32, height=32), omr.procedural(3, width=32, height=32),
              omr.procedural(0, width=32, height=32))
    sids = [ctx.upload(s) for s in scenes]
    spec = [(0, 40, 30, 2, 8, 11), (1, 36, 28, 1, 6, 12), (2, 40, 30, 2, 8, 13), (0, 24, 20, 3, 16, 14),
            (1, 20, 16, 2, 8, 15), (2, 48, 36, 1, 8, 16)]  # (scene, w, h, spp, depth, seed)
    cams = [_cam(omr, scenes[si], w, h) for si, w, h, _, _, _ in spec]
    prms = [omr.params(ss=1, spp_end=spp, max_depth=d, rr_depth=3, env_is=1, seed=seed)
            for _, _, _, spp, d, seed in spec]
    ids = [sids[si] for si, *_ in spec]
    alone = [ctx.render(ids[i], cams[i], prms[i]) for i in range(len(spec))]
    print("rounds with and without sky", flush=True)
    try:
        ctx.set_tuning(pool_paths=1024, solo_round_min=1 << 30)
        batch = ctx.render_batch(ids, cams, prms)
    finally:
        ctx.set_tuning(**saved)
    for i in range(len(spec)):
        assert np.array_equal(alone[i], batch[i]), f"job {i} differs"
        assert alone[i].mean() > 0
    for s in sids:
        ctx.release(s)


def test_batch_with_8wide_nodes_equals_4wide(ctx):
    """The batched (per-lane job) kernels over 8-wide nodes give the images
    of the 4-wide nodes, bit for bit."""
    import omr

    saved = ctx.get_tuning()
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    cams, prms, ids = [], [], []
    for si, w, h, ss, b, e, d, rr, seed in spec:
        cams.append(_cam(omr, scenes[si], w, h))
        prms.append(omr.params(ss=ss, spp_begin=b, spp_end=e, max_depth=d, rr_depth=rr, seed=seed))
        ids.append(sids[si])
    print("batch of", len(spec), "at node widths 4 and 8", flush=True)
    try:
        b4 = ctx.render_batch(ids, cams, prms)
        ctx.set_tuning(node_width=8)
        b8 = ctx.render_batch(ids, cams, prms)
    finally:
        ctx.set_tuning(**saved)
    for i in range(len(spec)):
        assert np.array_equal(b4[i], b8[i]), f"job {i} differs"
    for s in sids:
        ctx.release(s)


def test_progressive_batches_equal_one_shot(ctx):
    """PathTracer's progressive loop (samples min..max step) gives the same
    fixed-point sums as one call over the whole range."""
    import omr

    scene = omr.procedural(0, width=32, height=32)
    sid = ctx.upload(scene)
    full = ctx.render_sums(sid, scene.camera, omr.params(ss=2, spp_begin=0, spp_end=4, seed=9))
    parts = sum(ctx.render_sums(sid, scene.camera, omr.params(ss=2, spp_begin=k, spp_end=k + 1, seed=9))
                for k in range(4))
    assert np.array_equal(full, parts)
    ctx.release(sid)


@pytest.mark.parametrize("record_mb,solo_min", [(None, None), (0, None), (None, 1024)])
def test_served_equals_alone(record_mb, solo_min):
    """Requests submitted while others run; with solo_min, requests with
    1024 or more samples left end in placed rounds of their own."""
    import omr

    ctx = omr.Context(0)
    if record_mb is not None:
        ctx.set_tuning(record_mb=record_mb)
    if solo_min is not None:
        ctx.set_tuning(solo_round_min=solo_min)
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    jobs = []
    for si, w, h, ss, b, e, d, rr, seed in spec:
        jobs.append((sids[si], _cam(omr, scenes[si], w, h),
                     omr.params(ss=ss, spp_begin=b, spp_end=e, max_depth=d, rr_depth=rr, seed=seed)))
    alone = [ctx.render(*j) for j in jobs]
    pool_before = ctx.get_tuning()["pool_paths"]
    ctx.server_start(pool_paths=1 << 16)
    outs = [np.zeros_like(a) for a in alone]
    ids = []
    for k, (j, o) in enumerate(zip(jobs, outs)):
        ids.append(ctx.submit(j[0], j[1], j[2], o))
        if k == 2:
            # later requests arrive while the first ones are admitted or running
            while ctx.poll(ids[0]) == 0:
                time.sleep(0.001)
    print("submitted", len(ids), flush=True)
    for i in ids:
        assert ctx.wait(i, 60000) == 0
    st = ctx.server_stats()
    assert st["jobs_done"] == len(jobs) and st["lat_p99_ms"] > 0
    assert st["used_bytes"] == 0 and st["fixed_bytes"] > 0
    ctx.server_stop()
    # the server's pool size does not outlive it
    assert ctx.get_tuning()["pool_paths"] == pool_before
    for a, o in zip(alone, outs):
        assert np.array_equal(a, o)
    ctx.close()


def test_served_budget_and_scene_holds():
    """The HBM budget covers the pool and scenes: a start whose pool alone
    exceeds it fails; a request whose image cannot fit in what is left is
    rejected (state 5, OMR_ERR_BUDGET, with a message).  A scene held by a
    queued or running request cannot be released until the request ends."""
    import omr

    ctx = omr.Context(0)
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    with pytest.raises(omr.OmrError) as e:
        ctx.server_start(pool_paths=1 << 12, hbm_budget_bytes=1024)
    assert e.value.code == omr.ERR_BUDGET
    ctx.server_start(pool_paths=1 << 12)
    fixed = ctx.server_stats()["fixed_bytes"]
    ctx.server_stop()
    ctx.server_start(pool_paths=1 << 12, hbm_budget_bytes=fixed + 4096)
    big = np.zeros((48, 64, 3), np.float32)
    cam = _cam(omr, scenes[0], 64, 48)
    prm = omr.params(ss=1, spp_end=1, seed=3)
    with pytest.raises(omr.OmrError) as e:
        ctx.submit(sids[0], cam, prm, big)
    assert e.value.code == omr.ERR_BUDGET
    st = ctx.server_stats()
    assert st["jobs_rejected"] == 1
    ctx.server_stop()
    # a held scene: release is refused while its request is queued or running
    ctx.server_start(pool_paths=1 << 16)
    out = np.zeros((64, 80, 3), np.float32)
    cam1 = _cam(omr, scenes[1], 80, 64)
    job = ctx.submit(sids[1], cam1, omr.params(ss=4, spp_end=8, max_depth=16, seed=8), out)
    with pytest.raises(omr.OmrError) as e:
        ctx.release(sids[1])
    assert e.value.code == omr.ERR_STATE
    assert ctx.wait(job, 60000) == 0
    ctx.server_stop()
    ctx.release(sids[1])
    ctx.close()


def test_served_out_of_order_completion():
    """Requests that finish behind an unfinished earlier one (short
    environment-lit jobs queued after a long one whose deepest paths outlast
    them) stay in the live window until it passes them; the loop keeps
    serving and every result equals the job alone."""
    import omr

    ctx = omr.Context(0)
    scene = omr.procedural(3, width=64, height=48)
    sid = ctx.upload(scene)
    spec = [(320, 240, 4, 4, 16), (24, 16, 1, 1, 2), (32, 18, 2, 1, 2), (16, 9, 1, 1, 1), (48, 27, 2, 2, 2)]
    jobs = [(sid, _cam(omr, scene, w, h), omr.params(ss=ss, spp_end=spp, max_depth=d, env_is=1, seed=40 + k))
            for k, (w, h, ss, spp, d) in enumerate(spec)]
    alone = [ctx.render(*j) for j in jobs]
    ctx.server_start(pool_paths=1 << 16)
    outs = [np.zeros_like(a) for a in alone]
    ids = [ctx.submit(j[0], j[1], j[2], o) for j, o in zip(jobs, outs)]
    print("submitted", len(ids), flush=True)
    for i in ids:
        assert ctx.wait(i, 60000) == 0, ctx.job_error(i)
    lat = [ctx.job_latency(i) for i in ids]
    print("latencies", lat, flush=True)
    ctx.server_stop()
    for a, o in zip(alone, outs):
        assert np.array_equal(a, o)
    ctx.close()
