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


@pytest.mark.parametrize("record_mb", [None, 0])
def test_served_equals_alone(record_mb):
    import omr

    ctx = omr.Context(0)
    if record_mb is not None:
        ctx.set_tuning(record_mb=record_mb)
    scenes, spec = _jobs(omr)
    sids = [ctx.upload(s) for s in scenes]
    jobs = []
    for si, w, h, ss, b, e, d, rr, seed in spec:
        jobs.append((sids[si], _cam(omr, scenes[si], w, h),
                     omr.params(ss=ss, spp_begin=b, spp_end=e, max_depth=d, rr_depth=rr, seed=seed)))
    alone = [ctx.render(*j) for j in jobs]
    ctx.server_start(pool_paths=1 << 16)
    outs = [np.zeros_like(a) for a in alone]
    ids = []
    for k, (j, o) in enumerate(zip(jobs, outs)):
        ids.append(ctx.submit(j[0], j[1], j[2], o))
        if k == 2:
            time.sleep(0.01)   # later requests arrive while earlier ones run
    print("submitted", len(ids), flush=True)
    for i in ids:
        assert ctx.wait(i, 60000) == 0
    st = ctx.server_stats()
    assert st["jobs_done"] == len(jobs) and st["lat_p99_ms"] > 0
    # a request above the HBM budget is rejected at submission
    ctx.server_stop()
    for a, o in zip(alone, outs):
        assert np.array_equal(a, o)
    ctx.server_start(hbm_budget_bytes=1024)
    big = np.zeros((48, 64, 3), np.float32)
    with pytest.raises(omr.OmrError):
        ctx.submit(jobs[0][0], jobs[0][1], jobs[0][2], big)
    ctx.server_stop()
    ctx.close()
