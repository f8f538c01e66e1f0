"""Edge cases of the C ABI on the GPU: degenerate and ragged sizes give the
oracle's bits, and invalid arguments fail with an error code instead of
running (the reference's own conventions: bool returns, no exceptions)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cam(omr, scene, w, h):
    c = omr.Camera()
    C.memmove(C.byref(c), C.byref(scene.camera), C.sizeof(omr.Camera))
    c.width, c.height = w, h
    return c


@pytest.mark.parametrize("w,h,ss,spp", [(1, 1, 1, 1), (3, 1, 2, 1), (1, 5, 1, 3), (7, 3, 3, 2)])
def test_tiny_and_ragged_images_bit_exact(ctx, w, h, ss, spp):
    """Images smaller than one wave of samples and sizes that are no multiple
    of anything: the sample runs, the first fill and the per-sample records
    still give the oracle's sums."""
    import omr
    import pyoracle

    scene = omr.procedural(1, width=w, height=h)
    sid = ctx.upload(scene)
    cam = _cam(omr, scene, w, h)
    prm = omr.params(ss=ss, spp_end=spp, max_depth=16, rr_depth=3, seed=77 + w)
    g = ctx.render_sums(sid, cam, prm)
    c, seg, _ = pyoracle.Oracle(scene, ctx.scene_bvh(sid)).render_sums(cam, prm, pyoracle.WIDE_BVH)
    assert np.array_equal(g, c)
    assert ctx.last_stats()["segments"] == seg
    ctx.release(sid)


def test_trace_empty_ray_set(ctx):
    import omr

    scene = omr.procedural(0, width=8, height=8)
    sid = ctx.upload(scene)
    p, t = ctx.trace(sid, np.zeros((0, 8), np.float32))
    assert p.shape == (0,) and t.shape == (0,)
    ctx.release(sid)


def test_invalid_arguments_fail_cleanly(ctx):
    import omr

    scene = omr.procedural(0, width=8, height=8)
    sid = ctx.upload(scene)
    good = omr.params(ss=1, spp_end=1)
    bad_params = [omr.params(ss=0, spp_end=1), omr.params(ss=65, spp_end=1),
                  omr.params(ss=1, spp_begin=2, spp_end=2), omr.params(ss=1, spp_end=1, max_depth=256),
                  omr.params(ss=64, spp_end=1 << 20)]          # more than 2^32 samples
    for prm in bad_params:
        with pytest.raises(omr.OmrError):
            ctx.render(sid, scene.camera, prm)
    with pytest.raises(omr.OmrError):
        ctx.render(sid, _cam(omr, scene, 0, 8), good)
    with pytest.raises(omr.OmrError):
        ctx.render(sid + 1000, scene.camera, good)
    saved = ctx.get_tuning()
    for bad in (dict(wg_size=96), dict(stack_short=12), dict(record_mb=1 << 30), dict(pool_paths=16)):
        with pytest.raises(omr.OmrError):
            ctx.set_tuning(**bad)
    assert ctx.get_tuning() == saved
    # the context still works after every refusal
    out = ctx.render(sid, scene.camera, good)
    assert out.shape == (8, 8, 3) and np.isfinite(out).all()
    ctx.release(sid)
    with pytest.raises(omr.OmrError):
        ctx.render(sid, scene.camera, good)


def test_scene_slots_are_reused_under_new_ids(ctx):
    """More uploads than the 64 resident-scene slots, each released before
    the next: every id is new, a released id stays unknown after its slot
    holds another scene, and the reused slot renders its new scene."""
    import omr

    a = omr.procedural(0, width=8, height=8)
    b = omr.procedural(1, width=8, height=8)
    prm = omr.params(ss=1, spp_end=2, max_depth=4, rr_depth=3, seed=5)
    sa = ctx.upload(a)
    ref_a = ctx.render(sa, a.camera, prm)
    ctx.release(sa)
    sb = ctx.upload(b)
    ref_b = ctx.render(sb, b.camera, prm)
    ctx.release(sb)
    ids = set()
    print("cycling 80 uploads", flush=True)
    for i in range(80):
        s = a if i % 2 == 0 else b
        sid = ctx.upload(s)
        assert sid not in ids
        ids.add(sid)
        out = ctx.render(sid, s.camera, prm)
        assert np.array_equal(out, ref_a if i % 2 == 0 else ref_b), f"upload {i}"
        ctx.release(sid)
        with pytest.raises(omr.OmrError):
            ctx.render(sid, s.camera, prm)
    with pytest.raises(omr.OmrError):
        ctx.release(min(ids))
