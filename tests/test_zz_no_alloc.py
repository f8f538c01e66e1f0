"""The allocation-free closest-hit entry point against the allocating one."""
import numpy as np
import pytest

import helpers

pytestmark = pytest.mark.gpu


def _run(ctx, sid, rays, noalloc):
    n = rays.shape[0]
    d_rays = helpers.DeviceArray(rays.reshape(-1))
    d_prim = helpers.DeviceArray(np.full(n, -7, np.int32))
    d_t = helpers.DeviceArray(np.full(n, -1.0, np.float32))
    try:
        if noalloc:
            ctx.trace_device_noalloc(sid, d_rays.address, n, d_prim.address, d_t.address)
        else:
            ctx.trace_device(sid, d_rays.address, n, d_prim.address, d_t.address)
        return d_prim.get(), d_t.get()
    finally:
        for b in (d_rays, d_prim, d_t):
            b.free()


def test_noalloc_trace_matches_allocating_entry_point(ctx):
    """Two calls in a row (the second on a smaller ray set and another scene)
    give the ids and distances of omr_trace_closest_device bit for bit."""
    import omr

    big = omr.procedural(2, width=64, height=36)   # 1M-triangle terrain: the deepest stack bound
    small = omr.procedural(1, width=64, height=64)
    s_big, s_small = ctx.upload(big), ctx.upload(small)
    for sid, scene, n, seed in ((s_big, big, 300_000, 3), (s_small, small, 40_000, 4)):
        rays = helpers.hit_rays(scene, n, seed=seed)
        print(f"noalloc trace n={n}", flush=True)
        ap, at = _run(ctx, sid, rays, noalloc=False)
        np_, nt = _run(ctx, sid, rays, noalloc=True)
        assert np.array_equal(ap, np_), f"{int((ap != np_).sum())} ids differ"
        assert np.array_equal(at.view(np.uint32), nt.view(np.uint32))
        assert (np_ >= 0).mean() > 0.2
    # the empty ray set and an oversized count
    ctx.trace_device_noalloc(s_small, 0, 0, 0, 0)
    with pytest.raises(omr.OmrError):
        ctx.trace_device_noalloc(s_small, 8, 1 << 32, 8, 8)
    ctx.release(s_big)
    ctx.release(s_small)
