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
