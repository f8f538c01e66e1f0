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
