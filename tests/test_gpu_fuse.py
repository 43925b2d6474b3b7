"""Depth-map fusion on the device against the oracle (MI355X): bit-identical
positions, normals and colours, in view then row-major order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fuse_matches_oracle(pkg, orc, ctx, scenes):
    tt = scenes.turntable(n_views=6, W=320, H=240, f=500.0, threads=8)
    rng = np.random.default_rng(8)
    maps = []
    for v in range(6):
        d = tt["gt_depth"][v].copy()
        n = rng.normal(size=d.shape + (3,)).astype(np.float32)
        conf = rng.uniform(0, 1, d.shape).astype(np.float32)
        cam = tt["cameras"][v]
        maps.append(dict(depth=d, normal=n, confidence=conf, color=tt["images"][v], K=cam["K"],
                         R=cam["R"], t=cam["t"]))
    for mc in (0.0, 0.5):
        p, nn, c = ctx.fuse(maps, min_confidence=mc)
        op, on, oc = orc.fuse(maps, min_confidence=mc)
        assert len(p) == len(op) > 1000
        assert np.array_equal(p, op) and np.array_equal(nn, on) and np.array_equal(c, oc)
    # points of the ground-truth maps lie on the sphere (radius 0.05 at (0, 0, 0.25))
    p, _, _ = ctx.fuse(maps)
    r = np.linalg.norm(p - np.array([0.0, 0.0, 0.25], np.float32), axis=1)
    assert np.abs(r - 0.05).max() < 1e-4
    # count-only call and an empty result
    empty = [dict(maps[0], depth=np.zeros_like(maps[0]["depth"]))]
    p, _, _ = ctx.fuse(empty)
    assert len(p) == 0
