"""Depth-map fusion: the oracle pinned by closed forms (CPU).

A fronto-parallel depth map z = 4 seen by a camera at the origin must fuse
into points (z (x - cx)/f, z (y - cy)/f, z); the same map seen by a rotated,
translated camera must land on the plane R^T (z-plane - t); camera normals
(0, 0, -1) must become the world normal -R^T e_z; colours are the 8-bit value
/ 255; invalid and low-confidence pixels are skipped; order is view then
row-major.
"""
import math

import numpy as np


def _rot_y(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def test_fuse_oracle_closed_forms(orc):
    H, W, f = 6, 8, 10.0
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1.0]])
    depth = np.full((H, W), 4.0, np.float32)
    depth[0, 0] = 0.0                                     # invalid
    conf = np.ones((H, W), np.float32)
    conf[1, 2] = 0.2                                      # below the threshold
    normal = np.zeros((H, W, 3), np.float32)
    normal[..., 2] = -1.0
    gray = (np.arange(H * W) % 256).astype(np.uint8).reshape(H, W)
    m0 = dict(depth=depth, normal=normal, confidence=conf, color=gray, K=K, R=np.eye(3),
              t=np.zeros(3))
    R1, t1 = _rot_y(30.0), np.array([0.5, -0.2, 1.0])
    rgb = np.stack([gray, 255 - gray, gray // 2], axis=-1)
    m1 = dict(depth=depth, normal=normal, color=rgb, K=K, R=R1, t=t1)
    pos, nrm, col = orc.fuse([m0, m1], min_confidence=0.5)
    n0 = H * W - 2
    assert len(pos) == n0 + H * W - 1
    ys, xs = np.nonzero((depth > 0) & (conf >= 0.5))      # row-major order
    want0 = np.stack([4.0 * (xs - W / 2) / f, 4.0 * (ys - H / 2) / f, np.full(len(xs), 4.0)], 1)
    np.testing.assert_allclose(pos[:n0], want0, atol=2e-6)
    np.testing.assert_array_equal(nrm[:n0], np.tile([0, 0, -1.0], (n0, 1)))
    np.testing.assert_array_equal(col[:n0, 0], gray[ys, xs] / np.float32(255))
    ys1, xs1 = np.nonzero(depth > 0)
    Pc = np.stack([4.0 * (xs1 - W / 2) / f, 4.0 * (ys1 - H / 2) / f, np.full(len(xs1), 4.0)], 1)
    want1 = (Pc - t1) @ R1                                # R^T (P - t), row form
    np.testing.assert_allclose(pos[n0:], want1, atol=5e-6)
    np.testing.assert_allclose(nrm[n0:], np.tile(-R1.T[:, 2], (len(xs1), 1)), atol=1e-6)
    np.testing.assert_array_equal(col[n0:, 1], (255 - gray[ys1, xs1]).astype(np.float32) / 255)
