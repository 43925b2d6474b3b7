"""Host view set-up: source-view selection and SfM seed projection (CPU).

The product (hp3d_select_source_views, hp3d_project_seed_points) counts
shared points in one pass over the observation lists; the oracle restates
the reference's per-pair scans (src/dense_recon/patchmatch_mvs.cu:41-157,
1126-1155).  Both are compared exactly, and the oracle itself is pinned by
closed forms: a ring of cameras whose baseline angles are known, depth
percentiles of points placed on the optical axis, and the projection of a
point straight ahead of the camera.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def turntable(scenes):
    return scenes.turntable(threads=8)


def ring(scenes, angles_deg, radius=4.0, center_z=4.0, f=500.0, W=640, H=480):
    K = scenes.pinhole(f, W, H)
    return [dict(K=K, R=R, t=t) for R, t in
            (scenes.arc_camera(a, center_z=center_z, radius=radius) for a in angles_deg)]


@pytest.mark.parametrize("num_sources", [1, 5, 8])
def test_select_matches_oracle_turntable(pkg, orc, turntable, num_sources):
    prod = pkg.SfM(turntable["cameras"], turntable["points"]).select_source_views(num_sources)
    ref = orc.SfM(turntable["cameras"], turntable["points"]).select_source_views(num_sources)
    assert len(prod) == 36
    assert prod == ref
    # reference test gates (tests/test_patchmatch.cpp:72-118)
    for vp in prod:
        assert 0 < len(vp["sources"]) <= num_sources
        for sid, score, deg in vp["sources"]:
            assert 3.0 <= deg <= 120.0 and score > 0.0
        assert vp["depth_min"] > 0 and 1.1 < vp["depth_max"] / vp["depth_min"] < 10.0


def test_select_matches_oracle_random(pkg, orc, scenes):
    rng = np.random.default_rng(5)
    cams = ring(scenes, rng.uniform(-120, 120, 14))
    for i in (3, 9):
        cams[i]["registered"] = False
    pts = []
    for _ in range(600):
        X = np.array([0.0, 0.0, 4.0]) + rng.normal(0, 0.8, 3)
        ids = rng.choice(14, size=int(rng.integers(0, 6)), replace=False)
        pts.append((X, list(ids) + list(ids[:1])))   # a repeated id counts once
    for ns in (2, 5, 8):
        a = pkg.SfM(cams, pts).select_source_views(ns)
        b = orc.SfM(cams, pts).select_source_views(ns)
        assert a == b
        assert [v["ref_image_id"] for v in a] == [i for i in range(14) if i not in (3, 9)]


def test_select_closed_form_ring(pkg, orc, scenes):
    """Cameras at 0, 2, 10, 20, ..., 100 deg on a circle around the points'
    centroid: the baseline angle of camera k to camera 0 is its ring angle;
    2 deg (< 5) and 100 deg (> 90) are rejected; every point is seen by every
    view, so score = N sin(clamp(angle, 5, 30) deg) and ties (30..90 deg)
    keep camera order."""
    angles = [0, 2, 10, 20, 30, 40, 60, 90, 100]
    cams = ring(scenes, angles)
    n = 40
    pts = [(np.array([0.0, 0.0, 4.0]) + d, list(range(len(angles))))
           for d in (np.array([0.1, 0.0, 0.0]), np.array([-0.1, 0.0, 0.0]))
           for _ in range(n // 2)]
    for sfm in (pkg.SfM(cams, pts), orc.SfM(cams, pts)):
        vp = sfm.select_source_views(8)[0]
        ids = [s[0] for s in vp["sources"]]
        assert ids == [4, 5, 6, 7, 3, 2]
        for sid, score, deg in vp["sources"]:
            assert abs(deg - angles[sid]) < 1e-3
            want = n * math.sin(math.radians(min(30.0, max(5.0, angles[sid]))))
            assert abs(score - want) <= 1e-4 * want


def test_depth_range_percentiles_and_fallback(pkg, orc, scenes):
    cams = ring(scenes, [0, 20])
    # 100 points on camera 0's optical axis at z = 1..100, seen by both views
    pts = [(np.array([0.0, 0.0, float(z)]), [0, 1]) for z in range(1, 101)]
    for sfm in (pkg.SfM(cams, pts), orc.SfM(cams, pts)):
        vp = sfm.select_source_views(5)[0]
        # sorted depths d[i] = i + 1: P5 index 5 -> 6, P95 index 95 -> 96
        assert vp["depth_min"] == np.float32(6.0) * np.float32(0.5)
        assert vp["depth_max"] == np.float32(96.0) * np.float32(1.5)
    # one visible point: the centroid's depth x 0.3 and x 3
    pts1 = [(np.array([0.0, 0.0, 7.0]), [0, 1])]
    for sfm in (pkg.SfM(cams, pts1), orc.SfM(cams, pts1)):
        vp = sfm.select_source_views(5)[0]
        assert vp["depth_min"] == np.float32(7.0) * np.float32(0.3)
        assert vp["depth_max"] == np.float32(7.0) * np.float32(3.0)


def test_seed_projection(pkg, orc, scenes, turntable):
    cams = ring(scenes, [0, 20])
    pts = [(np.array([0.0, 0.0, 2.0]), [0, 1]),      # straight ahead: principal point
           (np.array([0.0, 0.0, 50.0]), [0, 1]),     # outside the depth range
           (np.array([0.0, 0.0, -2.0]), [0, 1]),     # behind the camera
           (np.array([0.4, -0.2, 4.0]), [1]),        # not observed by camera 0
           (np.array([0.4, -0.2, 4.0]), [1, 0])]
    arr, n = pkg.SfM(cams, pts).seed_points(0, 1.0, 10.0, 640, 480)
    got = [(arr[i].x, arr[i].y, arr[i].depth, tuple(arr[i].normal)) for i in range(n)]
    assert got == orc.SfM(cams, pts).seeds(0, 1.0, 10.0)
    assert got[0][:3] == (320, 240, 2.0) and got[0][3] == (0.0, 0.0, -1.0)
    x = 320 + 500 * 0.4 / 4.0
    y = 240 - 500 * 0.2 / 4.0
    assert len(got) == 2 and got[1][:2] == (round(x), round(y))
    nrm = np.array([-0.4, 0.2, -4.0]) / np.linalg.norm([0.4, 0.2, 4.0])
    assert np.allclose(got[1][3], nrm, atol=1e-7)
    # full turntable, every camera: the product keeps exactly the oracle's
    # seeds whose square touches the image, in the same order
    sp = pkg.SfM(turntable["cameras"], turntable["points"])
    so = orc.SfM(turntable["cameras"], turntable["points"])
    for v in (0, 7, 23):
        arr, n = sp.seed_points(v, 0.1, 0.4, 1280, 960, seed_radius=50)
        got = [(arr[i].x, arr[i].y, arr[i].depth, tuple(arr[i].normal)) for i in range(n)]
        ref = [s for s in so.seeds(v, 0.1, 0.4)
               if -50 <= s[0] <= 1279 + 50 and -50 <= s[1] <= 959 + 50]
        assert got == ref and n > 100


def test_apply_seeds_oracle_order():
    """Later seeds overwrite earlier ones inside their squares (.cu:1157-1169)."""
    import __graft_entry__ as entry
    orc = entry.load_oracle()
    d = np.zeros((20, 30), np.float32)
    n = np.zeros((20, 30, 3), np.float32)
    orc.apply_seeds(d, n, [(5, 5, 1.0, (0, 0, -1)), (8, 5, 2.0, (0, 1, 0)),
                           (-3, 18, 3.0, (1, 0, 0))], radius=3)
    assert d[5, 4] == 1.0 and d[5, 5] == 2.0 and d[2, 11] == 2.0 and d[5, 12] == 0.0
    assert d[19, 0] == 3.0 and d[15, 0] == 3.0 and d[14, 0] == 0.0 and d[19, 1] == 0.0
    assert tuple(n[5, 5]) == (0, 1, 0)


def test_view_setup_rejects_bad_arguments(pkg, scenes):
    cams = ring(scenes, [0, 20])
    pts = [(np.array([0.0, 0.0, 2.0]), [0, 1])]
    sfm = pkg.SfM(cams, pts)
    for ns in (0, 9):
        with pytest.raises(pkg.HP3DError):
            sfm.select_source_views(ns)
    sfm.cams[1].image_id = 0
    with pytest.raises(pkg.HP3DError):
        sfm.select_source_views(5)
    with pytest.raises(pkg.HP3DError):
        pkg.SfM(cams, pts).seed_points(0, 1.0, 10.0, 0, 480)
