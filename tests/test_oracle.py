"""CPU tests that pin the oracle (oracle/pm_oracle.c) before it is trusted.

How each oracle function is pinned (reference = src/dense_recon/patchmatch_mvs.cu):
  orc_philox4x32_10   Random123 published known-answer vectors (golden file).
  orc_compose         identity pose -> A = I, b = 0; pure translation with a
                      fronto-parallel plane -> known disparity f*B/z; agreement
                      with an independent numpy composition.
  orc_cost_fp64       closed forms (NCC(a, alpha*a + beta) = +-1, constant patch,
                      count < min_count, top-k mean), integer and half-pixel
                      known disparities (warp + bilinear), and a brute-force
                      evaluation written independently (ray-plane intersection
                      and projection instead of A/b, np.corrcoef for NCC).
  orc_cost_fp32       agreement with orc_cost_fp64 within the reference's own
                      FP32 rounding (its one-pass variance is not shifted).
  orc_init_fp32       reference test invariants (InitKernelDepthRange,
                      InitKernelNormals, test_patchmatch.cpp:120-190) plus the
                      camera-facing rule and recomputation of the Philox draws.
  orc_halfpass_fp32   non-increasing cost, only one colour written, result
                      independent of thread count / order, refinement steps
                      bounded by (dmax-dmin)/2 * 0.6^i * (1 + 1/2).
  orc_patchmatch_*    known answer: converges to the analytic plane depth;
                      PropagationReducesCost analogue (test_patchmatch.cpp:356-407).
  orc_geom_filter     consistent GT maps retained, inconsistent maps removed.
  orc_select_source_views, orc_project_seeds, orc_apply_seeds
                      tests/test_views.py: a camera ring with known baseline
                      angles and scores, depth percentiles of points on the
                      optical axis, the centroid fallback, projection of a
                      point straight ahead, overwrite order of seeds.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import oracle_view, rel_within

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# Philox and u01
# ---------------------------------------------------------------------------
def test_philox_known_answer_vectors(orc):
    kat = json.load(open(os.path.join(GOLDEN, "philox4x32_10_kat.json")))
    for v in kat["vectors"]:
        ctr = [int(h, 16) for h in v["ctr"]]
        key = [int(h, 16) for h in v["key"]]
        out = orc.philox(ctr, key)
        assert [f"{x:08x}" for x in out] == v["out"]


def test_u01_maps_into_half_open_unit_interval(orc):
    assert orc.u01(0) == 2.0 ** -24
    assert orc.u01(255) == 2.0 ** -24          # low 8 bits dropped
    assert orc.u01(0xFFFFFFFF) == 1.0
    assert orc.u01(0x80000000) == pytest.approx(0.5 + 2.0 ** -24, abs=0)


# ---------------------------------------------------------------------------
# Composition of A, b
# ---------------------------------------------------------------------------
def _rot(ax, ay, az):
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def test_compose_identity_pose(orc):
    K = np.array([[300.0, 0, 160], [0, 310.0, 120], [0, 0, 1]])
    R = _rot(0.1, -0.2, 0.05)
    t = np.array([0.3, -0.1, 2.0])
    A, b = orc.compose(K, R, t, K, R, t)
    np.testing.assert_allclose(A, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(b, 0.0, atol=1e-12)


def test_compose_matches_independent_composition(orc):
    rng = np.random.default_rng(3)
    for _ in range(10):
        Kr = np.array([[rng.uniform(200, 2000), 0, rng.uniform(100, 900)],
                       [0, rng.uniform(200, 2000), rng.uniform(100, 700)], [0, 0, 1]])
        Ks = Kr.copy()
        Ks[0, 0] *= rng.uniform(0.9, 1.1)
        Rr, Rs = _rot(*rng.normal(0, 0.3, 3)), _rot(*rng.normal(0, 0.3, 3))
        tr, ts = rng.normal(0, 1, 3), rng.normal(0, 1, 3)
        A, b = orc.compose(Kr, Rr, tr, Ks, Rs, ts)
        Rrel = Rs @ Rr.T
        np.testing.assert_allclose(A, Ks @ Rrel @ np.linalg.inv(Kr), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(b, Ks @ (ts - Rrel @ tr), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------------------
# Closed-form cost cases
# ---------------------------------------------------------------------------
def _identity_view(orc, ref, srcs, K, **kw):
    """All sources share the reference camera: the warp is the identity."""
    Kinv = np.linalg.inv(K)
    A = [np.eye(3)] * len(srcs)
    b = [np.zeros(3)] * len(srcs)
    return orc.View(ref, srcs, Kinv, A, b, 2.0, 10.0, **kw)


def _texture(H, W, lo, hi, seed):
    return np.random.default_rng(seed).integers(lo, hi + 1, size=(H, W)).astype(np.uint8)


K_SMALL = np.array([[50.0, 0, 16], [0, 50.0, 12], [0, 0, 1]])
FRONT = (0.0, 0.0, -1.0)


def test_ncc_affine_images_give_cost_0_and_2(orc):
    ref = _texture(24, 32, 20, 100, 0)
    up = (2 * ref.astype(int) + 10).astype(np.uint8)       # alpha > 0
    down = (255 - ref.astype(int)).astype(np.uint8)          # alpha < 0
    v = _identity_view(orc, ref, [up, down], K_SMALL, radius=3, top_k=3)
    c, f, per = v.cost_fp64(12, 10, 5.0, FRONT, per_src=True)
    assert per[0] == pytest.approx(0.0, abs=1e-12)
    assert per[1] == pytest.approx(2.0, abs=1e-12)
    assert c == pytest.approx(1.0, abs=1e-12)               # mean of min(3, S=2) entries
    assert f == 0


def test_constant_source_patch_costs_2(orc):
    ref = _texture(24, 32, 20, 100, 1)
    flat = np.full_like(ref, 128)
    v = _identity_view(orc, ref, [flat], K_SMALL, radius=2, top_k=3)
    c, _, per = v.cost_fp64(10, 10, 5.0, FRONT, per_src=True)
    assert per[0] == 2.0 and c == 2.0


def test_min_count_rule_at_image_corner(orc):
    ref = _texture(24, 32, 20, 200, 2)
    v = _identity_view(orc, ref, [ref.copy()], K_SMALL, radius=5, top_k=3)
    # corner: taps with dx, dy >= 0 only -> 36 valid taps < floor(121 / 2) = 60
    assert v.cost_fp64(0, 0, 5.0, FRONT)[0] == 2.0
    # x = 2: dx in [-2, 5] -> 8 x 6 (dy in [0, 5]) = 48 < 60 -> still 2
    assert v.cost_fp64(2, 0, 5.0, FRONT)[0] == 2.0
    # x = 2, y = 5: 8 x 11 = 88 >= 60 -> identical images -> 0
    assert v.cost_fp64(2, 5, 5.0, FRONT)[0] == pytest.approx(0.0, abs=1e-12)
    # last column is excluded by u < W - 1: x = W-1 keeps dx < 0 only
    assert v.cost_fp64(31, 12, 5.0, FRONT)[0] == 2.0        # 5 x 11 = 55 < 60


def test_top_k_mean_of_smallest_three(orc):
    ref = _texture(24, 32, 20, 100, 4)
    same = ref.copy()
    anti = (255 - ref.astype(int)).astype(np.uint8)
    flat = np.full_like(ref, 50)
    scaled = (2 * ref.astype(int) + 10).astype(np.uint8)
    v = _identity_view(orc, ref, [same, anti, flat, scaled], K_SMALL, radius=3, top_k=3)
    c, _, per = v.cost_fp64(12, 12, 5.0, FRONT, per_src=True)
    np.testing.assert_allclose(per, [0.0, 2.0, 2.0, 0.0], atol=1e-12)
    assert c == pytest.approx(np.sort(per)[:3].mean(), abs=1e-12)
    assert c == pytest.approx(2.0 / 3.0, abs=1e-12)


def _stereo_view(orc, ref, src, f, B, radius=3, stride=1):
    """Source camera displaced by +B along x (R = I): u = x - f B / z."""
    H, W = ref.shape
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]])
    A, b = orc.compose(K, np.eye(3), np.zeros(3), K, np.eye(3), np.array([-B, 0.0, 0.0]))
    return orc.View(ref, [src], np.linalg.inv(K), [A], [b], 1.0, 100.0, radius=radius,
                    stride=stride, top_k=3), K


def test_integer_disparity_known_answer(orc):
    ref = _texture(40, 64, 30, 220, 5)
    d = 3
    src = np.zeros_like(ref)
    src[:, :-d] = ref[:, d:]                 # src(x') = ref(x' + d)
    src[:, -d:] = 17
    f, B = 60.0, 0.5
    v, _ = _stereo_view(orc, ref, src, f, B)
    z = f * B / d
    c_true, flags, _ = v.cost_fp64(30, 20, z, FRONT, per_src=True)
    assert c_true == pytest.approx(0.0, abs=1e-12) and flags == 0
    for zz in (z * 0.8, z * 1.3):
        assert v.cost_fp64(30, 20, zz, FRONT)[0] > 0.05


def test_half_pixel_disparity_is_bilinear_average(orc):
    # src values of one parity so that neighbour averages are integers
    rng = np.random.default_rng(6)
    src = (2 * rng.integers(15, 110, size=(40, 64))).astype(np.uint8)
    ref = np.zeros_like(src)
    # ref(x) = (src(x - 1) + src(x)) / 2 : a warp u = x - 0.5 samples exactly that
    ref[:, 1:] = ((src[:, :-1].astype(int) + src[:, 1:].astype(int)) // 2).astype(np.uint8)
    f, B = 60.0, 0.25
    v, _ = _stereo_view(orc, ref, src, f, B)
    z = f * B / 0.5
    assert v.cost_fp64(30, 20, z, FRONT)[0] == pytest.approx(0.0, abs=1e-12)
    # nearest-texel sampling would not match; the true half-pixel shift must
    # beat a whole-pixel one
    assert v.cost_fp64(30, 20, f * B / 1.0, FRONT)[0] > 0.02


def test_stride_two_window_uses_36_taps(orc):
    ref = _texture(30, 40, 20, 200, 7)
    v = _identity_view(orc, ref, [ref.copy()], K_SMALL, radius=5, stride=2, top_k=3)
    # stride 2 -> taps at -5,-3,...,5: 6 x 6 = 36, min_count 18.  At x = 1 the
    # taps dx in {-5,-3} fall outside (u < 0): 4 x 6 = 24 >= 18 -> match -> 0
    assert v.cost_fp64(1, 15, 5.0, FRONT)[0] == pytest.approx(0.0, abs=1e-12)
    # at x = 0, y = 0: dx, dy in {1,3,5} only -> 9 < 18 -> 2
    assert v.cost_fp64(0, 0, 5.0, FRONT)[0] == 2.0


# ---------------------------------------------------------------------------
# Brute force: an independent evaluation of the reference cost
# ---------------------------------------------------------------------------
def brute_cost(ref, srcs, K, cams, x, y, d, n, radius, stride, top_k):
    """Reference cost (patchmatch_mvs.cu:222-343) computed another way: each tap's
    ray is intersected with the hypothesis plane in the reference camera frame
    and projected into the source; NCC via np.corrcoef (two-pass Pearson)."""
    Kinv = np.linalg.inv(K)
    n = np.asarray(n, dtype=np.float64)
    P = float(d) * (Kinv @ np.array([x, y, 1.0]))
    nP = n @ P
    if abs(nP) < 1e-8:
        return 2.0
    H, W = ref.shape
    R0, t0 = cams[0]
    offs = -radius + stride * np.arange(2 * radius // stride + 1)
    k = len(offs) ** 2
    costs = []
    # 8-bit images hold DN (gray = DN / 255), float images gray itself
    unit = 1.0 if np.issubdtype(ref.dtype, np.floating) else 255.0
    for Rs, ts in cams[1:]:
        src = srcs[len(costs)].astype(np.float64) / unit
        rs, ss = [], []
        for dy in offs:
            for dx in offs:
                p = np.array([x + dx, y + dy, 1.0])
                ray = Kinv @ p
                X = ray * nP / (n @ ray)                     # on the plane, ref frame
                Xs = Rs @ (R0.T @ (X - t0)) + ts
                q = K @ Xs
                wz = q[2] * (n @ ray) / nP                   # the reference's w_z
                if abs(wz) < 1e-8:
                    continue
                u, v = q[0] / q[2], q[1] / q[2]
                if not (0 <= u < W - 1 and 0 <= v < H - 1):
                    continue
                x0, y0 = int(math.floor(u)), int(math.floor(v))
                ax, ay = u - x0, v - y0
                s = ((1 - ax) * (1 - ay) * src[y0, x0] + ax * (1 - ay) * src[y0, x0 + 1]
                     + (1 - ax) * ay * src[y0 + 1, x0] + ax * ay * src[y0 + 1, x0 + 1])
                rs.append(float(ref[min(max(y + dy, 0), H - 1), min(max(x + dx, 0), W - 1)]) / unit)
                ss.append(s)
        c = len(rs)
        if c < k // 2:
            costs.append(2.0)
            continue
        rs, ss = np.array(rs), np.array(ss)
        if c * c * rs.var() < 1e-6 or c * c * ss.var() < 1e-6:
            costs.append(2.0)
            continue
        ncc = np.corrcoef(rs, ss)[0, 1]
        costs.append(1.0 - min(1.0, max(-1.0, ncc)))
    kk = min(top_k, len(costs))
    return float(np.sort(costs)[:kk].mean())


def test_cost_fp64_matches_brute_force(orc, scenes):
    job = scenes.sphere_job(48, 36, 3, scene_seed=11, noise_sigma=2.0, radius=3, stride=1,
                            iterations=0)
    v = oracle_view(orc, job)
    rng = np.random.default_rng(8)
    d0, n0 = v.init_fp32(99)
    checked = 0
    for _ in range(60):
        x, y = int(rng.integers(0, 48)), int(rng.integers(0, 36))
        for d, n in ((job.gt_depth[y, x], job.gt_normal[y, x]), (d0[y, x], n0[y, x])):
            c, flags = v.cost_fp64(x, y, d, n)
            if flags:
                continue
            b = brute_cost(job.ref, job.srcs, job.K, job.cams, x, y, d, n, 3, 1, 3)
            assert c == pytest.approx(b, abs=1e-9), (x, y, d, n)
            checked += 1
    assert checked > 80


def test_cost_fp64_matches_brute_force_stride2_and_topk(orc, scenes):
    job = scenes.sphere_job(40, 30, 5, scene_seed=12, radius=5, stride=2, iterations=0)
    for top_k in (1, 3, 5):
        v = oracle_view(orc, job, top_k=top_k)
        for x, y in ((20, 15), (5, 25), (33, 4)):
            c, flags = v.cost_fp64(x, y, job.gt_depth[y, x], job.gt_normal[y, x])
            if not flags:
                b = brute_cost(job.ref, job.srcs, job.K, job.cams, x, y, job.gt_depth[y, x],
                               job.gt_normal[y, x], 5, 2, top_k)
                assert c == pytest.approx(b, abs=1e-9)


def _float_gray(img, seed):
    """Float gray images with sub-DN detail that 8-bit levels cannot hold."""
    rng = np.random.default_rng(seed)
    return ((img.astype(np.float64) + rng.uniform(0.0, 1.0, img.shape)) / 256.0).astype(np.float32)


def test_float_images_cost_fp64_matches_brute_force(orc, scenes):
    """The oracle's float-image planes (gray given directly) against the
    independent evaluation on the same float32 values."""
    job = scenes.sphere_job(48, 36, 3, scene_seed=13, noise_sigma=2.0, radius=3, stride=1,
                            iterations=0)
    ref = _float_gray(job.ref, 1)
    srcs = [_float_gray(s, 2 + i) for i, s in enumerate(job.srcs)]
    v = orc.View(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=3, stride=1)
    v8 = oracle_view(orc, job)
    rng = np.random.default_rng(10)
    d0, n0 = v8.init_fp32(7)
    checked = moved = 0
    for _ in range(60):
        x, y = int(rng.integers(0, 48)), int(rng.integers(0, 36))
        for d, n in ((job.gt_depth[y, x], job.gt_normal[y, x]), (d0[y, x], n0[y, x])):
            c, flags = v.cost_fp64(x, y, d, n)
            if flags:
                continue
            b = brute_cost(ref, srcs, job.K, job.cams, x, y, d, n, 3, 1, 3)
            assert c == pytest.approx(b, abs=1e-9), (x, y, d, n)
            c8, f8 = v8.cost_fp64(x, y, d, n)
            moved += (not f8) and abs(c - c8) > 1e-6   # the float planes are the ones read
            checked += 1
    assert checked > 80 and moved > checked // 2


def test_float_view_of_8bit_levels_runs_like_the_8bit_view(orc, c1):
    """gray = DN / 255 rounded to float32, given as float planes, is what the
    oracle forms from 8-bit planes itself: the FP32 run is bit-identical."""
    v8 = oracle_view(orc, c1)
    g = lambda a: (a.astype(np.float32) / np.float32(255.0)).astype(np.float32)
    vf = orc.View(g(c1.ref), [g(s) for s in c1.srcs], c1.Kinv, c1.A, c1.b, c1.dmin, c1.dmax,
                  radius=c1.radius, stride=c1.stride, top_k=c1.top_k)
    a = v8.patchmatch_fp32(1, 5, threads=4)
    b = vf.patchmatch_fp32(1, 5, threads=4)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_cost_fp32_reference_order_within_its_rounding(orc, c1):
    v = oracle_view(orc, c1)
    d0, n0 = v.init_fp32(42)
    rng = np.random.default_rng(9)
    errs = []
    for _ in range(200):
        x, y = int(rng.integers(0, 320)), int(rng.integers(0, 240))
        c64, flags = v.cost_fp64(x, y, d0[y, x], n0[y, x])
        if flags:
            continue
        errs.append(abs(v.cost_fp32(x, y, d0[y, x], n0[y, x]) - c64))
    errs = np.array(errs)
    # one-pass variance in [0,1] units without a shift: ~1e-4 typical, a few 1e-3 worst
    assert np.median(errs) < 2e-4 and errs.max() < 5e-3


# ---------------------------------------------------------------------------
# Random initialisation (init_depth_normal_kernel, :164-216)
# ---------------------------------------------------------------------------
def test_init_invariants(orc, c1):
    v = oracle_view(orc, c1)
    d, n = v.init_fp32(42)
    assert np.all((d > c1.dmin) & (d <= c1.dmax))          # InitKernelDepthRange
    ln = np.linalg.norm(n.astype(np.float64), axis=-1)
    assert np.mean(np.abs(ln - 1) < 0.01) == 1.0             # InitKernelNormals (>= 99 %)
    H, W = d.shape
    ys, xs = np.mgrid[0:H, 0:W]
    rays = np.stack([xs, ys, np.ones_like(xs)], -1) @ c1.Kinv.T
    assert np.all((n * rays).sum(-1) <= 1e-7)                # faces the camera
    assert abs(d.mean() - 6.0) < 0.05                        # U(2, 10]
    assert abs(np.abs(n[..., 2]).mean() - 0.5) < 0.05        # cos(phi) ~ U(0, 1]


def test_init_draws_follow_the_counter_layout(orc, c1):
    v = oracle_view(orc, c1)
    d, _ = v.init_fp32(12345)
    for x, y in ((0, 0), (319, 239), (17, 101)):
        idx = y * 320 + x
        r = orc.philox([idx, 0, 0, orc.DOMAIN_INIT], [12345, 0])
        u = np.float32(orc.u01(int(r[0])))
        expect = np.float32(np.float32(2.0) + u * np.float32(8.0))
        assert d[y, x] == expect


# ---------------------------------------------------------------------------
# Half-pass properties (propagate_and_refine_kernel, :519-655)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c1_state(orc, c1):
    v = oracle_view(orc, c1)
    d, n, c = v.patchmatch_fp32(0, 42)
    return v, d, n, c


def test_halfpass_writes_one_colour_and_never_raises_cost(c1_state):
    v, d0, n0, c0 = c1_state
    H, W = d0.shape
    red = ((np.arange(H)[:, None] + np.arange(W)[None, :]) % 2) == 0
    for parity, mask in ((0, red), (1, ~red)):
        d, n, c = d0.copy(), n0.copy(), c0.copy()
        v.halfpass_fp32(0, parity, 42, d, n, c)
        assert np.array_equal(d[~mask], d0[~mask]) and np.array_equal(c[~mask], c0[~mask])
        assert np.all(c[mask] <= c0[mask])
        assert (c[mask] < c0[mask]).mean() > 0.3


def test_halfpass_independent_of_thread_count(c1_state):
    v, d0, n0, c0 = c1_state
    outs = []
    for threads in (1, 3, 8):
        d, n, c = d0.copy(), n0.copy(), c0.copy()
        v.halfpass_fp32(1, 0, 42, d, n, c, threads=threads)
        outs.append((d, n, c))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_refinement_steps_are_bounded(orc, c1):
    """With every neighbour out of range only the two refinement steps act:
    |d' - d| <= (dmax - dmin)/2 * 0.6^i * (1 + 1/2)."""
    v = oracle_view(orc, c1)
    d0, n0, c0 = v.patchmatch_fp32(0, 42)
    H, W = d0.shape
    red = ((np.arange(H)[:, None] + np.arange(W)[None, :]) % 2) == 0
    for it in (0, 3, 7):
        d, n, c = d0.copy(), n0.copy(), c0.copy()
        d[~red] = 1000.0                                     # black neighbours skipped
        v.halfpass_fp32(it, 0, 7, d, n, c)
        step = (c1.dmax - c1.dmin) / 2 * 0.6 ** it
        moved = np.abs(d[red] - d0[red])
        assert moved.max() <= step * 1.5 * (1 + 1e-5)
        assert (moved > 0).mean() > 0.05


# ---------------------------------------------------------------------------
# Full-run known answers
# ---------------------------------------------------------------------------
def test_fp32_run_converges_on_plane(orc, c1):
    v = oracle_view(orc, c1)
    d, _, c = v.patchmatch_fp32(8, 42)
    assert rel_within(d, c1.gt_depth).mean() > 0.88
    _, _, c0 = v.patchmatch_fp32(0, 42)
    assert c.mean() < 0.5 * c0.mean()                         # PropagationReducesCost


def test_fp32_and_fp64_cost_runs_agree_once_converged(orc, c1):
    v = oracle_view(orc, c1)
    d32, _, _ = v.patchmatch_fp32(8, 42)
    d64, _, _ = v.patchmatch_fp32(8, 42, cost_mode=orc.COST_FP64)
    assert rel_within(d32, d64).mean() > 0.995


# ---------------------------------------------------------------------------
# Geometric consistency filter (:668-781)
# ---------------------------------------------------------------------------
def test_geom_filter_keeps_consistent_and_drops_random(orc, pkg, scenes):
    W, H, f = 160, 120, 150.0
    K = scenes.pinhole(f, W, H)
    scene = scenes.sphere_scene(5, f)
    geo = []
    for a in (-4.0, 0.0, 4.0):
        R, t = scenes.arc_camera(a)
        _, gd, gn = pkg.render(scene, scenes._camera(K, R, t, W, H))
        geo.append((K, R, t, gd, gn))
    Kinv = np.linalg.inv(K)
    depth = geo[1][3].copy()
    conf = np.ones_like(depth)
    orc.geom_filter_fp32(depth, geo[1][4], conf, Kinv, geo[1][1], geo[1][2], [geo[0], geo[2]])
    valid = geo[1][3] > 0
    # everything seen by both neighbours is consistent; only the ~8 px side
    # strips leaving one neighbour's frame and nearest-pixel lookups across the
    # sphere's silhouette may drop
    assert (depth[valid] > 0).mean() > 0.8
    assert (depth[10:-10, 12:-12] > 0).mean() > 0.9
    noisy = geo[1][3] * np.random.default_rng(0).uniform(0.5, 1.5, depth.shape).astype(np.float32)
    conf = np.ones_like(noisy)
    orc.geom_filter_fp32(noisy, geo[1][4], conf, Kinv, geo[1][1], geo[1][2], [geo[0], geo[2]])
    assert (noisy[valid] > 0).mean() < 0.3
    assert np.all(conf[noisy == 0] == 0)
