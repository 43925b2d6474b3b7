"""Edge cases of the plain path against the oracle (MI355X).

Ragged and tiny images (odd widths, a width that is not a multiple of the
64-pixel tile, an image exactly the size of the window), every window the
kernels are built for, 1 and 8 sources (top-3 over fewer than three
sources), flat images (every window hits the variance guard), a source
that sees nothing of the reference (every tap out of bounds), and the largest
image plane accepted (2 GiB - 128 KiB) with one row more refused.  Gates: random
init bit-identical, cost of identical hypotheses within 1e-4 of the FP64
oracle on every unflagged pixel, and runs that never raise a cost and keep
depths in range.
"""
import numpy as np
import pytest

from conftest import oracle_view

pytestmark = pytest.mark.gpu

COST_TOL = 1e-4

CASES = [  # (W, H, S, radius, stride)
    (97, 61, 2, 5, 1),
    (97, 61, 1, 5, 2),
    (65, 33, 8, 3, 1),
    (33, 17, 3, 1, 2),
    (11, 11, 2, 5, 1),
    (130, 70, 4, 4, 2),
    (45, 29, 5, 2, 1),
]


def _desc(pkg, ctx, job, store, iterations=2, A=None, b=None, top_k=3):
    ref = ctx.upload(job.ref, store)
    srcs = [ctx.upload(s, store) for s in job.srcs]
    d = pkg.make_desc(ref, srcs, job.Kinv, job.A if A is None else A, job.b if b is None else b,
                      job.dmin, job.dmax, radius=job.radius, stride=job.stride,
                      iterations=iterations, top_k=top_k, seed=job.seed)
    return d, [ref] + srcs


@pytest.mark.parametrize("W,H,S,radius,stride", CASES)
def test_edge_shapes_match_oracle(pkg, orc, ctx, scenes, W, H, S, radius, stride):
    job = scenes.sphere_job(W, H, S, f=0.9375 * max(W, 64), radius=radius, stride=stride,
                            iterations=2, scene_seed=W + H + S, name=f"edge{W}x{H}")
    v = oracle_view(orc, job)
    for store in (0, 1):
        d, handles = _desc(pkg, ctx, job, store)
        d0, n0 = ctx.init_hypotheses(d, job.shape)
        o0, on = v.init_fp32(job.seed)
        assert np.array_equal(d0, o0) and np.array_equal(n0, on)
        c = ctx.cost_eval(d, d0, n0)
        c64, flags = v.cost_map_fp64(d0, n0)
        ok = flags == 0
        err = np.abs(c.astype(np.float64) - c64)
        assert ok.mean() > 0.6
        assert err[ok].max() <= COST_TOL, (store, float(err[ok].max()))
        out = ctx.run_view(d, job.shape)
        assert np.all(out["cost"] <= c)
        assert np.all((out["depth"] >= job.dmin) & (out["depth"] <= job.dmax))
        for h in handles:
            ctx.release(h)


def test_flat_images_cost_two_everywhere(pkg, ctx, scenes):
    """Constant images: every window fails the variance guard (cost 2), so
    no candidate is ever accepted and the run returns the random init."""
    job = scenes.sphere_job(120, 80, 3, radius=5, stride=1, iterations=2, name="flat")
    job.ref = np.full_like(job.ref, 100)
    job.srcs = [np.full_like(s, 100) for s in job.srcs]
    d, handles = _desc(pkg, ctx, job, 0)
    d0, n0 = ctx.init_hypotheses(d, job.shape)
    out = ctx.run_view(d, job.shape)
    assert np.all(out["cost"] == 2.0)
    assert np.array_equal(out["depth"], d0) and np.array_equal(out["normal"], n0)
    for h in handles:
        ctx.release(h)


def test_source_that_sees_nothing(pkg, orc, ctx, scenes):
    """A source whose homography sends every tap far outside the image: its
    per-source cost is 2 (count below min_count); with one such source of
    two and top-3, the cost is the mean of the other source's cost and 2."""
    job = scenes.sphere_job(96, 64, 2, radius=3, stride=1, iterations=1, name="blind")
    A = [job.A[0], np.asarray(job.A[1]) * 1.0]
    b = [job.b[0], np.asarray(job.b[1]) + np.array([1e6, 0.0, 0.0])]
    d, handles = _desc(pkg, ctx, job, 0, A=A, b=b)
    d0, n0 = ctx.init_hypotheses(d, job.shape)
    c = ctx.cost_eval(d, d0, n0)
    one = pkg.make_desc(handles[0], handles[1:2], job.Kinv, A[:1], b[:1], job.dmin, job.dmax,
                        radius=job.radius, stride=job.stride, iterations=1, seed=job.seed)
    c1 = ctx.cost_eval(one, d0, n0)
    want = (c1.astype(np.float64) + 2.0) / 2.0
    assert np.abs(c - want).max() <= 1e-6
    v = oracle_view(orc, job)
    v.c.b[1][0] = float(b[1][0])
    c64, flags = v.cost_map_fp64(d0, n0)
    ok = flags == 0
    assert np.abs(c.astype(np.float64) - c64)[ok].max() <= COST_TOL
    for h in handles:
        ctx.release(h)


def test_pitched_uploads_equal_packed(pkg, ctx, scenes):
    """Images handed over with padded rows (any byte pitch for 8-bit, a
    multiple of 4 for float) give the same planes as packed rows."""
    import ctypes
    job = scenes.sphere_job(75, 41, 2, radius=2, stride=1, iterations=1, name="pitch")
    H, W = job.shape

    def upload_pitched(img, pad, store):
        buf = np.zeros((H, W + pad), dtype=img.dtype)
        buf[:, :W] = img
        buf[:, W:] = 255 if img.dtype == np.uint8 else 7.0      # garbage in the padding
        out = ctypes.c_uint64(0)
        fmt = pkg.FMT_U8 if img.dtype == np.uint8 else pkg.FMT_F32
        rc = pkg.lib().hp3d_pm_image_upload(ctx.h, buf.ctypes.data, W, H, buf.strides[0], fmt,
                                            store, ctypes.byref(out))
        assert rc == 0, ctx.last_error()
        return out.value

    for store in (0, 1):
        packed = [ctx.upload(job.ref, store)] + [ctx.upload(s, store) for s in job.srcs]
        d = pkg.make_desc(packed[0], packed[1:], job.Kinv, job.A, job.b, job.dmin, job.dmax,
                          radius=2, stride=1, seed=job.seed)
        d0, n0 = ctx.init_hypotheses(d, job.shape)
        want = ctx.cost_eval(d, d0, n0)
        for imgs, pad in (([job.ref] + job.srcs, 13),
                          ([(i.astype(np.float32) / 255.0) for i in [job.ref] + job.srcs], 3)):
            h = [upload_pitched(i, pad, store) for i in imgs]
            dp = pkg.make_desc(h[0], h[1:], job.Kinv, job.A, job.b, job.dmin, job.dmax,
                               radius=2, stride=1, seed=job.seed)
            got = ctx.cost_eval(dp, d0, n0)
            if imgs[0].dtype == np.uint8:
                assert np.array_equal(got, want)
            else:   # gray / 255 * 255 is within an ulp of the DN
                assert np.abs(got - want).max() <= 1e-5
            for x in h:
                ctx.release(x)
        for x in packed:
            ctx.release(x)


@pytest.mark.parametrize("top_k", [1, 2, 4, 5, 8])
def test_top_k_variants_match_oracle(pkg, orc, ctx, scenes, top_k):
    """Every top-k the boundary accepts (1..8; the reference fixes 3), with 8
    sources: the kernels keep a 3-entry list when at most 3 costs are
    averaged and an 8-entry one otherwise.  The selection and averaging are
    what this checks, on a scene chosen to be hard for FP32 moments (7x7
    windows of low contrast far from mid-gray, where unshifted FP32 moments
    put a single source's cost 2.4e-4 off FP64): with one or two sources
    averaged nothing dilutes a source's error, so this holds the kernels'
    shifted moments to the same 1e-4 gate as the workloads, and to a tenth
    of the reference's own FP32 operation order on a sample.  A 2-iteration
    run must put >= 99 % of pixels within 1 % of the oracle's depth."""
    tol = COST_TOL
    job = scenes.sphere_job(160, 120, 8, radius=3, stride=1, iterations=2, scene_seed=31,
                            name="topk")
    v = oracle_view(orc, job, top_k=top_k)
    for store in (0, 1):
        d, handles = _desc(pkg, ctx, job, store, top_k=top_k)
        d0, n0 = ctx.init_hypotheses(d, job.shape)
        c = ctx.cost_eval(d, d0, n0)
        c64, flags = v.cost_map_fp64(d0, n0)
        ok = flags == 0
        assert ok.mean() > 0.6
        err = np.abs(c.astype(np.float64) - c64)
        ys, xs = np.nonzero(ok)
        pick = np.random.default_rng(top_k).choice(len(xs), 1500, replace=False)
        ref32 = max(abs(v.cost_fp32(int(xs[i]), int(ys[i]), float(d0[ys[i], xs[i]]),
                                    n0[ys[i], xs[i]]) - c64[ys[i], xs[i]]) for i in pick)
        wy, wx = np.unravel_index(np.argmax(np.where(ok, err, 0)), err.shape)
        per_src = np.sort(np.asarray(v.cost_fp64(int(wx), int(wy), float(d0[wy, wx]), n0[wy, wx],
                                                 per_src=True)[2]))
        print(f"top_k={top_k} store={store}: max|err| {err[ok].max():.2e} at ({wx},{wy}): gpu "
              f"{c[wy, wx]:.7f} fp64 {c64[wy, wx]:.7f}, per-source fp64 {np.round(per_src, 6)}; "
              f"reference-order FP32 on a sample {ref32:.2e}")
        assert err[ok].max() <= tol, (top_k, store)
        assert err[ys[pick], xs[pick]].max() <= 0.1 * ref32
        out = ctx.run_view(d, job.shape)
        for h in handles:
            ctx.release(h)
    o64, _, _ = v.patchmatch_fp32(job.iterations, job.seed, cost_mode=orc.COST_FP64)
    agree = (np.abs(out["depth"].astype(np.float64) - o64) <= 0.01 * np.abs(o64)).mean()
    print(f"top_k={top_k}: depth within 1 % of the FP64 oracle on {agree:.4f}")
    assert agree >= 0.99


def test_largest_plane_accepted_one_row_more_refused(pkg, ctx):
    """Taps address a plane with 32-bit offsets, so planes are capped 64 KiB
    below 2 GiB: a 16384-wide FP32 image has an 8-byte-per-pixel row-pair
    plane of 131072 B per row, which fits at 16383 rows and not at 16384
    (refused before anything is allocated).  FP16 storage halves the plane,
    so 16384 rows fit there."""
    W = 16384
    big = np.zeros((W, W), np.uint8)            # untouched pages: no resident memory
    with pytest.raises(pkg.HP3DError) as e:
        ctx.upload(big, pkg.STORE_F32)
    assert e.value.code == pkg.ERR_UNSUPPORTED
    print("16384 x 16384 FP32 refused", flush=True)
    for rows, store in ((W - 1, pkg.STORE_F32), (W, pkg.STORE_F16)):
        h = ctx.upload(big[:rows], store)
        w_, h_, st, nb = ctx.image_info(h)
        assert (w_, h_, st) == (W, rows, store)
        assert nb == rows * W * (12 if store == pkg.STORE_F32 else 6)
        ctx.release(h)
        print(f"16384 x {rows} store {store} accepted ({nb} B)", flush=True)


def test_solve_on_the_largest_plane(pkg, ctx):
    """A solve over a 16384 x 16383 image, whose FP32 row-pair plane is the
    largest accepted (2 GiB - 128 KiB): with the source equal to the
    reference and the identity homography (A = I, b = 0) every tap samples
    the source at the reference pixel itself, so every window that lies
    inside the image has NCC 1 -- cost 0 up to FP32 rounding -- for any
    hypothesis.  A gather addressed wrongly anywhere in the plane (an offset
    that wraps, a row pitch off by one) breaks that."""
    W, H, R = 16384, 16383, 5
    img = np.random.default_rng(7).integers(0, 256, (H, W), dtype=np.uint8)
    ref = ctx.upload(img, pkg.STORE_F32)
    f = 16000.0
    Kinv = np.linalg.inv(np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1.0]]))
    d = pkg.make_desc(ref, [ref], Kinv, [np.eye(3)], [np.zeros(3)], 2.0, 12.0, radius=R,
                      stride=2, iterations=1, top_k=1, seed=11)
    print("solving 16384 x 16383, 1 source, 1 iteration", flush=True)
    out = ctx.run_view(d, (H, W))
    ctx.release(ref)
    inner = out["cost"][R + 1:H - R - 1, R + 1:W - R - 1]
    print(f"max cost inside {inner.max():.2e}", flush=True)
    assert inner.max() < 1e-5
    assert np.all((out["depth"] >= 2.0) & (out["depth"] <= 12.0))


def test_widest_row_pitch_accepted_and_solved(pkg, ctx):
    """Source taps form row offsets with a 24-bit multiply, so a plane's row
    pitch (the width rounded up to 64 elements) must stay below 2^24: width
    2^24 - 64 is accepted, one pixel more is refused before anything is
    allocated.  On the widest image, with the source equal to the reference
    and the identity homography, every window inside the image must cost 0
    (up to FP32 rounding): a row offset that wrapped would break that."""
    W, H, R = (1 << 24) - 64, 12, 3
    wide = np.zeros((H, W + 1), np.uint8)
    with pytest.raises(pkg.HP3DError) as e:
        ctx.upload(wide, pkg.STORE_F16)
    assert e.value.code == pkg.ERR_UNSUPPORTED
    del wide
    img = np.random.default_rng(3).integers(0, 256, (H, W), dtype=np.uint8)
    ref = ctx.upload(img, pkg.STORE_F16)
    f = float(W)
    Kinv = np.linalg.inv(np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1.0]]))
    d = pkg.make_desc(ref, [ref], Kinv, [np.eye(3)], [np.zeros(3)], 2.0, 12.0, radius=R,
                      stride=2, iterations=1, top_k=1, seed=5)
    print(f"solving {W} x {H}, 1 source, 1 iteration", flush=True)
    out = ctx.run_view(d, (H, W), with_normal=False)
    ctx.release(ref)
    inner = out["cost"][R:H - R - 1, R:W - R - 1]
    print(f"max cost inside {inner.max():.2e}", flush=True)
    assert inner.max() < 1e-5


def test_non_finite_homography_terms_refused(pkg, ctx, scenes):
    """A source whose A or b holds a NaN or an infinity is refused before
    any launch (by the per-view, multi-view and batched entry points)."""
    job = scenes.sphere_job(64, 48, 3, scene_seed=3, radius=2, stride=1, iterations=1)
    ref = ctx.upload(job.ref)
    srcs = [ctx.upload(s) for s in job.srcs]
    for which, val in (("A", float("nan")), ("b", float("inf"))):
        A = [np.array(a, dtype=np.float64) for a in job.A]
        b = [np.array(x, dtype=np.float64) for x in job.b]
        (A if which == "A" else b)[1].flat[0] = val
        d = pkg.make_desc(ref, srcs, job.Kinv, A, b, job.dmin, job.dmax, radius=2, stride=1,
                          iterations=1)
        with pytest.raises(pkg.HP3DError) as e:
            ctx.run_view(d, job.shape)
        assert e.value.code == pkg.ERR_INVALID_ARG and "not finite" in str(e.value)
        with pytest.raises(pkg.HP3DError) as e:
            ctx.batch_run([d], [job.shape])
        assert e.value.code == pkg.ERR_INVALID_ARG
