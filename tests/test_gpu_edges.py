"""Edge cases of the plain path against the oracle (MI355X).

Ragged and tiny images (odd widths, a width that is not a multiple of the
64-pixel tile, an image exactly the size of the window), every window the
kernels are built for, 1 and 8 sources (top-3 over fewer than three
sources), flat images (every window hits the variance guard) and a source
that sees nothing of the reference (every tap out of bounds).  Gates: random
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
