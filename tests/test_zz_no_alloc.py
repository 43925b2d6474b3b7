"""GPU test of hp3d_pm_run_view_noalloc, the solve that allocates nothing:
its view state lives in the library's fixed device scratch array.  Called
through the C ABI twice in a row -- the second time on a smaller view, so
the array holds the first call's larger state -- each output must equal
hp3d_pm_run_view's on the same view (device outputs) bit for bit; a view too
large for the array and host outputs are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_both(pkg, ctx, job):
    ref = ctx.upload(job.ref)
    srcs = [ctx.upload(s) for s in job.srcs]
    d = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=job.radius,
                      stride=job.stride, iterations=job.iterations, seed=job.seed)
    H, W = job.shape
    res = {}
    for name in ("alloc", "noalloc"):
        bufs = [ctx.device_alloc(W * H * 4 * m) for m in (1, 3, 1)]
        out = pkg.ViewOut(ctypes.cast(bufs[0], pkg._f32p), ctypes.cast(bufs[1], pkg._f32p),
                          ctypes.cast(bufs[2], pkg._f32p), None, 1)
        fn = pkg.lib().hp3d_pm_run_view if name == "alloc" else pkg.lib().hp3d_pm_run_view_noalloc
        print(f"{name}: {W}x{H}, {job.S} sources, {job.iterations} iterations", flush=True)
        pkg._check(fn(ctx.h, ctypes.byref(d), ctypes.byref(out)), ctx.h)
        res[name] = {"depth": ctx.download(bufs[0], (H, W)),
                     "normal": ctx.download(bufs[1], (H, W, 3)),
                     "cost": ctx.download(bufs[2], (H, W))}
        for b in bufs:
            ctx.device_free(b)
    for h in [ref] + srcs:
        ctx.release(h)
    return res


def test_noalloc_equals_run_view_twice_second_smaller(pkg, ctx, scenes):
    big = scenes.sphere_job(640, 480, 4, scene_seed=61, radius=5, stride=2, iterations=3,
                            seed=101, name="noalloc_big")
    small = scenes.sphere_job(200, 150, 3, scene_seed=62, radius=5, stride=2, iterations=2,
                              seed=202, name="noalloc_small")
    for job in (big, small):
        r = _solve_both(pkg, ctx, job)
        for key in ("depth", "normal", "cost"):
            assert np.array_equal(r["noalloc"][key], r["alloc"][key]), (job.name, key)
        assert np.all((r["noalloc"]["depth"] >= job.dmin) & (r["noalloc"]["depth"] <= job.dmax))


def test_noalloc_refuses_oversized_view_and_host_outputs(pkg, ctx):
    W, H = 2400, 1800                       # 4.3 Mpixel: above the scratch array
    img = np.zeros((H, W), np.uint8)
    ref = ctx.upload(img)
    d = pkg.make_desc(ref, [ref], np.eye(3), [np.eye(3)], [np.zeros(3)], 2.0, 12.0, radius=5,
                      stride=2, iterations=1)
    dd = ctx.device_alloc(W * H * 4)
    out = pkg.ViewOut(ctypes.cast(dd, pkg._f32p), None, None, None, 1)
    code = pkg.lib().hp3d_pm_run_view_noalloc(ctx.h, ctypes.byref(d), ctypes.byref(out))
    assert code == pkg.ERR_UNSUPPORTED, ctx.last_error()
    host = np.zeros((H, W), np.float32)
    out_h = pkg.ViewOut(pkg._p(host, pkg._f32p), None, None, None, 0)
    code = pkg.lib().hp3d_pm_run_view_noalloc(ctx.h, ctypes.byref(d), ctypes.byref(out_h))
    assert code == pkg.ERR_INVALID_ARG
    ctx.device_free(dd)
    ctx.release(ref)
