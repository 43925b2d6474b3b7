"""GPU tests of the multi-view call hp3d_pm_run_views: views solved back to
back with each view's download overlapping the next view's solve.

Every view's depth, normal and cost must equal hp3d_pm_run_view's on the
same view bit for bit (run_view is checked against the oracle in
test_gpu_parity.py), whatever the mix of sizes, output kinds (page-locked,
pageable, device) and the reuse of the two staging slots.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def views(pkg, scenes, ctx):
    specs = [(200, 150, 3, 3), (96, 64, 5, 2), (256, 192, 4, 2), (120, 90, 8, 1), (64, 96, 2, 3)]
    out = []
    for k, (W, H, S, it) in enumerate(specs):
        job = scenes.sphere_job(W, H, S, scene_seed=70 + k, radius=5, stride=2, iterations=it,
                                seed=11 + 1000000 * k, name=f"pipe{k}")
        ref = ctx.upload(job.ref)
        srcs = [ctx.upload(s) for s in job.srcs]
        d = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=5,
                          stride=2, iterations=it, seed=job.seed)
        out.append((job, d, ctx.run_view(d, job.shape)))
    return out


def _same(res, exp):
    for key in ("depth", "normal", "cost"):
        assert np.array_equal(res[key], exp[key]), key


@pytest.mark.parametrize("pinned", [True, False], ids=["pinned", "pageable"])
def test_run_views_equals_run_view(pkg, ctx, views, pinned):
    order = [0, 1, 2, 3, 4, 2, 0]      # each staging slot reused, sizes up and down
    outs = [ctx.host_outputs(views[i][0].shape, pinned=pinned) for i in order]
    ctx.run_views([views[i][1] for i in order], outs)
    for i, o in zip(order, outs):
        _same(o, views[i][2])
    if pinned:
        for o in outs:
            for a in o.values():
                ctx.host_free(a)


def test_run_views_device_and_host_outputs_mixed(pkg, ctx, views):
    """A view with device outputs between host ones takes no staging slot."""
    job, d, exp = views[2]
    H, W = job.shape
    dd, dn, dc = (ctx.device_alloc(W * H * 4 * m) for m in (1, 3, 1))
    o0 = ctx.host_outputs(views[0][0].shape, pinned=True)
    o2 = ctx.host_outputs(views[1][0].shape)
    descs = (pkg.ViewDesc * 3)(views[0][1], d, views[1][1])
    outs = (pkg.ViewOut * 3)()
    outs[0] = pkg.ViewOut(pkg._p(o0["depth"], pkg._f32p), pkg._p(o0["normal"], pkg._f32p),
                          pkg._p(o0["cost"], pkg._f32p), None, 0)
    outs[1] = pkg.ViewOut(pkg.ctypes.cast(dd, pkg._f32p), pkg.ctypes.cast(dn, pkg._f32p),
                          pkg.ctypes.cast(dc, pkg._f32p), None, 1)
    outs[2] = pkg.ViewOut(pkg._p(o2["depth"], pkg._f32p), pkg._p(o2["normal"], pkg._f32p),
                          pkg._p(o2["cost"], pkg._f32p), None, 0)
    pkg._check(pkg.lib().hp3d_pm_run_views(ctx.h, descs, 3, outs), ctx.h)
    _same(o0, views[0][2])
    _same(o2, views[1][2])
    _same({"depth": ctx.download(dd, (H, W)), "normal": ctx.download(dn, (H, W, 3)),
           "cost": ctx.download(dc, (H, W))}, exp)
    for p in (dd, dn, dc):
        ctx.device_free(p)


def test_run_views_rejects_bad_views(pkg, ctx, views):
    job, d, _ = views[0]
    bad = pkg.make_desc(d.ref, [d.src[0]], job.Kinv, job.A[:1], job.b[:1], 5.0, 5.0, radius=5,
                        stride=2, iterations=1)
    o = ctx.host_outputs(job.shape)
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_views([d, bad], [o, ctx.host_outputs(job.shape)])
    assert e.value.code == pkg.ERR_INVALID_ARG
    assert "view 1" in ctx.last_error()


def _seeds(pkg, pts):
    arr = (pkg.SeedPoint * len(pts))()
    for i, (x, y, dep, n) in enumerate(pts):
        arr[i].x, arr[i].y, arr[i].depth = x, y, dep
        for k in range(3):
            arr[i].normal[k] = n[k]
    return arr, len(pts)


def test_run_views_with_caller_hypotheses_and_seeds(pkg, ctx, views):
    """Views that bring their own starting hypotheses or SfM seeds, in
    growing sizes, so the buffers those pass through are the largest view's
    from the start; each equals run_view on the same descriptor."""
    descs, exps, shapes = [], [], []
    for k in (1, 0, 2, 4, 2):                       # sizes down, up, down, up
        job, d, _ = views[k]
        H, W = job.shape
        kw = dict(radius=5, stride=2, iterations=2, seed=job.seed + 7)
        if k % 2 == 0:                               # start from another seed's hypotheses
            d0, n0 = ctx.init_hypotheses(d, job.shape)
            kw["init"] = (d0[::-1].copy(), n0[::-1].copy())
        else:                                        # seeds painted over the random start
            kw["seeds"] = _seeds(pkg, [(W // 3, H // 2, float(job.dmin + 1.0), (0.0, 0.0, -1.0)),
                                       (2 * W // 3, H // 3, float(job.dmax - 1.0), (0.0, 0.6, -0.8))])
            kw["seed_radius"] = 6
        dd = pkg.make_desc(d.ref, [d.src[i] for i in range(d.num_src)], job.Kinv, job.A, job.b,
                           job.dmin, job.dmax, **kw)
        descs.append(dd)
        shapes.append(job.shape)
        exps.append(ctx.run_view(dd, job.shape))
    outs = [ctx.host_outputs(s, pinned=True) for s in shapes]
    ctx.run_views(descs, outs)
    for o, e in zip(outs, exps):
        _same(o, e)
    for o in outs:
        for a in o.values():
            ctx.host_free(a)
