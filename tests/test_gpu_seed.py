"""SfM seeding on the device against the oracle's sequential painting (MI355X).

The reference paints a (2r+1)^2 square of (z, n) around every seed in point
order over the random init (src/dense_recon/patchmatch_mvs.cu:1111-1177);
the device resolves each pixel to the last covering seed with an atomic max.
The hypotheses after init + seeding must be bit-identical to the oracle's
init followed by its in-order painting, and a seeded run must then follow
the same schedule as a run given those hypotheses explicitly.
"""
import numpy as np
import pytest

from conftest import oracle_view

pytestmark = pytest.mark.gpu


def _desc(pkg, ctx, job, seeds=None, radius=50, iterations=0, init=None):
    ref = ctx.upload(job.ref)
    srcs = [ctx.upload(s) for s in job.srcs]
    return pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                         radius=job.radius, stride=job.stride, iterations=iterations,
                         top_k=job.top_k, seed=job.seed, seeds=seeds, seed_radius=radius, init=init)


def _seed_array(pkg, seeds):
    arr = (pkg.SeedPoint * len(seeds))()
    for i, (x, y, d, n) in enumerate(seeds):
        arr[i].x, arr[i].y, arr[i].depth = x, y, d
        for j in range(3):
            arr[i].normal[j] = n[j]
    return arr, len(seeds)


@pytest.mark.parametrize("radius", [0, 3, 50])
def test_seeded_init_bit_exact(pkg, orc, ctx, c1, radius):
    job = c1
    H, W = job.shape
    rng = np.random.default_rng(radius)
    seeds = []
    for k in range(300):
        x = int(rng.integers(-60, W + 60))
        y = int(rng.integers(-60, H + 60))
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        seeds.append((x, y, float(np.float32(rng.uniform(job.dmin, job.dmax))),
                      tuple(np.float32(n))))
    seeds += [seeds[0][:2] + (float(job.dmin), (0.0, 0.0, -1.0))]   # overwrites seed 0
    d_gpu, n_gpu = ctx.init_hypotheses(_desc(pkg, ctx, job, _seed_array(pkg, seeds), radius),
                                       job.shape)
    d_orc, n_orc = oracle_view(orc, job).init_fp32(job.seed)
    orc.apply_seeds(d_orc, n_orc, seeds, radius)
    assert np.array_equal(d_gpu, d_orc)
    assert np.array_equal(n_gpu, n_orc)


def test_seeded_run_equals_given_hypotheses(pkg, ctx, scenes):
    tt = scenes.turntable(n_views=12, W=320, H=240, f=500.0, threads=8)
    sfm = pkg.SfM(tt["cameras"], tt["points"])
    vp = sfm.select_source_views(4)[0]
    seeds = sfm.seed_points(0, vp["depth_min"], vp["depth_max"], 320, 240, seed_radius=5)
    assert seeds[1] > 10
    K = tt["cameras"][0]["K"]
    Kinv = pkg.invert3(K)
    A, b = [], []
    for sid, _, _ in vp["sources"]:
        c0, cs = tt["cameras"][0], tt["cameras"][sid]
        a_, b_ = pkg.compose(K, c0["R"], c0["t"], cs["K"], cs["R"], cs["t"])
        A.append(a_)
        b.append(b_)
    ref = ctx.upload(tt["images"][0])
    srcs = [ctx.upload(tt["images"][s]) for s, _, _ in vp["sources"]]

    def desc(**kw):
        return pkg.make_desc(ref, srcs, Kinv, A, b, vp["depth_min"], vp["depth_max"], radius=5,
                             stride=1, seed=42, **kw)

    d0, n0 = ctx.init_hypotheses(desc(seeds=seeds, seed_radius=5), (240, 320))
    a = ctx.run_view(desc(iterations=3, seeds=seeds, seed_radius=5), (240, 320))
    b_ = ctx.run_view(desc(iterations=3, init=(d0, n0)), (240, 320))
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(a[k], b_[k])


def test_seeds_rejected_by_batch_and_checked(pkg, ctx, c1):
    job = c1
    bad = _seed_array(pkg, [(1 << 25, 0, 3.0, (0.0, 0.0, -1.0))])
    with pytest.raises(pkg.HP3DError):
        ctx.init_hypotheses(_desc(pkg, ctx, job, bad), job.shape)
    ok = _seed_array(pkg, [(10, 10, 3.0, (0.0, 0.0, -1.0))])
    with pytest.raises(pkg.HP3DError):
        ctx.init_hypotheses(_desc(pkg, ctx, job, ok, radius=5000), job.shape)
    with pytest.raises(pkg.HP3DError):
        ctx.batch_run([_desc(pkg, ctx, job, ok, iterations=1)], [job.shape])
