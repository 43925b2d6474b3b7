"""The batched and serving workloads at their real sizes, checked against the
oracle (MI355X).

* A batch shaped like the batched workload: views of 640x480, 1280x960,
  1600x1200, 2400x1800 and 4000x3000 with 3-8 sources, 11x11 windows at
  stride 1, 8 iterations, FP16 storage, through hp3d_pm_batch_run (graph of
  steps and persistent kernel).  Every view equals its per-view solve bit for
  bit; the 640x480 view puts >= 99 % of pixels within 1 % of the FP64-cost
  oracle's depth; on the 4000x3000 view, too large for the oracle's full run,
  size-independent properties hold: the reported cost is exactly the cost the
  cost pass gives the final hypotheses, that cost is within 1e-4 of the FP64
  oracle on 20000 sampled pixels, depths stay in range, and the share of
  pixels within 1 % of the ground truth is no lower than on the 1600x1200 view.
* A serving window shaped like the serving workload: Poisson arrivals at
  20 jobs/s, each job 1 reference + 3-8 sources at a resolution of the same
  mix and 4-12 iterations, FP16, stride 2.  Every job equals its per-view
  solve bit for bit, and the smallest job's depth agrees with the oracle.
"""
import time

import numpy as np
import pytest

from conftest import oracle_view, rel_within

pytestmark = pytest.mark.gpu

MB = 1 << 20
MIX = [(640, 480, 3), (1280, 960, 8), (1600, 1200, 5), (2400, 1800, 4), (4000, 3000, 6)]


def _descs(pkg, ctx, jobs, store, **kw):
    out, handles = [], []
    for j in jobs:
        ref = ctx.upload(j.ref, store)
        srcs = [ctx.upload(s, store) for s in j.srcs]
        handles += [ref] + srcs
        out.append(pkg.make_desc(ref, srcs, j.Kinv, j.A, j.b, j.dmin, j.dmax, radius=j.radius,
                                 stride=j.stride, iterations=kw.get("iterations", j.iterations),
                                 seed=j.seed))
    return out, handles


@pytest.mark.timeout(600)
def test_batched_workload_shape_against_oracle(pkg, orc, ctx, scenes):
    jobs = [scenes.sphere_job(W, H, S, scene_seed=300 + k, radius=5, stride=1, iterations=8,
                              seed=42 + 1000000 * k, name=f"c4_{k}")
            for k, (W, H, S) in enumerate(MIX)]
    print("rendered", flush=True)
    descs, handles = _descs(pkg, ctx, jobs, pkg.STORE_F16)
    shapes = [j.shape for j in jobs]
    singles = [ctx.run_view(d, s) for d, s in zip(descs, shapes)]
    for persistent in (False, True):
        print(f"batch persistent={persistent}", flush=True)
        outs = ctx.batch_run(descs, shapes, persistent=persistent)
        for k, (o, r) in enumerate(zip(outs, singles)):
            for key in ("depth", "normal", "cost"):
                assert np.array_equal(o[key], r[key]), (persistent, k, key)
    # the 640x480 view against the oracle's FP64-cost run
    small = jobs[0]
    o64, _, _ = oracle_view(orc, small).patchmatch_fp32(small.iterations, small.seed,
                                                         cost_mode=orc.COST_FP64)
    agree = rel_within(singles[0]["depth"], o64).mean()
    print(f"640x480: depth within 1 % of the FP64 oracle on {agree:.4f}", flush=True)
    assert agree >= 0.99
    # the 4000x3000 view: size-independent properties
    big, out = jobs[4], singles[4]
    c_again = ctx.cost_eval(descs[4], out["depth"], out["normal"])
    assert np.array_equal(c_again, out["cost"])
    H, W = big.shape
    rng = np.random.default_rng(1)
    ys, xs = rng.integers(0, H, 20000), rng.integers(0, W, 20000)
    v = oracle_view(orc, big)
    c64, flags = v.cost_list_fp64(xs, ys, out["depth"][ys, xs], out["normal"][ys, xs])
    ok = flags == 0
    err = np.abs(out["cost"][ys, xs].astype(np.float64) - c64)[ok]
    print(f"4000x3000: final cost vs FP64 oracle max {err.max():.2e} over {ok.sum()} pixels",
          flush=True)
    assert ok.mean() > 0.9 and err.max() <= 1e-4
    assert np.all((out["depth"] >= big.dmin) & (out["depth"] <= big.dmax))
    gt = [rel_within(singles[k]["depth"], jobs[k].gt_depth).mean() for k in (2, 4)]
    print(f"within 1 % of ground truth: 1600x1200 {gt[0]:.4f}, 4000x3000 {gt[1]:.4f}")
    assert gt[1] >= gt[0]
    for h in handles:
        ctx.release(h)


@pytest.mark.timeout(600)
def test_serving_window_against_per_view_and_oracle(pkg, orc, ctx, scenes):
    rng = np.random.default_rng(2026)
    pool = []
    for k in range(8):
        W, H = MIX[k % 5][:2]
        S = int(rng.integers(3, 9))
        pool.append(scenes.sphere_job(W, H, S, scene_seed=400 + k, radius=5, stride=2,
                                      iterations=8, seed=7 + 1000000 * k, name=f"c5_{k}"))
    print("rendered", flush=True)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / 20.0)
        if t > 1.5:
            break
        arrivals.append((t, int(rng.integers(len(pool))), int(rng.integers(4, 13))))
    with pkg.Server(24 * 1024 * MB, max_jobs=16, stride=2) as srv:
        t0 = time.perf_counter()
        ids = []
        for at, k, it in arrivals:
            while time.perf_counter() - t0 < at:
                time.sleep(0.0005)
            ids.append(srv.submit_job(pool[k], iterations=it))
        lat = []
        for jid in ids:
            st = srv.wait(jid, 300)
            assert st["state"] == pkg.JOB_DONE, st
            lat.append(st["t_done"] - st["t_submit"])
        s = srv.stats()
    print(f"{len(ids)} jobs, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, p99 "
          f"{np.percentile(lat, 99) * 1e3:.1f} ms, max active {s['max_active']}", flush=True)
    assert s["done"] == len(ids) and s["max_active"] >= 2
    # every job against its per-view solve
    cache = {}
    for (at, k, it), jid in zip(arrivals, ids):
        if (k, it) not in cache:
            d, hs = _descs(pkg, ctx, [pool[k]], pkg.STORE_F16, iterations=it)
            cache[(k, it)] = ctx.run_view(d[0], pool[k].shape)
            for h in hs:
                ctx.release(h)
        exp, got = cache[(k, it)], srv.result(jid)
        assert np.array_equal(got["depth"], exp["depth"]) and np.array_equal(got["cost"], exp["cost"])
    # the job of least work against the oracle's FP64-cost run
    at, k, it = min(arrivals, key=lambda a: pool[a[1]].shape[0] * pool[a[1]].shape[1] *
                    pool[a[1]].S * a[2])
    job = pool[k]
    o64, _, _ = oracle_view(orc, job).patchmatch_fp32(it, job.seed, cost_mode=orc.COST_FP64)
    agree = rel_within(cache[(k, it)]["depth"], o64).mean()
    print(f"job {job.shape} S={job.S} {it} iterations: within 1 % of the oracle on {agree:.4f}")
    assert agree >= 0.99
