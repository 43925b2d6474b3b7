"""GPU tests of the batched form (hp3d_pm_batch_run) and its autotuner.

The step kernel computes every pixel with the same device functions and the
same random counters as the per-view kernels, so a batched view must match
its per-view solve bit for bit -- whatever the batch mix, the tile shape
(kernel variant) or whether the step sequence runs from a hipGraph.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed_jobs(scenes):
    """Uneven sizes, source counts and iteration counts."""
    specs = [(160, 120, 3, 2), (96, 200, 5, 3), (240, 180, 8, 1), (64, 48, 4, 0), (200, 100, 6, 4)]
    jobs = []
    for k, (W, H, S, it) in enumerate(specs):
        jobs.append(scenes.sphere_job(W, H, S, scene_seed=20 + k, radius=5, stride=1,
                                      iterations=it, seed=42 + 1000000 * k, name=f"mix{k}"))
    return jobs


def descs_for(pkg, ctx, jobs, store=0):
    out = []
    for job in jobs:
        ref = ctx.upload(job.ref, store)
        srcs = [ctx.upload(s, store) for s in job.srcs]
        out.append(pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                                 radius=job.radius, stride=job.stride, iterations=job.iterations,
                                 seed=job.seed))
    return out


@pytest.fixture(scope="module")
def singles(pkg, ctx, mixed_jobs):
    descs = descs_for(pkg, ctx, mixed_jobs)
    return descs, [ctx.run_view(d, j.shape) for d, j in zip(descs, mixed_jobs)]


@pytest.mark.parametrize("variant", list(range(6)))
def test_batch_matches_per_view_bit_for_bit(pkg, ctx, mixed_jobs, singles, variant):
    descs, ref_out = singles
    outs = ctx.batch_run(descs, [j.shape for j in mixed_jobs], variant=variant)
    for k, (o, r) in enumerate(zip(outs, ref_out)):
        for key in ("depth", "normal", "cost"):
            assert np.array_equal(o[key], r[key]), (variant, k, key)


def test_batch_without_graph_and_graph_replay(pkg, ctx, mixed_jobs, singles):
    descs, ref_out = singles
    shapes = [j.shape for j in mixed_jobs]
    for use_graph in (False, True, True):          # the second graph run replays
        outs = ctx.batch_run(descs, shapes, variant=0, use_graph=use_graph)
        for o, r in zip(outs, ref_out):
            assert np.array_equal(o["depth"], r["depth"])


def test_batch_order_does_not_matter(pkg, ctx, mixed_jobs, singles):
    descs, ref_out = singles
    order = [3, 1, 4, 0, 2]
    outs = ctx.batch_run([descs[i] for i in order], [mixed_jobs[i].shape for i in order])
    for pos, i in enumerate(order):
        assert np.array_equal(outs[pos]["depth"], ref_out[i]["depth"])


def test_batch_f16_storage_matches_per_view_f16(pkg, ctx, mixed_jobs):
    descs16 = descs_for(pkg, ctx, mixed_jobs, store=1)
    outs = ctx.batch_run(descs16, [j.shape for j in mixed_jobs])
    for d, o, j in zip(descs16, outs, mixed_jobs):
        r = ctx.run_view(d, j.shape)
        assert np.array_equal(o["depth"], r["depth"]) and np.array_equal(o["cost"], r["cost"])


def test_autotune_picks_a_variant_and_is_persisted(pkg, ctx, mixed_jobs, singles, tmp_path):
    descs, ref_out = singles
    best, times = ctx.autotune(descs)
    print("autotune times (ms):", times, "best", best)
    assert 0 <= best < pkg.NUM_VARIANTS and times[best] > 0
    assert times[best] == times[times > 0].min()
    p = str(tmp_path / "tune.txt")
    ctx.tune_save(p)
    other = pkg.Context()
    other.tune_load(p)
    other.close()
    outs = ctx.batch_run(descs, [j.shape for j in mixed_jobs])   # uses the tuned variant
    for o, r in zip(outs, ref_out):
        assert np.array_equal(o["depth"], r["depth"])


def test_batch_rejects_mixed_windows(pkg, ctx, mixed_jobs, singles):
    descs = list(singles[0][:2])
    bad = pkg.make_desc(descs[1].ref, [descs[1].src[i] for i in range(descs[1].num_src)],
                        mixed_jobs[1].Kinv, mixed_jobs[1].A, mixed_jobs[1].b, 2.0, 12.0,
                        radius=3, stride=1, iterations=1)
    with pytest.raises(pkg.HP3DError) as e:
        ctx.batch_run([descs[0], bad], [mixed_jobs[0].shape, mixed_jobs[1].shape])
    assert e.value.code == pkg.ERR_INVALID_ARG


# ---------------------------------------------------------------------------
# Persistent form: one kernel for the whole batch, tiles released by per-tile
# counters (a half-pass of a tile waits for that tile and its 4 neighbours
# to finish the previous pass).  Same pixels, same arithmetic: every view
# must equal its per-view solve bit for bit, for every tile shape and chunk
# size, from a graph replay too (the counters are reset on every replay).
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("variant,chunk", [(0, 1), (1, 8), (2, 3), (3, 16), (4, 8), (5, 64)])
def test_persistent_batch_matches_per_view(pkg, ctx, mixed_jobs, singles, variant, chunk):
    descs, ref_out = singles
    shapes = [j.shape for j in mixed_jobs]
    print(f"persistent variant {variant} chunk {chunk}", flush=True)
    outs = ctx.batch_run(descs, shapes, variant=variant, persistent=True, chunk=chunk)
    for k, (o, r) in enumerate(zip(outs, ref_out)):
        for key in ("depth", "normal", "cost"):
            assert np.array_equal(o[key], r[key]), (variant, chunk, k, key)


def test_persistent_batch_graph_replays_and_direct(pkg, ctx, mixed_jobs, singles):
    descs, ref_out = singles
    shapes = [j.shape for j in mixed_jobs]
    for use_graph in (True, True, True, False):
        outs = ctx.batch_run(descs, shapes, variant=1, use_graph=use_graph, persistent=True)
        for o, r in zip(outs, ref_out):
            assert np.array_equal(o["depth"], r["depth"]) and np.array_equal(o["cost"], r["cost"])


def test_persistent_batch_f16_and_other_window(pkg, ctx, scenes, mixed_jobs):
    """FP16 storage, and a 5x5 window at stride 2 (the 16x16-tile kernel)."""
    descs16 = descs_for(pkg, ctx, mixed_jobs, store=1)
    outs = ctx.batch_run(descs16, [j.shape for j in mixed_jobs], persistent=True)
    for d, o, j in zip(descs16, outs, mixed_jobs):
        r = ctx.run_view(d, j.shape)
        assert np.array_equal(o["depth"], r["depth"]) and np.array_equal(o["cost"], r["cost"])
    jobs = [scenes.sphere_job(W, H, S, scene_seed=90 + k, radius=2, stride=2, iterations=it,
                              seed=5 + k, name=f"r2s2_{k}")
            for k, (W, H, S, it) in enumerate([(130, 90, 3, 2), (77, 140, 6, 3), (200, 64, 2, 1)])]
    descs = descs_for(pkg, ctx, jobs)
    outs = ctx.batch_run(descs, [j.shape for j in jobs], persistent=True, chunk=2)
    for d, o, j in zip(descs, outs, jobs):
        r = ctx.run_view(d, j.shape)
        assert np.array_equal(o["depth"], r["depth"]) and np.array_equal(o["normal"], r["normal"])
