"""GPU tests of the serving form (hp3d_server_*): continuous batching.

Each job's result must equal its per-view solve bit for bit, alone and in a
mix of jobs that arrive while others run; admission respects the HBM budget;
a cancelled or failed job leaves the others untouched.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.fixture(scope="module")
def serve_jobs(scenes):
    specs = [(200, 150, 3, 4), (128, 96, 5, 2), (256, 192, 4, 6), (96, 72, 8, 3), (160, 240, 6, 5),
             (180, 120, 3, 1)]
    return [scenes.sphere_job(W, H, S, scene_seed=40 + k, radius=5, stride=2, iterations=it,
                              seed=7 + 1000000 * k, name=f"srv{k}")
            for k, (W, H, S, it) in enumerate(specs)]


@pytest.fixture(scope="module")
def expected(pkg, ctx, serve_jobs):
    """Per-view solves with the server's storage (FP16)."""
    out = []
    for job in serve_jobs:
        ref = ctx.upload(job.ref, pkg.STORE_F16)
        srcs = [ctx.upload(s, pkg.STORE_F16) for s in job.srcs]
        d = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=5,
                          stride=2, iterations=job.iterations, seed=job.seed)
        out.append(ctx.run_view(d, job.shape))
    return out


def _check(res, exp):
    assert np.array_equal(res["depth"], exp["depth"])
    assert np.array_equal(res["cost"], exp["cost"])


def test_single_job_matches_run_view(pkg, serve_jobs, expected):
    with pkg.Server(256 * MB, max_jobs=4, stride=2) as srv:
        jid = srv.submit_job(serve_jobs[0])
        st = srv.wait(jid, timeout=60)
        assert st["state"] == pkg.JOB_DONE
        assert st["steps"] == 1 + 2 * serve_jobs[0].iterations
        _check(srv.result(jid), expected[0])


def test_mix_matches_alone(pkg, serve_jobs, expected):
    with pkg.Server(512 * MB, max_jobs=8, stride=2) as srv:
        ids = [srv.submit_job(j) for j in serve_jobs]
        for k, jid in enumerate(ids):
            st = srv.wait(jid, timeout=120)
            assert st["state"] == pkg.JOB_DONE, st
            _check(srv.result(jid), expected[k])
        s = srv.stats()
        assert s["done"] == len(serve_jobs) and s["max_active"] >= 2
        print("serve stats:", s)


def test_staggered_arrivals_and_few_slots(pkg, serve_jobs, expected):
    with pkg.Server(512 * MB, max_jobs=2, stride=2) as srv:
        ids = []
        for j in serve_jobs:
            ids.append(srv.submit_job(j))
            time.sleep(0.003)
        for k, jid in enumerate(ids):
            assert srv.wait(jid, timeout=120)["state"] == pkg.JOB_DONE
            _check(srv.result(jid), expected[k])
        assert srv.stats()["max_active"] <= 2


def test_budget_admission(pkg, serve_jobs, expected):
    small = 40 * MB
    with pkg.Server(small, max_jobs=8, stride=2) as srv:
        big = serve_jobs[2]
        ids = [srv.submit_job(j) for j in serve_jobs]
        for k, jid in enumerate(ids):
            assert srv.wait(jid, timeout=120)["state"] == pkg.JOB_DONE
            _check(srv.result(jid), expected[k])
        s = srv.stats()
        assert s["arena_peak"] <= small
        # a job that can never fit is refused at submit
        huge = np.zeros((3000, 4000), np.uint8)
        with pytest.raises(pkg.HP3DError) as e:
            srv.submit(huge, [huge, huge], big.Kinv, big.A[:2], big.b[:2], 2.0, 12.0, 1)
        assert e.value.code == pkg.ERR_OOM_BUDGET
        assert srv.stats()["rejected"] == 1


def test_cancel_and_injected_failure_leave_others_intact(pkg, serve_jobs, expected):
    with pkg.Server(512 * MB, max_jobs=1, stride=2) as srv:
        first = srv.submit_job(serve_jobs[2])                    # occupies the only slot
        failing = srv.submit_job(serve_jobs[1], fail_at_step=2)
        cancelled = srv.submit_job(serve_jobs[3])
        srv.cancel(cancelled)
        last = srv.submit_job(serve_jobs[0])
        assert srv.wait(first, 120)["state"] == pkg.JOB_DONE
        st_f = srv.wait(failing, 120)
        assert st_f["state"] == pkg.JOB_FAILED and st_f["steps"] == 2
        assert srv.wait(cancelled, 120)["state"] == pkg.JOB_CANCELLED
        assert srv.wait(last, 120)["state"] == pkg.JOB_DONE
        _check(srv.result(first), expected[2])
        _check(srv.result(last), expected[0])
        s = srv.stats()
        assert (s["done"], s["failed"], s["cancelled"]) == (2, 1, 1)


def test_forget_only_finished_jobs(pkg, serve_jobs, expected):
    """A queued or running job cannot be forgotten (its output buffers are
    still to be written); a finished one is dropped and no longer found."""
    with pkg.Server(512 * MB, max_jobs=1, stride=2) as srv:
        first = srv.submit_job(serve_jobs[2])
        queued = srv.submit_job(serve_jobs[0])      # waits for the only slot
        with pytest.raises(pkg.HP3DError) as e:
            srv.forget(queued)
        assert e.value.code == pkg.ERR_AGAIN
        assert srv.wait(first, 120)["state"] == pkg.JOB_DONE
        assert srv.wait(queued, 120)["state"] == pkg.JOB_DONE
        _check(srv.result(queued), expected[0])
        srv.forget(first)
        srv.forget(queued)
        with pytest.raises(pkg.HP3DError) as e:
            srv.poll(first)
        assert e.value.code == pkg.ERR_NOT_FOUND
        with pytest.raises(pkg.HP3DError) as e:
            srv.forget(first)
        assert e.value.code == pkg.ERR_NOT_FOUND


def test_invalid_submissions(pkg, serve_jobs):
    with pkg.Server(64 * MB, max_jobs=2, stride=2) as srv:
        j = serve_jobs[0]
        with pytest.raises(pkg.HP3DError) as e:
            srv.submit(j.ref, [], j.Kinv, [], [], 2.0, 12.0, 1)
        assert e.value.code == pkg.ERR_INVALID_ARG
        with pytest.raises(pkg.HP3DError) as e:
            srv.submit(j.ref, j.srcs, j.Kinv, j.A, j.b, 5.0, 5.0, 1)
        assert e.value.code == pkg.ERR_INVALID_ARG
        for dmax, iters in ((float("inf"), 1), (12.0, (1 << 20) + 1)):
            with pytest.raises(pkg.HP3DError) as e:
                srv.submit(j.ref, j.srcs, j.Kinv, j.A, j.b, 2.0, dmax, iters)
            assert e.value.code == pkg.ERR_INVALID_ARG, (dmax, iters)
        bad_b = [np.array(b, dtype=np.float64) for b in j.b]
        bad_b[-1][2] = float("nan")
        with pytest.raises(pkg.HP3DError) as e:
            srv.submit(j.ref, j.srcs, j.Kinv, j.A, bad_b, 2.0, 12.0, 1)
        assert e.value.code == pkg.ERR_INVALID_ARG
        with pytest.raises(pkg.HP3DError):
            srv.poll(123456)


def test_cancel_queued_job_releases_nothing(pkg, ctx, serve_jobs, expected):
    """A job cancelled while it waits in the queue owns no arena block, so
    cancelling it must not free one: here job A runs at arena offset 0, job
    B (A's size) waits because the budget leaves room only for a smaller
    job C, B is cancelled and C submitted.  C must be placed beside A, not
    over it: both results equal their per-view solves and the arena's peak
    stays inside the budget."""
    big, small = serve_jobs[2], serve_jobs[0]
    fp = {}
    for name, job in (("big", big), ("small", small)):
        with pkg.Server(256 * MB, max_jobs=1, stride=2) as srv:
            assert srv.wait(srv.submit_job(job, iterations=1), 60)["state"] == pkg.JOB_DONE
            fp[name] = srv.stats()["arena_peak"]
    assert fp["big"] > fp["small"] + 256
    budget = fp["big"] + fp["small"] + 256
    # A runs long enough for B and C to be queued and cancelled meanwhile
    iters_a = 40
    ref = ctx.upload(big.ref, pkg.STORE_F16)
    srcs = [ctx.upload(s, pkg.STORE_F16) for s in big.srcs]
    exp_a = ctx.run_view(pkg.make_desc(ref, srcs, big.Kinv, big.A, big.b, big.dmin, big.dmax,
                                       radius=5, stride=2, iterations=iters_a, seed=big.seed),
                         big.shape)
    for h in [ref] + srcs:
        ctx.release(h)
    with pkg.Server(budget, max_jobs=4, stride=2) as srv:
        a = srv.submit_job(big, iterations=iters_a)
        b = srv.submit_job(big, seed=12345)
        srv.cancel(b)
        c = srv.submit_job(small)
        assert srv.wait(b, 120)["state"] == pkg.JOB_CANCELLED
        assert srv.wait(a, 120)["state"] == pkg.JOB_DONE
        assert srv.wait(c, 120)["state"] == pkg.JOB_DONE
        _check(srv.result(a), exp_a)
        _check(srv.result(c), expected[0])
        s = srv.stats()
        print("arena", fp, budget, s["arena_peak"])
        assert s["arena_peak"] <= budget


# ---------------------------------------------------------------------------
# Resident images and pyramid levels
# ---------------------------------------------------------------------------
def test_resident_images_match_host_images(pkg, serve_jobs, expected):
    """Jobs that name resident images give the same bits as jobs that send
    their images; each resident image is uploaded once however many jobs
    use it, and the stats count the reuses."""
    job = serve_jobs[2]
    with pkg.Server(256 * MB, max_jobs=4, stride=2) as srv:
        ref = srv.image_add(job.ref)
        srcs = [srv.image_add(s) for s in job.srcs]
        ids = [srv.submit(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, job.iterations,
                          seed=job.seed, shape=job.shape) for _ in range(3)]
        mixed = srv.submit(job.ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                           job.iterations, seed=job.seed)      # host reference, resident sources
        for jid in ids + [mixed]:
            assert srv.wait(jid, 60)["state"] == pkg.JOB_DONE
            _check(srv.result(jid), expected[2])
        s = srv.stats()
        print("resident stats", s)
        assert s["image_uploads"] == 1 + len(srcs) + 1      # resident images + one host image
        assert s["image_reuses"] == 3 * (1 + len(srcs)) + len(srcs)
        assert s["resident_images"] == 1 + len(srcs)


def test_remove_resident_image_while_in_use(pkg, serve_jobs, expected):
    job = serve_jobs[0]
    with pkg.Server(256 * MB, max_jobs=2, stride=2) as srv:
        hs = [srv.image_add(im) for im in [job.ref] + job.srcs]
        jid = srv.submit(hs[0], hs[1:], job.Kinv, job.A, job.b, job.dmin, job.dmax,
                         job.iterations, seed=job.seed, shape=job.shape)
        for h in hs:
            srv.image_remove(h)                    # still used by the job: freed afterwards
        assert srv.wait(jid, 60)["state"] == pkg.JOB_DONE
        _check(srv.result(jid), expected[0])
        assert srv.stats()["resident_images"] == 0
        with pytest.raises(pkg.HP3DError) as e:
            srv.submit(hs[0], hs[1:], job.Kinv, job.A, job.b, job.dmin, job.dmax, 1,
                       shape=job.shape)
        assert e.value.code == pkg.ERR_NOT_FOUND


@pytest.mark.parametrize("store", [0, 1], ids=["f32", "f16"])
def test_pyramid_levels_are_box_means(pkg, scenes, store):
    """Level l is the 2x2 mean of level l - 1 (sizes halved, odd edges
    dropped): exact in FP32 at every level, in FP16 for the first two."""
    img = np.random.default_rng(4).integers(0, 256, (203, 310), dtype=np.uint8)
    with pkg.Server(64 * MB, max_jobs=1, stride=2, store=store) as srv:
        h = srv.image_add(img, levels=4)
        prev = img.astype(np.float64)
        assert np.array_equal(srv.image_read(h, 0), prev)
        for lv in range(1, 4):
            Hl, Wl = prev.shape[0] // 2, prev.shape[1] // 2
            want = prev[:2 * Hl, :2 * Wl].reshape(Hl, 2, Wl, 2).mean(axis=(1, 3))
            got = srv.image_read(h, lv).astype(np.float64)
            assert got.shape == want.shape
            if store == 0 or lv <= 2:
                assert np.array_equal(got, want), lv
            else:
                assert np.abs(got - want).max() <= 0.0625, lv
            prev = got


def test_job_on_a_pyramid_level_matches_the_small_image(pkg, orc, ctx, scenes):
    """A 2x pixel-replicated image has the original as its first pyramid
    level, exactly: a job on level 1 of replicated images must give the
    bits of the per-view solve on the originals, and that solve agrees with
    the oracle's run."""
    job = scenes.sphere_job(160, 120, 3, scene_seed=71, radius=5, stride=2, iterations=2,
                            seed=9, name="level1")
    up = [np.repeat(np.repeat(im, 2, axis=0), 2, axis=1) for im in [job.ref] + job.srcs]
    ref = ctx.upload(job.ref, pkg.STORE_F16)
    srcs = [ctx.upload(s, pkg.STORE_F16) for s in job.srcs]
    exp = ctx.run_view(pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                                     radius=5, stride=2, iterations=job.iterations, seed=job.seed),
                       job.shape)
    for h in [ref] + srcs:
        ctx.release(h)
    with pkg.Server(64 * MB, max_jobs=2, stride=2) as srv:
        hs = [srv.image_add(im, levels=2) for im in up]
        jid = srv.submit(hs[0], hs[1:], job.Kinv, job.A, job.b, job.dmin, job.dmax,
                         job.iterations, seed=job.seed, level=1, shape=job.shape)
        assert srv.wait(jid, 60)["state"] == pkg.JOB_DONE
        _check(srv.result(jid), exp)
    v = orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=5,
                 stride=2, top_k=3)
    o64, _, _ = v.patchmatch_fp32(job.iterations, job.seed, cost_mode=orc.COST_FP64)
    agree = (np.abs(exp["depth"].astype(np.float64) - o64) <= 0.01 * np.abs(o64)).mean()
    assert agree >= 0.99, agree
