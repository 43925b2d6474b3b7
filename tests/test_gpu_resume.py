"""Saving a run to PFM and resuming it (MI355X).

A run of 2k iterations must equal a run of k iterations whose depth and
normals are written as PFM files ("Pf" depth, "PF" normals), read back and
resumed with first_iteration = k for k more: the cost of a hypothesis is a
deterministic function of it, the refinement ranges and random streams
depend only on the absolute iteration index, so the two are bit-identical.
The reference writes depth only and cannot resume (patchmatch_mvs.cu:1254-1257).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("store", [0, 1], ids=["f32", "f16"])
def test_resume_from_pfm_is_bit_identical(pkg, ctx, scenes, tmp_path, store):
    job = scenes.sphere_job(320, 240, 4, scene_seed=21, name="resume")
    ref = ctx.upload(job.ref, store)
    srcs = [ctx.upload(s, store) for s in job.srcs]

    def desc(iterations, **kw):
        return pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                             radius=job.radius, stride=job.stride, iterations=iterations,
                             seed=job.seed, **kw)

    full = ctx.run_view(desc(6), job.shape)
    half = ctx.run_view(desc(3), job.shape)
    pkg.pfm_save_ex(str(tmp_path / "depth.pfm"), half["depth"])
    pkg.pfm_save_ex(str(tmp_path / "normal.pfm"), half["normal"])
    d = pkg.pfm_load_ex(str(tmp_path / "depth.pfm"))
    n = pkg.pfm_load_ex(str(tmp_path / "normal.pfm"))
    resumed = ctx.run_view(desc(3, init=(d, n), first_iteration=3), job.shape)
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(resumed[k], full[k]), k
    # resuming at the wrong index changes the schedule (sanity of the test)
    wrong = ctx.run_view(desc(3, init=(d, n), first_iteration=0), job.shape)
    assert not np.array_equal(wrong["depth"], full["depth"])
    for h in [ref] + srcs:
        ctx.release(h)
