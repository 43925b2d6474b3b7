"""GPU parity tests of the plain HIP path against the oracle (run on an MI355X).

Gates (tolerances written in the tests):
  * random init: bit-identical to the oracle (same Philox counters, same
    single-precision operations);
  * cost of identical hypotheses: |GPU - FP64 oracle| <= 1e-4 on every pixel
    the oracle does not flag as sitting on a discrete threshold; the flagged
    fraction is bounded and printed;
  * full run: >= 99 % of pixels within 1 % of the oracle's depth, where the
    oracle is the reference schedule with each cost evaluated in FP64 (the
    reference's own FP32 operation order is also compared: the GPU must be at
    least as close to it as that FP32 run is to its own FP64 twin);
  * error against the analytic ground truth no worse than the oracle's.
"""
import numpy as np
import pytest

from conftest import oracle_view, rel_within

pytestmark = pytest.mark.gpu

COST_TOL = 1e-4


def upload_job(ctx, job, store=0):
    ref = ctx.upload(job.ref, store)
    srcs = [ctx.upload(s, store) for s in job.srcs]
    return ref, srcs


def desc_for(pkg, job, ref, srcs, iterations=None, **kw):
    args = dict(radius=job.radius, stride=job.stride, top_k=job.top_k,
                iterations=job.iterations if iterations is None else iterations, seed=job.seed)
    args.update(kw)
    return pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, **args)


@pytest.fixture(scope="module")
def sphere_small(scenes):
    """C2-shaped view at quarter size: 400x300, 4 sources, stride 2, 8 iterations."""
    return scenes.sphere_job(400, 300, 4, name="sphere400")


@pytest.fixture(scope="module")
def sphere_r3(scenes):
    """3 sources, radius 3, stride 1 (other template instances)."""
    return scenes.sphere_job(200, 150, 3, radius=3, stride=1, iterations=4, scene_seed=9,
                             name="sphere200r3")


def _jobs(request, names):
    return [request.getfixturevalue(n) for n in names]


# ---------------------------------------------------------------------------
# init
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("jobname", ["c1", "sphere_small"])
def test_init_bit_exact(request, pkg, orc, ctx, jobname):
    job = request.getfixturevalue(jobname)
    ref, srcs = upload_job(ctx, job)
    desc = desc_for(pkg, job, ref, srcs)
    d_gpu, n_gpu = ctx.init_hypotheses(desc, job.shape)
    d_orc, n_orc = oracle_view(orc, job).init_fp32(job.seed)
    assert np.array_equal(d_gpu, d_orc)
    assert np.array_equal(n_gpu, n_orc)


# ---------------------------------------------------------------------------
# cost of identical hypotheses vs FP64
# ---------------------------------------------------------------------------
def hypothesis_sets(job, v, rng):
    d0, n0 = v.init_fp32(job.seed)
    gt_d, gt_n = job.gt_depth.copy(), job.gt_normal.copy()
    gt_n[gt_d <= 0] = (0, 0, -1)
    gt_d = np.where(gt_d > 0, gt_d, job.dmax)
    pert_d = (gt_d * rng.uniform(0.97, 1.03, gt_d.shape)).astype(np.float32)
    pert_n = gt_n + rng.normal(0, 0.15, gt_n.shape).astype(np.float32)
    pert_n /= np.linalg.norm(pert_n, axis=-1, keepdims=True)
    return {"random": (d0, n0), "gt": (gt_d, gt_n), "perturbed": (pert_d, pert_n.astype(np.float32))}


@pytest.mark.parametrize("jobname,store", [("c1", 0), ("sphere_small", 0), ("sphere_r3", 0),
                                           ("sphere_small", 1)])
def test_cost_matches_fp64_oracle(request, pkg, orc, ctx, jobname, store):
    job = request.getfixturevalue(jobname)
    ref, srcs = upload_job(ctx, job, store)
    desc = desc_for(pkg, job, ref, srcs)
    v = oracle_view(orc, job)
    rng = np.random.default_rng(0)
    for name, (d, n) in hypothesis_sets(job, v, rng).items():
        c_gpu = ctx.cost_eval(desc, d, n)
        c64, flags = v.cost_map_fp64(d, n)
        ok = flags == 0
        err = np.abs(c_gpu.astype(np.float64) - c64)
        print(f"{jobname} store={store} {name}: max|err| {err[ok].max():.2e} "
              f"p99.9 {np.quantile(err[ok], 0.999):.2e} flagged {(~ok).mean():.4f}")
        assert (~ok).mean() < 0.10
        assert err[ok].max() <= COST_TOL, (name, np.unravel_index(np.argmax(np.where(ok, err, 0)), err.shape))


def test_f16_storage_matches_f32(pkg, ctx, sphere_small):
    """8-bit gray levels (shifted by -128 DN) are integers, exact in FP16, and
    both storages evaluate the same FP32 bilinear lerps on the same values:
    costs and whole runs are bit-identical."""
    job = sphere_small
    r32, s32 = upload_job(ctx, job, 0)
    r16, s16 = upload_job(ctx, job, 1)
    d32 = desc_for(pkg, job, r32, s32, iterations=8)
    d16 = desc_for(pkg, job, r16, s16, iterations=8)
    d0, n0 = ctx.init_hypotheses(d32, job.shape)
    assert np.array_equal(ctx.cost_eval(d32, d0, n0), ctx.cost_eval(d16, d0, n0))
    out32 = ctx.run_view(d32, job.shape)
    out16 = ctx.run_view(d16, job.shape)
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(out16[k], out32[k]), k


def test_f16_storage_of_float_images_within_tolerance(pkg, orc, ctx, sphere_small):
    """Gray images given as floats in [0, 1] are rounded to FP16 on upload:
    costs then stay within the cost tolerance of the FP64 oracle on the
    unrounded image, except where the oracle flags a threshold."""
    job = sphere_small
    ref = ctx.upload((job.ref.astype(np.float32) + 0.37) / 255.0, 1)
    srcs = [ctx.upload((s.astype(np.float32) + 0.37) / 255.0, 1) for s in job.srcs]
    desc = desc_for(pkg, job, ref, srcs)
    d0, n0 = ctx.init_hypotheses(desc, job.shape)
    c16 = ctx.cost_eval(desc, d0, n0)
    c32 = ctx.cost_eval(desc_for(pkg, job, *upload_job(ctx, job, 0)), d0, n0)
    # an offset of 0.37 DN on every pixel leaves the NCC unchanged; FP16 of
    # (DN + 0.37) - 128 carries at most 2^-4 DN of rounding per texel
    err = np.abs(c16.astype(np.float64) - c32)
    print(f"f16 float-image cost: max|d| {err.max():.2e} mean {err.mean():.2e}")
    assert np.quantile(err, 0.999) <= 5e-3


# ---------------------------------------------------------------------------
# full runs vs the oracle's schedule
# ---------------------------------------------------------------------------
def _run_compare(pkg, orc, ctx, job, iterations, label, with_fp32=True, threads=None):
    ref, srcs = upload_job(ctx, job)
    out = ctx.run_view(desc_for(pkg, job, ref, srcs, iterations=iterations), job.shape)
    for h in [ref] + srcs:
        ctx.release(h)
    v = oracle_view(orc, job)
    d64, _, _ = v.patchmatch_fp32(iterations, job.seed, threads=threads, cost_mode=orc.COST_FP64)
    runs = [("gpu", out["depth"]), ("fp64", d64)]
    if with_fp32:
        d32, _, _ = v.patchmatch_fp32(iterations, job.seed, threads=threads,
                                      cost_mode=orc.COST_FP32)
        runs.append(("fp32", d32))
    dg = out["depth"]
    gpu_vs64 = rel_within(dg, d64).mean()
    gpu_vs32 = rel_within(dg, d32).mean() if with_fp32 else None
    o32_vs64 = rel_within(d32, d64).mean() if with_fp32 else None
    gt = job.gt_depth
    valid = gt > 0
    acc = {k: rel_within(d[valid], gt[valid]).mean() for k, d in runs}
    mre = {k: np.minimum(np.abs(d[valid] - gt[valid]) / gt[valid], 1.0).mean() for k, d in runs}
    print(f"{label}: gpu~fp64 {gpu_vs64:.4f} gpu~fp32 {gpu_vs32} fp32~fp64 {o32_vs64} "
          f"| GT<1%: {acc} | mean rel err: {mre}")
    return out, gpu_vs64, gpu_vs32, o32_vs64, acc, mre


@pytest.mark.parametrize("jobname", ["c1", "sphere_small", "sphere_r3"])
def test_depth_parity(request, pkg, orc, ctx, jobname):
    job = request.getfixturevalue(jobname)
    out, g64, g32, o3264, acc, mre = _run_compare(pkg, orc, ctx, job, job.iterations, jobname)
    assert g64 >= 0.99
    assert g32 >= o3264 - 0.003
    # error against the synthetic ground truth no worse than the oracle's
    assert acc["gpu"] >= min(acc["fp64"], acc["fp32"]) - 0.003
    assert mre["gpu"] <= max(mre["fp64"], mre["fp32"]) * 1.02 + 1e-4
    assert np.all((out["depth"] >= job.dmin) & (out["depth"] <= job.dmax))


def test_depth_parity_converged_c1(pkg, orc, ctx, c1):
    """C1 at 8 iterations: also >= 99 % against the reference-order FP32 run."""
    out, g64, g32, o3264, acc, mre = _run_compare(pkg, orc, ctx, c1, 8, "c1-8it")
    assert g64 >= 0.99 and g32 >= 0.99


def test_depth_parity_c2_full_size(pkg, orc, ctx, scenes):
    """The benchmark workload itself (1600x1200, 4 sources, 11x11 at stride 2,
    8 iterations): >= 99 % of pixels within 1 % of the FP64-cost oracle run,
    error against the ground truth no worse.  The oracle takes about half a
    minute on 16 host threads here."""
    import os
    job = scenes.c2()
    threads = min(16, os.cpu_count() or 1)
    out, g64, _, _, acc, mre = _run_compare(pkg, orc, ctx, job, job.iterations, "c2-full",
                                            with_fp32=False, threads=threads)
    assert g64 >= 0.99
    assert acc["gpu"] >= acc["fp64"] - 0.003
    assert mre["gpu"] <= mre["fp64"] * 1.02 + 1e-4


# ---------------------------------------------------------------------------
# run-level properties
# ---------------------------------------------------------------------------
def test_runs_are_deterministic(pkg, ctx, sphere_small):
    job = sphere_small
    ref, srcs = upload_job(ctx, job)
    desc = desc_for(pkg, job, ref, srcs, iterations=2)
    a = ctx.run_view(desc, job.shape)
    b = ctx.run_view(desc, job.shape)
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(a[k], b[k])


def test_iterations_never_raise_cost_and_keep_invariants(pkg, ctx, sphere_small):
    job = sphere_small
    ref, srcs = upload_job(ctx, job)
    d0, n0 = ctx.init_hypotheses(desc_for(pkg, job, ref, srcs), job.shape)
    c0 = ctx.cost_eval(desc_for(pkg, job, ref, srcs), d0, n0)
    prev = c0
    for iters in (1, 2, 4):
        out = ctx.run_view(desc_for(pkg, job, ref, srcs, iterations=iters), job.shape,
                           with_confidence=True)
        assert np.all(out["cost"] <= prev)
        prev = out["cost"]
        ln = np.linalg.norm(out["normal"], axis=-1)
        assert np.mean(np.abs(ln - 1) < 1e-3) > 0.999
        np.testing.assert_array_equal(out["confidence"], 1.0 - out["cost"] * 0.5)
    assert prev.mean() < 0.5 * c0.mean()


def test_given_init_reproduces_random_init_run(pkg, ctx, sphere_small):
    job = sphere_small
    ref, srcs = upload_job(ctx, job)
    desc = desc_for(pkg, job, ref, srcs, iterations=2)
    d0, n0 = ctx.init_hypotheses(desc, job.shape)
    a = ctx.run_view(desc, job.shape)
    b = ctx.run_view(desc_for(pkg, job, ref, srcs, iterations=2, init=(d0, n0)), job.shape)
    assert np.array_equal(a["depth"], b["depth"]) and np.array_equal(a["cost"], b["cost"])


def test_device_resident_outputs(pkg, ctx, sphere_small):
    job = sphere_small
    ref, srcs = upload_job(ctx, job)
    desc = desc_for(pkg, job, ref, srcs, iterations=1)
    H, W = job.shape
    before = ctx.stats()["hbm_in_use"]
    dd, dn, dc = (ctx.device_alloc(H * W * 4 * m) for m in (1, 3, 1))
    assert ctx.stats()["hbm_in_use"] == before + H * W * 20
    ctx.run_view_device(desc, dd, dn, dc)
    host = ctx.run_view(desc, job.shape)
    assert np.array_equal(ctx.download(dd, (H, W)), host["depth"])
    assert np.array_equal(ctx.download(dn, (H, W, 3)), host["normal"])
    assert np.array_equal(ctx.download(dc, (H, W)), host["cost"])
    for p in (dd, dn, dc):
        ctx.device_free(p)
    with pytest.raises(pkg.HP3DError) as e:
        ctx.device_free(dd)
    assert e.value.code == pkg.ERR_NOT_FOUND


def test_full_size_c2_properties(pkg, ctx, scenes):
    job = scenes.c2()
    ref, srcs = upload_job(ctx, job)
    out = ctx.run_view(desc_for(pkg, job, ref, srcs), job.shape)
    d = out["depth"]
    assert np.all((d >= 2.0) & (d <= 12.0))
    acc = rel_within(d, job.gt_depth).mean()
    print(f"C2 1600x1200 8 iters: {acc:.4f} of pixels within 1 % of GT, mean cost {out['cost'].mean():.4f}")
    assert acc > 0.6
    st = ctx.stats()
    assert st["n_propagate"] == 16 and st["taps"] == job.taps()


# ---------------------------------------------------------------------------
# error paths through the C ABI
# ---------------------------------------------------------------------------
def test_error_codes(pkg, ctx, c1):
    job = c1
    ref, srcs = upload_job(ctx, job)
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_view(desc_for(pkg, job, ref, []), job.shape)
    assert e.value.code == pkg.ERR_INVALID_ARG
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_view(desc_for(pkg, job, ref, srcs, radius=6), job.shape)
    assert e.value.code == pkg.ERR_UNSUPPORTED
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_view(desc_for(pkg, job, ref, [srcs[0], 987654]), job.shape)
    assert e.value.code == pkg.ERR_NOT_FOUND
    bad = desc_for(pkg, job, ref, srcs)
    bad.depth_min, bad.depth_max = 5.0, 5.0
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_view(bad, job.shape)
    assert e.value.code == pkg.ERR_INVALID_ARG
    # iteration counts past 2^20 are refused before anything is launched
    for field in ("iterations", "first_iteration"):
        bad = desc_for(pkg, job, ref, srcs)
        setattr(bad, field, (1 << 20) + 1)
        with pytest.raises(pkg.HP3DError) as e:
            ctx.run_view(bad, job.shape)
        assert e.value.code == pkg.ERR_INVALID_ARG, field
    small = pkg.Context(hbm_budget=64 << 10)     # below one 320x240 plane
    with pytest.raises(pkg.HP3DError) as e:
        small.upload(job.ref)
    assert e.value.code == pkg.ERR_OOM_BUDGET
    small.close()
    other = ctx.upload(np.zeros((10, 10), np.uint8))
    with pytest.raises(pkg.HP3DError) as e:
        ctx.run_view(desc_for(pkg, job, ref, [other]), job.shape)
    assert e.value.code == pkg.ERR_INVALID_ARG
    ctx.release(other)
    with pytest.raises(pkg.HP3DError):
        ctx.release(other)
