"""Seeded random configurations of the plain path against the oracle (MI355X).

880 small views from nineteen fixed generator seeds (16 from one, 48 from
each of eighteen more): size (odd and even, down to a few windows across), 1..8
sources at random arc angles, window radius 1..5, stride 1 or 2, top-k 1..8,
both storages, 1..3 iterations, and a first iteration of 0 or 3.  Each must
give the oracle's random init bit for bit, costs of identical hypotheses
within 1e-4 of the FP64 oracle on every unflagged pixel, whatever the window
size and however few sources are averaged (the reference's own FP32
operation order is printed beside it, on a sample), and either a short run
within 1 % of the FP64-cost oracle's depth on >= 99 % of pixels (or, below
that, on at least the share the reference's FP32 operation order reaches on
the same configuration) or, for first iteration 3, a resumed run equal to the
uninterrupted one bit for bit.  Every configuration is printed before it runs.

HP3D_RANDOM_CONFIG_SEEDS (a comma-separated list) adds 48 configurations from
each further generator seed, for wider sweeps.
"""
import os

import numpy as np
import pytest

from conftest import oracle_view, rel_within

pytestmark = pytest.mark.gpu


def _configs(n=16, seed=20261202):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        radius = int(rng.integers(1, 6))
        stride = int(rng.choice([1, 2])) if radius >= 2 else 1
        W = int(rng.integers(2 * radius + 12, 150))
        H = int(rng.integers(2 * radius + 8, 110))
        S = int(rng.integers(1, 9))
        out.append(dict(idx=i, W=W, H=H, S=S, radius=radius, stride=stride,
                        top_k=int(rng.integers(1, 9)), store=int(rng.integers(0, 2)),
                        iterations=int(rng.integers(1, 4)),
                        first=int(rng.choice([0, 0, 3])),
                        angles=sorted(float(a) for a in rng.uniform(-15.0, 15.0, S)),
                        scene_seed=int(rng.integers(1, 1000))))
    return out


DEFAULT_SWEEP_SEEDS = tuple(range(7, 25))


def _all_configs():
    out = _configs()
    extra = [int(s) for s in filter(None, os.environ.get("HP3D_RANDOM_CONFIG_SEEDS", "").split(","))]
    for s in list(DEFAULT_SWEEP_SEEDS) + [s for s in extra if s not in DEFAULT_SWEEP_SEEDS]:
        out += [dict(c, sweep_seed=s) for c in _configs(48, seed=s)]
    return out


def _id(c):
    return f"s{c['sweep_seed']}-cfg{c['idx']}" if "sweep_seed" in c else f"cfg{c['idx']}"


@pytest.mark.parametrize("cfg", _all_configs(), ids=_id)
def test_random_config_matches_oracle(pkg, orc, ctx, scenes, cfg):
    print(f"config {cfg}", flush=True)
    job = scenes.sphere_job(cfg["W"], cfg["H"], cfg["S"], f=0.9375 * max(cfg["W"], 64),
                            radius=cfg["radius"], stride=cfg["stride"],
                            iterations=cfg["iterations"], angles=cfg["angles"],
                            scene_seed=cfg["scene_seed"], name=f"rand{cfg['idx']}")
    tol = 1e-4
    v = oracle_view(orc, job, top_k=cfg["top_k"])
    ref = ctx.upload(job.ref, cfg["store"])
    srcs = [ctx.upload(s, cfg["store"]) for s in job.srcs]
    d = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                      radius=job.radius, stride=job.stride, iterations=job.iterations,
                      top_k=cfg["top_k"], seed=job.seed)
    d0, n0 = ctx.init_hypotheses(d, job.shape)
    o0, on = v.init_fp32(job.seed)
    assert np.array_equal(d0, o0) and np.array_equal(n0, on)
    c = ctx.cost_eval(d, d0, n0)
    c64, flags = v.cost_map_fp64(d0, n0)
    ok = flags == 0
    err = np.abs(c.astype(np.float64) - c64)[ok]
    # enough unflagged pixels for the gate to mean something (small views
    # with an 11x11 window have most taps near a border)
    assert ok.sum() >= min(200, ok.size // 3), int(ok.sum())
    ys, xs = np.nonzero(ok)
    pick = np.random.default_rng(cfg["idx"]).choice(len(xs), min(len(xs), 1500), replace=False)
    ref32 = max(abs(v.cost_fp32(int(xs[i]), int(ys[i]), float(d0[ys[i], xs[i]]),
                                n0[ys[i], xs[i]]) - c64[ys[i], xs[i]]) for i in pick)
    print(f"  cost max|err| {err.max():.2e}; reference-order FP32 on a sample {ref32:.2e}")
    assert err.max() <= tol, float(err.max())
    if cfg["first"]:
        # resumed: `first` iterations, then the rest from the saved state with
        # first_iteration set, must equal the uninterrupted run bit for bit
        first, iters = cfg["first"], job.iterations
        part = ctx.run_view(pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                                          radius=job.radius, stride=job.stride, iterations=first,
                                          top_k=cfg["top_k"], seed=job.seed), job.shape)
        rest = ctx.run_view(pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                                          radius=job.radius, stride=job.stride, iterations=iters,
                                          top_k=cfg["top_k"], seed=job.seed,
                                          init=(part["depth"], part["normal"]),
                                          first_iteration=first), job.shape)
        full = ctx.run_view(pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                                          radius=job.radius, stride=job.stride,
                                          iterations=first + iters, top_k=cfg["top_k"],
                                          seed=job.seed), job.shape)
        for key in ("depth", "normal", "cost"):
            assert np.array_equal(rest[key], full[key]), key
        out = full
    else:
        out = ctx.run_view(d, job.shape)
        o64, _, _ = v.patchmatch_fp32(job.iterations, job.seed, cost_mode=orc.COST_FP64)
        agree = rel_within(out["depth"], o64).mean()
        print(f"  depth within 1 % of the oracle on {agree:.4f}")
        if agree < 0.99:
            # a near-tie decided the other way early in a short run on a small
            # view moves every pixel it propagates to; the bar is then the
            # reference's own FP32 operation order on the same configuration
            o32, _, _ = v.patchmatch_fp32(job.iterations, job.seed, cost_mode=orc.COST_FP32)
            agree32 = rel_within(o32, o64).mean()
            print(f"  reference-order FP32 run within 1 % of the FP64 run on {agree32:.4f}")
            assert agree >= agree32, (agree, agree32)
    for h in [ref] + srcs:
        ctx.release(h)
    assert np.all((out["depth"] >= job.dmin) & (out["depth"] <= job.dmax))
