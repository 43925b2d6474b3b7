"""Benchmark of the MI355X PatchMatch-MVS path (one JSON line on stdout).

Workload (default): one 1600x1200 reference view with 4 sources, 11x11
window at stride 2, depth range [2, 12], 8 red/black iterations, FP32
(the "plain GPU drop-in" configuration).  A step is one full solve of that
view -- random init, initial cost, 8 iterations, outputs written to device
buffers -- with the images already resident in HBM.

value = steps * W * H * iterations / timed seconds / 1e6   (Mpixel-iters/s)

roofline: the dominant kernel is the half-pass (pm_halfpass_kernel).
  algorithmic work per launch = (W*H/2 active pixels) * 6 candidates * S
  sources * k taps * 34 FLOP/tap; its average launch duration is measured
  with HIP events recorded by the library on its own stream during the timed
  steps.  Peak: 157.3 TFLOP/s (MI355X FP32 vector, datasheet).
cpu_baseline: the oracle's FP32 restatement of the reference algorithm on
  host threads, timed on a bounded sample of the same view (2 iterations),
  scaled through its measured tap rate to this workload's taps per
  Mpixel-iteration.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
FLOP_PER_TAP = 34
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_halfpass_c2.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--store", choices=["f32", "f16"], default="f32")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink W, H (not for reporting)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (counter calibration)")
    return ap.parse_args()


def copy_probe(mib):
    """One plain device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the traffic counters."""
    import __graft_entry__ as entry
    pkg = entry.load_package()
    ctx = pkg.Context()
    n = mib << 20
    src, dst = ctx.device_alloc(n), ctx.device_alloc(n)
    ctx.upload_bytes(src, np.zeros(n, np.uint8))
    import ctypes
    rc = pkg.lib().hp3d_pm_device_copy(ctx.h, ctypes.c_void_p(dst), ctypes.c_void_p(src), n, 2)
    assert rc == 0, ctx.last_error()
    ctx.close()
    print(json.dumps({"copy_probe_bytes": n}))


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this path runs on one GPU; multi-GPU sharding is a design note only, "
              "so --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    import __graft_entry__ as entry
    from importlib import import_module
    pkg = entry.load_package()
    scenes = import_module(entry.PKG_NAME + ".scenes")

    job = scenes.c2(threads=a.cpu_threads, scale=a.scale)
    H, W = job.shape
    store = pkg.STORE_F16 if a.store == "f16" else pkg.STORE_F32
    n_taps = (2 * job.radius) // job.stride + 1
    k = n_taps * n_taps

    ctx = pkg.Context(timing=True)
    ref = ctx.upload(job.ref, store)
    srcs = [ctx.upload(s, store) for s in job.srcs]
    desc = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                         radius=job.radius, stride=job.stride, iterations=job.iterations,
                         seed=job.seed)
    # device-resident outputs; every run_view returns after its stream drained,
    # so the host clock brackets exactly the timed solves
    dd, dn, dc = (ctx.device_alloc(W * H * 4 * m) for m in (1, 3, 1))

    for _ in range(a.warmup):
        ctx.run_view_device(desc, dd, dn, dc)
    lat, prop_ms, prop_n = [], 0.0, 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.run_view_device(desc, dd, dn, dc)
        st = ctx.stats()
        lat.append(st["ms_total"])
        prop_ms += st["ms_propagate"]
        prop_n += st["n_propagate"]
    dt = time.perf_counter() - t0
    mpix_iters = a.steps * W * H * job.iterations / 1e6
    value = mpix_iters / dt

    # roofline of the half-pass kernel
    taps_per_launch = (W * H / 2.0) * 6 * job.S * k
    avg_launch_s = prop_ms / prop_n / 1e3
    achieved = taps_per_launch * FLOP_PER_TAP / avg_launch_s / 1e12
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and a.scale == 1.0 and a.store == "f32":
        traffic = json.load(open(TRAFFIC_FILE)).get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "pm_halfpass_kernel",
                "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "flop_per_launch": taps_per_launch * FLOP_PER_TAP}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(entry, job, a.cpu_threads, k)

    lat = np.array(lat)
    out = {
        "metric": "PatchMatch Mpixel-iters/s at <=1% depth dev; p50/p99 job latency",
        "value": round(value, 2),
        "unit": "Mpixel-iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded plane+sphere scene, uint8 images rendered in-process)",
        "config": {"workload": "plain GPU drop-in: 1 ref + 4 src, 1600x1200 uint8, 11x11 "
                               "window stride 2, depth [2,12], 8 iterations",
                   "width": W, "height": H, "sources": job.S, "radius": job.radius,
                   "stride": job.stride, "iterations": job.iterations,
                   "storage": a.store, "outputs": "device"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


def cpu_baseline(entry, job, threads, k):
    """Oracle FP32 restatement on host threads: 2 iterations of the same view."""
    orc = entry.load_oracle()
    v = orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                 radius=job.radius, stride=job.stride, top_k=job.top_k)
    H, W = job.shape
    iters = 2
    t0 = time.perf_counter()
    v.patchmatch_fp32(iters, job.seed, threads=threads, cost_mode=orc.COST_FP32)
    dt = time.perf_counter() - t0
    taps = W * H * (job.S * k + iters * 6 * job.S * k)
    taps_per_s = taps / dt
    taps_per_mpix_iter = (job.S * k + job.iterations * 6 * job.S * k) / job.iterations * 1e6
    return {"value": round(taps_per_s / taps_per_mpix_iter, 3), "unit": "Mpixel-iters/s",
            "cores": threads, "kind": "port",
            "sample": f"same {W}x{H} view, {iters} iterations ({taps / 1e9:.2f} G taps) in "
                      f"{dt:.1f} s; scaled by the tap rate to the {job.iterations}-iteration workload",
            "taps_per_s": round(taps_per_s)}


if __name__ == "__main__":
    main()
