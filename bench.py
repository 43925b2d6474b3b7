"""Benchmark of the MI355X PatchMatch-MVS path (one JSON line on stdout).

Workload (default): one 1600x1200 reference view with 4 sources, 11x11
window at stride 2, depth range [2, 12], 8 red/black iterations, FP32
(the "plain GPU drop-in" configuration).  A step is one full solve of that
view -- random init, initial cost, 8 iterations, finalize -- with the images
already resident in HBM, and its depth, normal and cost maps downloaded to
host memory (the reference downloads them per view, patchmatch_mvs.cu:
1243-1252).  The timed steps run as one multi-view call (hp3d_pm_run_views):
step i's download overlaps step i + 1's solve, every download lands in its
own page-locked host buffers, and the timed region ends when the last one
has.

value = steps * W * H * iterations / timed seconds / 1e6   (Mpixel-iters/s)
Also reported: device_outputs_value (outputs left in HBM) and
serial_host_value (one single-view call per step, download not overlapped).

roofline: the dominant kernel is the half-pass (pm_halfpass_kernel).
  algorithmic work per launch = (W*H/2 active pixels) * 6 candidates * S
  sources * k taps * 34 FLOP/tap (the reference's work: candidates the kernel
  skips as provably rejected are counted); its average launch duration is measured
  with HIP events recorded by the library on its own stream during the timed
  steps.  Peak: 157.3 TFLOP/s (MI355X FP32 vector, datasheet).  The
  kernel issues no MFMA and is bound by VALU issue, so bound = "valu".
  achieved_executed: the same over the taps the kernel executes (counter
  SQ_INSTS_VALU_TRANS_F32, one v_rcp_f32 per tap, x 64 lanes).
  traffic, hbm_gbs, l2_read_gbs: HBM bytes (FETCH_SIZE + WRITE_SIZE) and
  L2-to-L1 read bytes (TCP_TCC_READ_REQ) per launch from rocprofv3 counter
  runs of this workload, calibrated on a 1 GiB copy and kept in
  profiles/traffic_halfpass_c2.json, over the live launch time; peaks
  8 TB/s (HBM3E) and 39.3 TB/s (L2: 8 XCDs x 16 channels x 128 B/clk x
  2.4 GHz).
cpu_baseline: the oracle's FP32 restatement of the reference algorithm on
  host threads, the whole 8-iteration solve of the same view.
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
PEAK_HBM_GBS = 8000.0
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
    ap.add_argument("--form", choices=["plain", "precision", "batched", "serving", "geom"],
                    default="plain",
                    help="plain: the headline line; the others print their own JSON line")
    ap.add_argument("--views", type=int, default=48, help="batched: views per batch")
    ap.add_argument("--chunk", type=int, default=16,
                    help="batched: units a persistent workgroup takes at a time")
    ap.add_argument("--rate", type=float, default=20.0, help="serving: arrivals per second")
    ap.add_argument("--duration", type=float, default=120.0, help="serving: arrival window (s)")
    ap.add_argument("--stride", type=int, default=None,
                    help="batched/serving window stride (default: 1 batched, 2 serving)")
    ap.add_argument("--max-jobs", type=int, default=64, help="serving: concurrent job slots")
    ap.add_argument("--variant", type=int, default=-1,
                    help="batched/serving: step-kernel variant (-1: autotuned / default)")
    ap.add_argument("--form-store", choices=["f32", "f16"], default="f16",
                    help="batched/serving image storage")
    ap.add_argument("--no-forms", action="store_true",
                    help="plain line only: skip the batched and serving sub-objects")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="arrival window of each serving sub-object of the default line")
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
    if a.form != "plain":
        {"precision": form_precision, "batched": form_batched,
         "serving": form_serving, "geom": form_geom}[a.form](a, entry, pkg, scenes)
        return

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
    descs = [desc] * a.steps
    # one set of page-locked output buffers per timed step (the outputs of
    # every step stay valid after the run), one more set for the warm-up
    outs = [ctx.host_outputs(job.shape, pinned=True) for _ in range(a.steps)]
    warm = [ctx.host_outputs(job.shape, pinned=True)]
    if a.warmup:
        ctx.run_views([desc] * a.warmup, warm * a.warmup)
    t0 = time.perf_counter()
    ctx.run_views(descs, outs)
    dt = time.perf_counter() - t0
    st = ctx.stats()
    prop_ms, prop_n = st["ms_propagate"], st["n_propagate"]
    mpix_iters = a.steps * W * H * job.iterations / 1e6
    value = mpix_iters / dt
    # every step's outputs arrived: all identical (same view, same seed) and
    # in range
    for o in outs[1:]:
        assert np.array_equal(o["depth"], outs[0]["depth"])
    assert np.all((outs[0]["depth"] >= job.dmin) & (outs[0]["depth"] <= job.dmax))

    # the same solves one call per step, download not overlapped (latency of
    # a single view), and with outputs left in device memory
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t1 = time.perf_counter()
        ctx.run_views([desc], [outs[0]])
        lat.append((time.perf_counter() - t1) * 1e3)
    serial_value = mpix_iters / (time.perf_counter() - t0)
    dd, dn, dc = (ctx.device_alloc(W * H * 4 * m) for m in (1, 3, 1))
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.run_view_device(desc, dd, dn, dc)
    device_value = mpix_iters / (time.perf_counter() - t0)

    # roofline of the half-pass kernel
    taps_per_launch = (W * H / 2.0) * 6 * job.S * k
    avg_launch_s = prop_ms / prop_n / 1e3
    achieved = taps_per_launch * FLOP_PER_TAP / avg_launch_s / 1e12
    traffic, prof = None, {}
    if os.path.exists(TRAFFIC_FILE) and a.scale == 1.0 and a.store == "f32":
        prof = json.load(open(TRAFFIC_FILE))
        traffic = prof.get("hbm_bytes_per_launch")
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic, "kernel": "pm_halfpass_kernel",
                "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "flop_per_launch": taps_per_launch * FLOP_PER_TAP}
    # memory-side rates of the same launches: counter bytes per launch (profiled
    # once, rocprofv3) over the live launch time
    if traffic:
        gbs = traffic / avg_launch_s / 1e9
        roofline.update(hbm_gbs=round(gbs, 1), hbm_frac=round(gbs / PEAK_HBM_GBS, 4))
    if prof.get("executed_taps_per_launch"):
        # the taps the kernel executes (one v_rcp_f32 each, counted by
        # SQ_INSTS_VALU_TRANS_F32 x 64 lanes; candidates it proves rejected
        # are not evaluated), beside the algorithmic taps above
        ex = prof["executed_taps_per_launch"] * FLOP_PER_TAP / avg_launch_s / 1e12
        roofline.update(executed_taps_per_launch=prof["executed_taps_per_launch"],
                        algorithmic_taps_per_launch=taps_per_launch,
                        achieved_executed=round(ex, 3),
                        frac_executed=round(ex / PEAK_FP32_TFLOPS, 4))
    if prof.get("l2_read_bytes_per_launch"):
        l2 = prof["l2_read_bytes_per_launch"] / avg_launch_s / 1e9
        roofline.update(l2_read_gbs=round(l2, 1), l2_frac=round(l2 / prof["l2_peak_gbs"], 4),
                        l2_hit_rate=prof.get("l2_hit_rate"), l1_hit_rate=prof.get("l1_hit_rate"))

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(entry, job, a.cpu_threads, k)
    forms = {}
    if not a.no_forms:
        ctx.close()
        print("bench: batched workload", file=sys.stderr, flush=True)
        forms["batched"] = measure_batched(a, entry, pkg, scenes, views=48, stride=1, reps=1,
                                           with_fp32=False, with_cpu=False)
        for stride in (2, 1):
            print(f"bench: serving workload, stride {stride}", file=sys.stderr, flush=True)
            forms[f"serving_stride{stride}"] = measure_serving(
                a, entry, pkg, scenes, rate=20.0, duration=a.serve_seconds, stride=stride)

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
                   "storage": a.store,
                   "outputs": "depth, normal, cost downloaded to page-locked host memory "
                              "inside the timed region"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "device_outputs_value": round(device_value, 2),
        "serial_host_value": round(serial_value, 2),
    }
    out.update(forms)
    print(json.dumps(out))
    if a.no_forms:
        ctx.close()


def cpu_baseline(entry, job, threads, k):
    """Oracle FP32 restatement on host threads: the whole solve of the same view."""
    orc = entry.load_oracle()
    v = orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                 radius=job.radius, stride=job.stride, top_k=job.top_k)
    H, W = job.shape
    iters = job.iterations
    t0 = time.perf_counter()
    v.patchmatch_fp32(iters, job.seed, threads=threads, cost_mode=orc.COST_FP32)
    dt = time.perf_counter() - t0
    taps = W * H * (job.S * k + iters * 6 * job.S * k)
    out = {"value": round(W * H * iters / dt / 1e6, 3), "unit": "Mpixel-iters/s",
           "cores": threads, "kind": "port",
           "sample": f"the whole {W}x{H} {iters}-iteration solve ({taps / 1e9:.2f} G taps) "
                     f"in {dt:.1f} s",
           "taps_per_s": round(taps / dt)}
    out.update(host_cpu())
    # the CPU-runnable workload (320x240, 2 sources, 11x11 stride 1, 3
    # iterations) run whole, on one core and on `threads`
    from importlib import import_module
    c1 = import_module(entry.PKG_NAME + ".scenes").c1(threads=threads)
    v1 = orc.View(c1.ref, c1.srcs, c1.Kinv, c1.A, c1.b, c1.dmin, c1.dmax, radius=c1.radius,
                  stride=c1.stride, top_k=c1.top_k)
    H1, W1 = c1.shape
    for n in (1, threads):
        t0 = time.perf_counter()
        v1.patchmatch_fp32(c1.iterations, c1.seed, threads=n, cost_mode=orc.COST_FP32)
        dt1 = time.perf_counter() - t0
        out[f"c1_{n}core" + ("s" if n > 1 else "")] = {
            "value": round(W1 * H1 * c1.iterations / dt1 / 1e6, 4), "unit": "Mpixel-iters/s",
            "seconds": round(dt1, 2), "taps_per_s": round(c1.taps() / dt1)}
    return out


def host_cpu():
    """CPU model and logical CPU count of the host the baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}



# ---------------------------------------------------------------------------
# The other forms (each prints one JSON line of its own)
# ---------------------------------------------------------------------------
def _taps_per_launch(job):
    n = (2 * job.radius) // job.stride + 1
    H, W = job.shape
    return (W * H / 2.0) * 6 * job.S * n * n


def _cpu_tap_rate(entry, threads):
    """CPU baseline tap rate: the oracle's FP32 restatement on a 640x480
    4-source view for 1 iteration (a few seconds), in taps per second."""
    from importlib import import_module
    scenes = import_module(entry.PKG_NAME + ".scenes")
    orc = entry.load_oracle()
    job = scenes.sphere_job(640, 480, 4, radius=5, stride=1, iterations=1, threads=threads)
    v = orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax, radius=5,
                 stride=1, top_k=3)
    t0 = time.perf_counter()
    v.patchmatch_fp32(1, job.seed, threads=threads, cost_mode=orc.COST_FP32)
    return job.taps() / (time.perf_counter() - t0)


def _device_outputs(ctx, shape):
    H, W = shape
    return tuple(ctx.device_alloc(W * H * 4 * m) for m in (1, 3, 1))


def form_precision(a, entry, pkg, scenes):
    """3200x2400 view, 6 sources at +-5/10/15 deg, stride 2, 8 iterations:
    FP16 image storage (FP32 accumulation, FP64 host homographies) against
    the FP32 plain storage on the same view."""
    job = scenes.c3(threads=a.cpu_threads)
    H, W = job.shape
    ctx = pkg.Context(timing=True)
    res = {}
    for name, store in (("fp32", pkg.STORE_F32), ("fp16", pkg.STORE_F16)):
        ref = ctx.upload(job.ref, store)
        srcs = [ctx.upload(s, store) for s in job.srcs]
        desc = pkg.make_desc(ref, srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                             radius=job.radius, stride=job.stride, iterations=job.iterations,
                             seed=job.seed)
        outs = _device_outputs(ctx, job.shape)
        ctx.run_view_device(desc, *outs)
        steps, prop_ms, prop_n = max(2, a.steps // 4), 0.0, 0
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.run_view_device(desc, *outs)
            st = ctx.stats()
            prop_ms += st["ms_propagate"]
            prop_n += st["n_propagate"]
        dt = (time.perf_counter() - t0) / steps
        depth = ctx.download(outs[0], (H, W))
        achieved = _taps_per_launch(job) * FLOP_PER_TAP / (prop_ms / prop_n / 1e3) / 1e12
        res[name] = {"mpix_iters_per_s": round(job.mpix_iters() / dt, 2),
                     "ms_per_view": round(dt * 1e3, 3),
                     "halfpass_tflops": round(achieved, 3),
                     "gt_within_1pct": round(float((np.abs(depth - job.gt_depth)
                                                    <= 0.01 * job.gt_depth).mean()), 4),
                     "depth": depth}
        for p in outs:
            ctx.device_free(p)
        ctx.release(ref)
        for s in srcs:
            ctx.release(s)
    d32, d16 = res["fp32"].pop("depth"), res["fp16"].pop("depth")
    agree = float((np.abs(d16 - d32) <= 0.01 * np.abs(d32)).mean())
    out = {"form": "precision", "metric": "PatchMatch Mpixel-iters/s",
           "value": res["fp16"]["mpix_iters_per_s"], "unit": "Mpixel-iters/s",
           "dtype": "fp16 storage / fp32 accumulation", "data": "synthetic",
           "config": {"workload": "precision form: 1 ref + 6 src, 3200x2400 uint8, 11x11 stride 2, "
                                  "depth [2,12], 8 iterations", "width": W, "height": H,
                      "sources": job.S},
           "fp16": res["fp16"], "fp32": res["fp32"],
           "speedup_fp16_vs_fp32": round(res["fp16"]["mpix_iters_per_s"]
                                         / res["fp32"]["mpix_iters_per_s"], 3),
           "depth_within_1pct_fp16_vs_fp32": round(agree, 5),
           "roofline": {"bound": "valu", "achieved": res["fp16"]["halfpass_tflops"],
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(res["fp16"]["halfpass_tflops"] / PEAK_FP32_TFLOPS, 4)}}
    if not a.no_cpu:
        # too large to run on the host: extrapolated from the measured tap rate
        rate = _cpu_tap_rate(entry, a.cpu_threads)
        cpu = rate / job.taps() * job.mpix_iters()
        out["cpu_baseline"] = {"value": round(cpu, 3), "unit": "Mpixel-iters/s",
                               "cores": a.cpu_threads, "kind": "port",
                               "sample": "640x480 4-source view, 1 iteration, extrapolated by "
                                         "taps to this view"}
        out["speedup_vs_cpu"] = round(out["value"] / cpu, 1)
    print(json.dumps(out))
    ctx.close()


def measure_batched(a, entry, pkg, scenes, views=48, stride=1, reps=2, with_fp32=True,
                    with_cpu=True):
    """The batched workload: `views` views, resolutions uniform over {640x480 ...
    4000x3000}, 3-8 sources, 11x11 window, 8 iterations, FP16 storage, the
    whole batch in one launch captured in a hipGraph -- the persistent kernel
    and the graph of step launches both timed, the faster reported -- tile
    shape chosen by the autotuner; against the plain kernel run view by view
    over the same views.  Outputs stay in device memory."""
    specs = scenes.batch_specs(views, iterations=8, stride=stride)
    t0 = time.perf_counter()
    jobs = [scenes.job_from_spec(s, threads=a.cpu_threads) for s in specs]
    render_s = time.perf_counter() - t0
    ctx = pkg.Context(timing=True)
    mpix = sum(j.mpix_iters() for j in jobs)
    taps = sum(j.taps() for j in jobs)

    def descs_for(store):
        out = []
        for j in jobs:
            ref = ctx.upload(j.ref, store)
            srcs = [ctx.upload(s, store) for s in j.srcs]
            out.append(pkg.make_desc(ref, srcs, j.Kinv, j.A, j.b, j.dmin, j.dmax, radius=j.radius,
                                     stride=j.stride, iterations=j.iterations, seed=j.seed))
        return out

    fstore = pkg.STORE_F16 if a.form_store == "f16" else pkg.STORE_F32
    d16 = descs_for(fstore)
    outs = [_device_outputs(ctx, j.shape) for j in jobs]
    if a.variant >= 0:
        best, times = a.variant, np.full(pkg.NUM_VARIANTS, -1.0)
    else:
        best, times = ctx.autotune(d16)

    def timed_batch(**kw):
        ctx.batch_run_device(d16, outs, variant=best, **kw)     # warm-up (captures the graph)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.batch_run_device(d16, outs, variant=best, **kw)
        return (time.perf_counter() - t0) / reps
    # one persistent kernel for the whole batch, and the graph of step launches
    persist_s = timed_batch(persistent=True, chunk=a.chunk)
    steps_s = timed_batch(persistent=False)
    batched_s = min(persist_s, steps_s)

    # plain kernel one view after the other: the batch's storage (and FP32)
    def serial(descs):
        ctx.run_view_device(descs[0], *outs[0])
        t0 = time.perf_counter()
        for d, o in zip(descs, outs):
            ctx.run_view_device(d, *o)
        return time.perf_counter() - t0
    plain_same_s = serial(d16)
    plain_s = None
    if with_fp32 and fstore != pkg.STORE_F32:
        plain_s = serial(descs_for(pkg.STORE_F32))
    cpu = None
    if with_cpu and not a.no_cpu:
        rate = _cpu_tap_rate(entry, a.cpu_threads)
        cpu = {"value": round(mpix / (taps / rate), 3), "unit": "Mpixel-iters/s",
               "cores": a.cpu_threads, "kind": "port",
               "sample": "640x480 4-source view, 1 iteration, extrapolated by taps to the batch"}
    import ctypes
    ci = [ctypes.c_int() for _ in range(3)]
    pkg.lib().hp3d_pm_variant_info(best, *[ctypes.byref(c) for c in ci])
    value = mpix / batched_s
    out = {"value": round(value, 2), "unit": "Mpixel-iters/s",
           "dtype": f"{a.form_store} storage / fp32 accumulation",
           "config": {"workload": f"batched form: {views} views, resolutions over "
                                  "{640x480,1280x960,1600x1200,2400x1800,4000x3000}, 3-8 sources, "
                                  f"11x11 stride {stride}, 8 iterations, one launch in a hipGraph",
                      "views": views, "mpixel_iters": round(mpix, 2), "taps": taps,
                      "render_s": round(render_s, 1)},
           "ms_per_batch": round(batched_s * 1e3, 2),
           "launch_form": "persistent kernel" if persist_s <= steps_s else "graph of step launches",
           "persistent_value": round(mpix / persist_s, 2), "persistent_chunk": a.chunk,
           "step_graph_value": round(mpix / steps_s, 2),
           "plain_serial_same_storage_value": round(mpix / plain_same_s, 2),
           "speedup_vs_plain_serial_same_storage": round(plain_same_s / batched_s, 3),
           "autotune": {"variant": best, "tile": [ci[0].value, ci[1].value],
                        "times_ms": [round(float(t), 3) for t in times]},
           "tflops_algorithmic": round(taps * FLOP_PER_TAP / batched_s / 1e12, 3),
           "cpu_baseline": cpu}
    if plain_s:
        out["plain_serial_value"] = round(mpix / plain_s, 2)
        out["speedup_vs_plain_serial"] = round(plain_s / batched_s, 3)
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    ctx.close()
    return out


def form_batched(a, entry, pkg, scenes):
    out = measure_batched(a, entry, pkg, scenes, views=a.views, stride=a.stride or 1)
    out = dict(form="batched", metric="PatchMatch Mpixel-iters/s", data="synthetic", **out)
    print(json.dumps(out))


def measure_serving(a, entry, pkg, scenes, rate=20.0, duration=120.0, stride=2):
    """The serving workload: Poisson arrivals (rate, duration); each job 1
    reference + 3-8 sources from a pool of 16 views of the batched resolution
    mix, 4-12 iterations, FP16 storage, continuous batching with HBM-budgeted
    admission; the pool's images are resident on the device (added once,
    named by handle).  Latency = completion - submission.  rho = offered
    load / the plain kernel's rate run serially on the pool (same storage);
    the device's busy fraction is the measured utilisation."""
    rng = np.random.default_rng(2024)
    pool_specs = scenes.batch_specs(16, batch_seed=77, iterations=8, stride=stride)
    pool = [scenes.job_from_spec(s, threads=a.cpu_threads) for s in pool_specs]
    n_arr = rng.poisson(rate * duration)
    arrivals = np.sort(rng.uniform(0, duration, n_arr))
    picks = rng.integers(0, len(pool), n_arr)
    iters = rng.integers(4, 13, n_arr)
    offered = sum(pool[p].shape[0] * pool[p].shape[1] * it
                  for p, it in zip(picks, iters)) / 1e6 / duration
    store = pkg.STORE_F16 if a.form_store == "f16" else pkg.STORE_F32
    srv = pkg.Server(hbm_budget=96 << 30, max_jobs=a.max_jobs, radius=5, stride=stride,
                     store=store, variant=a.variant)
    handles = [[srv.image_add(j.ref)] + [srv.image_add(x) for x in j.srcs] for j in pool]

    def submit(p, it, **kw):
        j = pool[p]
        return srv.submit(handles[p][0], handles[p][1:], j.Kinv, j.A, j.b, j.dmin, j.dmax, it,
                          top_k=j.top_k, shape=j.shape, **kw)
    for w in range(2):                                   # warm-up
        srv.wait(submit(w, 4), timeout=120)
    s0 = srv.stats()
    lat, done_mpix, t_first, t_last = [], 0.0, None, None
    pending = []

    def collect(st, p, it):
        nonlocal done_mpix, t_first, t_last
        if st is None or st["state"] != pkg.JOB_DONE:
            return
        lat.append(st["t_done"] - st["t_submit"])
        H, W = pool[p].shape
        done_mpix += W * H * it / 1e6
        t_first = st["t_submit"] if t_first is None else min(t_first, st["t_submit"])
        t_last = st["t_done"] if t_last is None else max(t_last, st["t_done"])

    t0, dropped, max_queue, failed = time.perf_counter(), 0, 0, 0
    for k in range(n_arr):
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        # collect finished jobs so host output buffers do not pile up
        keep = []
        for jid, p, it in pending:
            st = srv.poll(jid)
            if st["state"] in (pkg.JOB_QUEUED, pkg.JOB_RUNNING):
                keep.append((jid, p, it))
            else:
                failed += st["state"] != pkg.JOB_DONE
                collect(st, p, it)
                srv.forget(jid)
        pending = keep
        max_queue = max(max_queue, len(pending))
        if len(pending) > 400:                           # host memory guard when overloaded
            dropped += 1
            continue
        pending.append((submit(int(picks[k]), int(iters[k]), seed=1000 + k, with_cost=False),
                        int(picks[k]), int(iters[k])))
        if k % 200 == 0:
            print(f"serving: {k}/{n_arr} submitted at t={now:.1f}s, in flight {len(pending)}",
                  file=sys.stderr, flush=True)
    for jid, p, it in pending:
        st = srv.wait(jid, timeout=600)
        failed += st is None or st["state"] != pkg.JOB_DONE
        if st is None:                 # still running: its buffers stay with the server
            srv.cancel(jid)
            continue
        collect(st, p, it)
        srv.forget(jid)
    s = srv.stats()
    srv.close()
    wall = t_last - t_first
    # the plain kernel, one job after the other, on the pool at 8 iterations
    ctx = pkg.Context()
    plain_mpix, plain_s = 0.0, 0.0
    for j in pool:
        ref = ctx.upload(j.ref, store)
        srcs = [ctx.upload(x, store) for x in j.srcs]
        d = pkg.make_desc(ref, srcs, j.Kinv, j.A, j.b, j.dmin, j.dmax, radius=5, stride=stride,
                          iterations=8, seed=j.seed)
        outs = _device_outputs(ctx, j.shape)
        ctx.run_view_device(d, *outs)
        t1 = time.perf_counter()
        ctx.run_view_device(d, *outs)
        plain_s += time.perf_counter() - t1
        plain_mpix += j.mpix_iters()
        for q in outs:
            ctx.device_free(q)
        ctx.release(ref)
        for x in srcs:
            ctx.release(x)
    ctx.close()
    lat = np.array(lat) * 1e3
    value = done_mpix / wall
    plain_rate = plain_mpix / plain_s
    steps = s["steps"] - s0["steps"]
    return {"value": round(value, 2), "unit": "Mpixel-iters/s",
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                           "p99": round(float(np.percentile(lat, 99)), 2),
                           "mean": round(float(lat.mean()), 2)},
            "dtype": f"{a.form_store} storage / fp32 accumulation",
            "config": {"workload": f"serving form: Poisson {rate}/s for {duration}s, 1 ref + "
                                   "3-8 src from the batched resolution mix, 4-12 iterations, "
                                   f"11x11 stride {stride}, resident images",
                       "pool_jobs": len(pool), "max_jobs": a.max_jobs},
            "arrivals": int(n_arr), "completed": len(lat), "failed": int(failed),
            "dropped_overload": dropped,
            "offered_mpix_iters_per_s": round(offered, 2),
            "rho": round(offered / plain_rate, 3),
            "stable": bool(offered < plain_rate),
            "utilisation": round((s["busy_ms"] - s0["busy_ms"]) / 1e3 / wall, 3),
            "scheduler_wait_ms_per_step": round((s["wait_ms"] - s0["wait_ms"]) / max(steps, 1), 3),
            "max_in_flight": int(max_queue), "max_active_per_step": s["max_active"],
            "image_uploads": s["image_uploads"], "image_reuses": s["image_reuses"],
            "plain_serial_same_storage_value": round(plain_rate, 2),
            "speedup_vs_plain_serial": round(value / plain_rate, 3),
            "steps": steps}


def form_serving(a, entry, pkg, scenes):
    out = measure_serving(a, entry, pkg, scenes, rate=a.rate, duration=a.duration,
                          stride=a.stride or 2)
    out = dict(form="serving", metric="PatchMatch Mpixel-iters/s; p50/p99 job latency",
               data="synthetic", **out)
    print(json.dumps(out))


def form_geom(a, entry, pkg, scenes):
    """The HBM-bound rows beside the hot path on the 36-view turntable at
    1280x960 (ground-truth depth with 1 % noise, 8 sources per view):
    geometric consistency filter over all views, then fusion into points.
    Host buffers in and out, as the C ABI takes them (PCIe included); the
    kernels' own durations come from rocprofv3."""
    tt = scenes.turntable(threads=a.cpu_threads)
    sfm = pkg.SfM(tt["cameras"], tt["points"])
    vps = sfm.select_source_views(8)
    rng = np.random.default_rng(3)
    base = []
    for v in range(len(tt["cameras"])):
        d = tt["gt_depth"][v] * rng.normal(1.0, 0.01, tt["gt_depth"][v].shape).astype(np.float32)
        d[tt["gt_depth"][v] <= 0] = rng.uniform(0.1, 0.4, int((tt["gt_depth"][v] <= 0).sum()))
        base.append(d.astype(np.float32))
    normal = np.zeros(base[0].shape + (3,), np.float32)
    normal[..., 2] = -1.0
    ctx = pkg.Context()

    def maps():
        return [dict(depth=base[v].copy(), normal=normal, confidence=np.ones_like(base[v]),
                     K=c["K"], R=c["R"], t=c["t"], src=[s[0] for s in vps[v]["sources"]],
                     color=tt["images"][v])
                for v, c in enumerate(tt["cameras"])]

    m = maps()
    ctx.geom_filter(m)
    reps, tg, tf = 3, 0.0, 0.0
    for _ in range(reps):
        m = maps()
        t0 = time.perf_counter()
        kept = ctx.geom_filter(m)
        tg += time.perf_counter() - t0
        t0 = time.perf_counter()
        pos, _, _ = ctx.fuse(m)
        tf += time.perf_counter() - t0
    H, W = base[0].shape
    mpix = len(base) * W * H / 1e6
    out = {"form": "geom", "metric": "Mpixel/s (host buffers, PCIe included)",
           "geom_filter_mpix_per_s": round(mpix / (tg / reps), 1),
           "fuse_mpix_per_s": round(mpix / (tf / reps), 1),
           "retained_fraction": round(float(sum(max(0, k) for k in kept)) / (len(base) * W * H), 4),
           "points": int(len(pos)),
           "config": {"workload": "36-view turntable 1280x960, 8 sources per view, GT depth "
                                  "+1 % noise, random background depth"}}
    print(json.dumps(out))
    ctx.close()

if __name__ == "__main__":
    main()
