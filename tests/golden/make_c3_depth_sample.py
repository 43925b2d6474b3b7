"""Writes tests/golden/c3_fp64_depth_sample.npz: the oracle's full run (every
cost in FP64, the reference schedule) on the 3200x2400 6-source view (11x11
window at stride 2, 8 iterations, seed 42), sampled at 20000 seeded pixels,
with checksums of the rendered inputs so a changed renderer is detected.
The run takes minutes of CPU, too long for a test; the GPU test compares
its own full-size solves with this sample.

    python tests/golden/make_c3_depth_sample.py [threads]
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

N_SAMPLE = 20000


def inputs_digest(job):
    h = hashlib.sha256()
    for im in [job.ref] + list(job.srcs):
        h.update(np.ascontiguousarray(im).tobytes())
    return h.hexdigest()


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    from importlib import import_module
    entry.load_package()
    scenes = import_module(entry.PKG_NAME + ".scenes")
    orc = entry.load_oracle()
    job = scenes.c3(threads=threads)
    H, W = job.shape
    v = orc.View(job.ref, job.srcs, job.Kinv, job.A, job.b, job.dmin, job.dmax,
                 radius=job.radius, stride=job.stride, top_k=job.top_k)
    t0 = time.time()
    d64, n64, c64 = v.patchmatch_fp32(job.iterations, job.seed, threads=threads,
                                      cost_mode=orc.COST_FP64)
    print(f"oracle FP64-cost run {W}x{H}: {time.time() - t0:.0f} s on {threads} threads")
    rng = np.random.default_rng(20261203)
    idx = rng.choice(H * W, N_SAMPLE, replace=False).astype(np.int64)
    gt = job.gt_depth.reshape(-1)[idx]
    od = d64.reshape(-1)[idx]
    acc = float((np.abs(od - gt) <= 0.01 * np.abs(gt)).mean())
    np.savez_compressed(os.path.join(HERE, "c3_fp64_depth_sample.npz"), index=idx,
                        depth=od.astype(np.float32), gt_within_1pct=np.float64(acc),
                        inputs_sha256=np.array(inputs_digest(job)),
                        shape=np.array([H, W]), iterations=np.int64(job.iterations),
                        seed=np.int64(job.seed))
    print(f"wrote {N_SAMPLE} pixels; oracle within 1 % of ground truth on {acc:.4f}")


if __name__ == "__main__":
    main()
