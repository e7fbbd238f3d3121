#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the Barnes-Hut step (sort + refresh + walk)
on one MI355X, headline workload = collision of two 2,097,152-particle disks
(N = 4,194,304), theta 0.5, eps 0.005, bucket 32, one topology rebuild per
step.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

A step is gs_step(sim, 1, 1): bbox + Morton encode, 4 radix passes (last one
gathers the particle records), 11 BFS build levels, FP64 refresh, walk with
halos + kick-drift + FP64 reductions.  Inputs are resident in HBM before the
timed region; every step is bracketed by a stream synchronisation so the
per-step latencies give p50/p99.  Stage durations come from HIP events the
library records on its own stream around each stage.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "galaxy-sim_amd"))

METRIC = "particle-steps/sec (sort+refresh+walk) at N=4M θ=0.5; p50/p99 step latency"
WALK_FLOP_PER_INTERACTION = 21  # 3 sub, 3 mul + 3 add, sqrt, mul, mul, div, 3 mul + 3 add, mul + sub
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def cpu_baseline(data, n, theta, eps, bucket, sample):
    """The oracle (C restatement of mortonTree.js + FP32 transliteration of the
    WGSL walk) on this host: sortAndBuild + FP64 refresh on all N particles,
    single-threaded as the reference's worker; the walk over all host cores on
    a seeded sample of particles, extrapolated to N."""
    from oracle import oracle
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ref = oracle.sort_and_build(data.reshape(-1), n, bucket=bucket)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.refresh_f64(ref["sorted"], ref["nodes"])
    t_refresh = time.perf_counter() - t0
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(n, sample, replace=False)).astype(np.uint32)
    op = oracle.pack_params(n, dt=0.005, G=1.0, epsilon=eps, theta=theta)
    t0 = time.perf_counter()
    _, inter = oracle.walk_subset(ref["sorted"], ref["nodes"], op, idx, nthreads=cores)
    t_walk = (time.perf_counter() - t0) * n / sample
    step = t_build + t_refresh + t_walk
    return {
        "value": n / step, "unit": "particle-steps/s", "cores": cores, "kind": "port",
        "sample": (f"sortAndBuild + FP64 refresh of all {n} particles on 1 thread "
                   f"({t_build:.2f} s + {t_refresh:.2f} s); walk of {sample} seeded-sampled "
                   f"particles on {cores} threads extrapolated to N ({t_walk:.2f} s)"),
        "step_s": step, "interactions_per_particle": float(inter.mean()),
    }


def copy_probe(mib):
    """Device-to-device copy of a known size (hipMemcpy), the calibration
    kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--exact", action="store_true", help="IEEE sqrt/div force path")
    ap.add_argument("--lib", default=None, help="path of the libgalaxy_hip.so to load")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.lib:
        os.environ["GALAXY_HIP_LIB"] = args.lib
    import galaxy_hip
    from galaxy_hip import ic

    n, theta, eps, bucket = args.n, 0.5, 0.005, 32
    data = ic.make("collision_4m", n=n)
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, timing=True, exact_force=args.exact)
    sim.update_params(dt=0.005, G=1.0, epsilon=eps, theta=theta)
    for _ in range(args.warmup):
        sim.step(1, sort_every=1)
    sim.sync()
    sim.stage_times(reset=True)

    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        sim.step(1, sort_every=1)
        sim.sync()
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    stages = sim.stage_times(reset=True)
    info = sim.tree_info()
    inter = sim.count_interactions()
    sim.destroy()

    lat_ms = np.array(lat) * 1e3
    walk_ms, walk_cnt = stages["walk"]
    walk_avg_ms = walk_ms / max(walk_cnt, 1)
    nu = info["num_nodes"] / n
    achieved_tflops = inter * WALK_FLOP_PER_INTERACTION / (walk_avg_ms * 1e-3) / 1e12
    step_bytes = (236 + 160 * nu) * n
    ms_per_step = total / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": n * args.steps / total,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-disk collision, 2 x 2,097,152 particles)",
        "config": {"workload": "collision_4m: two exponential disks + Hernquist bulges, "
                               "N=4,194,304, theta=0.5, eps=0.005, bucket=32, dt=0.005, "
                               "rebuild every step",
                   "n": n, "theta": theta, "epsilon": eps, "bucket": bucket,
                   "sort_every": 1, "force_path": "exact" if args.exact else "rsqrt"},
        "roofline": {
            "bound": "valu", "kernel": "walk_kernel",
            "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS, "traffic": None,
            "interactions_per_step": inter, "flop_per_interaction": WALK_FLOP_PER_INTERACTION,
            "walk_avg_ms": walk_avg_ms,
        },
        "step_hbm": {
            "bytes_per_particle": 236 + 160 * nu, "nu": nu,
            "achieved_gbs": step_bytes / (ms_per_step * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "stages_ms": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()},
        "interactions_per_s": inter / (ms_per_step * 1e-3),
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        cb = cpu_baseline(data, n, theta, eps, bucket, min(args.cpu_sample, n))
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
