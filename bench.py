#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the Barnes-Hut step (sort + refresh + walk)
on one MI355X, headline workload = collision of two 2,097,152-particle disks
(N = 4,194,304), theta 0.5, eps 0.005, bucket 32, one topology rebuild per
step.  Prints ONE JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                  [--no-configs] [--no-batch] [--no-serve]
                  [--only headline|configs|batch|serve]

A step is gs_step(sim, 1, 1): bbox + Morton encode, 4 radix passes (last one
gathers the particle records), 11 BFS build levels, FP64 refresh, walk with
halos + kick-drift + FP64 reductions.  Inputs are resident in HBM before the
timed region.  Throughput (value, ms_per_step) is K steps launched back to
back with a synchronisation at each end; K more steps, each bracketed by a
stream synchronisation, give the p50/p99 step latency.  Each step replays two captured graphs
(rebuild + refresh, then the walk) with HIP events the library records on its
own stream between them, which time the walk kernel over the timed region;
a few more steps with direct launches give the per-stage breakdown.

Beside the headline the line carries, per plain-form workload, the GPU rate
with p50/p99, the reference step timed on the host cores through the oracle
(cpu_baseline / "cpu": one warm-up step, then full steps of every particle),
and the FP64 direct-sum force error on a seeded sample:
  configs.cpu_disk_16k  N=16,384 disk, theta 0.7, eps 0.02, bucket 16;
  configs.disk_500k     N=500,000 disk with 30 % gas, Hernquist halo (M 5,
                        a 3) and drag 0.25, theta 0.7, eps 0.01, 100 steps;
  configs.collision_4m  the headline.
and two more objects (skipped with --no-batch/--no-serve):
  batched  64 galaxies of N log-uniform in [4096, 262144] stepped as one slot
           pool (one graph replay per step), walk workgroup / bucket / digit
           width from the autotuner, against the same galaxies stepped with
           per-galaxy launches;
  served   a Poisson job stream (20 jobs/s, N log-uniform in [8192, 2^20],
           200 steps each) served by continuous batching in real time, with
           step p50/p99, queue wait, HBM budget use, the one-job-at-a-time
           throughput and the FP64 force error of three of its jobs.
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


def _halo_params(c):
    """Halo / drag parameters of a workload: the 500,000-particle disk carries
    a Hernquist halo (M 5, a 3) at the origin and gas drag 0.25."""
    if "halo_mass" not in c:
        return {}, ()
    return (dict(haloMass=c["halo_mass"], haloA=c["halo_a"], gasDrag=c["gas_drag"]),
            [(0.0, 0.0, 0.0)])


def cpu_reference_steps(data, c, steps, warmup=1):
    """The reference step on this host's cores through the oracle (C restatement
    of mortonTree.js + FP32 transliteration of the WGSL walk): sortAndBuild and
    the FP64 refresh on one thread, as the reference's worker runs them, then
    the walk with halos, drag and kick-drift over every particle on all the
    cores this process may use.  Each step starts from the previous one's
    particles.  Returns per-step wall times (s) of the timed steps."""
    from oracle import oracle
    n = data.shape[0]
    extra, halo = _halo_params(c)
    op = oracle.pack_params(n, dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"],
                            halo_mass=extra.get("haloMass", 0.0), halo_a=extra.get("haloA", 1.0),
                            gas_drag=extra.get("gasDrag", 0.0), halo_centers=halo)
    threads = oracle.default_threads()
    cur = data.reshape(-1).copy()
    lat, parts = [], {"build": 0.0, "refresh": 0.0, "walk": 0.0}
    for k in range(warmup + steps):
        t0 = time.perf_counter()
        ref = oracle.sort_and_build(cur, n, bucket=c["bucket"])
        t1 = time.perf_counter()
        nodes = oracle.refresh_f64(ref["sorted"], ref["nodes"])
        t2 = time.perf_counter()
        cur, _, _, _, _ = oracle.walk(ref["sorted"], nodes, op, nthreads=threads)
        t3 = time.perf_counter()
        if k >= warmup:
            lat.append(t3 - t0)
            parts["build"] += t1 - t0
            parts["refresh"] += t2 - t1
            parts["walk"] += t3 - t2
    lat = np.array(lat)
    return {
        "value": n * steps / lat.sum(), "unit": "particle-steps/s", "cores": threads,
        "kind": "port", "steps": steps, "warmup": warmup,
        "step_s": float(lat.mean()), "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "stages_s": {k: v / steps for k, v in parts.items()},
        "sample": (f"{steps} full steps of all {n} particles after {warmup} warm-up step(s) "
                   f"from the initial particles: "
                   f"sortAndBuild + FP64 refresh on 1 thread, walk + kick-drift on {threads} "
                   f"threads"),
    }


def force_error(acc, parts, eps, sample, seed):
    """Relative error |a - a_ds| / |a_ds| of the GPU forces against the FP64
    direct sum (oracle.direct_sum_f64) on a seeded sample of particles (all of
    them when sample is None or at least N)."""
    from oracle import oracle
    n = parts.shape[0]
    if sample is None or sample >= n:
        idx = np.arange(n, dtype=np.uint32)
    else:
        idx = np.sort(np.random.default_rng(seed).choice(n, sample, replace=False)).astype(np.uint32)
    ds = oracle.direct_sum_f64(parts.reshape(-1), 1.0, float(np.float32(eps * eps)), idx)
    a = acc[idx, :3].astype(np.float64)
    err = np.linalg.norm(a - ds, axis=1) / np.linalg.norm(ds, axis=1)
    return {"particles": int(idx.size), "median": float(np.median(err)),
            "p99": float(np.quantile(err, 0.99)), "max": float(err.max()),
            "rms": float(np.sqrt(np.mean(err * err)))}


def gpu_forces(data, bucket, theta, eps, exact=False):
    """Default-path accelerations of one force evaluation (sort, refresh, walk)
    and the sorted particles they belong to."""
    import galaxy_hip
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact)
    try:
        sim.update_params(dt=0.005, G=1.0, epsilon=eps, theta=theta)
        sim.sort()
        sim.refresh()
        sim.force()
        acc, _ = sim.accel()
        return acc, sim.particles()
    finally:
        sim.destroy()


def bench_config(name, steps, cpu_steps, fe_sample, args):
    """One plain-form workload: GPU steps (each step one sort + refresh + walk
    with its own synchronisation, for p50/p99), the reference step on the host
    cores, and the FP64 direct-sum force error."""
    import galaxy_hip
    from galaxy_hip import ic
    c = ic.CONFIGS[name]
    data = ic.make(name)
    n = data.shape[0]
    extra, halo = _halo_params(c)
    log(f"{name}: N={n}, {steps} GPU steps")
    sim = galaxy_hip.NBodyGPU(data, bucket=c["bucket"], exact_force=args.exact)
    try:
        sim.update_params(dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"],
                          haloCenters=halo, **extra)
        # every pass covers the same steps: the initial particles, the
        # untimed warm-up steps, then the measured ones
        def restart():
            sim.upload(data)
            sim.step(args.warmup, sort_every=1)
            sim.sync()

        restart()
        t0 = time.perf_counter()
        sim.step(steps, sort_every=1)  # back to back
        sim.sync()
        total = time.perf_counter() - t0
        restart()
        lat = _timed_steps(sim, steps)  # each step synchronised: latency
        restart()
        counts = []
        for _ in range(steps):
            sim.step(1, sort_every=1)
            counts.append(sim.count_interactions())
        inter = float(np.mean(counts))
    finally:
        sim.destroy()
    lat_ms = lat * 1e3
    gpu = {"value": n * steps / total, "unit": "particle-steps/s", "steps": steps,
           "warmup": args.warmup, "ms_per_step": total / steps * 1e3,
           "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
           "interactions_per_step": inter}
    log(f"{name}: GPU {gpu['value'] / 1e6:.1f} M particle-steps/s, p50 {gpu['p50_ms']:.3f} ms "
        f"p99 {gpu['p99_ms']:.3f} ms")
    out = {"n": n, "theta": c["theta"], "epsilon": c["epsilon"], "bucket": c["bucket"],
           "dt": c["dt"], **({"halo": extra} if extra else {}), "gpu": gpu}
    acc, parts = gpu_forces(data, c["bucket"], c["theta"], c["epsilon"], args.exact)
    out["force_error_fp64"] = force_error(acc, parts, c["epsilon"], fe_sample,
                                          0x6A1A5EED00000100 + n)
    log(f"{name}: FP64 direct-sum error median {out['force_error_fp64']['median']:.3e} "
        f"p99 {out['force_error_fp64']['p99']:.3e}")
    if cpu_steps and not args.no_cpu_baseline:
        cpu = cpu_reference_steps(data, c, cpu_steps)
        out["cpu"] = cpu
        out["speedup_vs_cpu"] = gpu["value"] / cpu["value"]
        log(f"{name}: CPU reference {cpu['value'] / 1e6:.3f} M particle-steps/s on "
            f"{cpu['cores']} threads ({cpu['step_s'] * 1e3:.1f} ms/step)")
    return out


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def _timed_steps(obj, steps):
    """Per-step wall latencies (s) of obj.step(1, sort_every=1) + sync."""
    lat = []
    for _ in range(steps):
        t0 = time.perf_counter()
        obj.step(1, sort_every=1)
        obj.sync()
        lat.append(time.perf_counter() - t0)
    return np.array(lat)


def bench_batch(args):
    """64 galaxies of uneven N stepped as one slot pool (one captured graph
    per step, every stage one launch over all galaxies), against the plain
    per-galaxy form: each galaxy on its own handle with its own launches,
    one galaxy after another, and (second plain form) all 64 handles queued
    on their own streams before one synchronisation."""
    import galaxy_hip
    from galaxy_hip import autotune, ic
    c = ic.CONFIGS["batch64"]
    gals = ic.batch64(c["count"], c["lo"], c["hi"])
    datas = [g for _, g in gals]
    sizes = [d.shape[0] for d in datas]
    total_n = int(sum(sizes))
    params = dict(dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"])
    log(f"batch64: {len(datas)} galaxies, N total {total_n}, min {min(sizes)}, max {max(sizes)}")
    t0 = time.perf_counter()
    if args.tuned:
        wg, bucket, radix = (int(x) for x in args.tuned.split(","))
        tuned = {"walk_wg": wg, "bucket": bucket, "radix_bits": radix, "trials": []}
    else:
        tuned = autotune.tune(datas, params, log=log)
    t_tune = time.perf_counter() - t0
    wg, bucket, radix = tuned["walk_wg"], tuned["bucket"], tuned["radix_bits"]
    log(f"autotune: wg={wg} bucket={bucket} radix={radix} ({len(tuned['trials'])} trials, "
        f"{t_tune:.1f} s)")

    pool = galaxy_hip.GalaxyPool(sizes, buckets=[bucket] * len(sizes), walk_wg=wg,
                                 radix_bits=radix)
    try:
        # every pass, the plain forms' included, covers the same steps: the
        # initial particles, the untimed warm-up steps, then the measured ones
        def restart():
            for k, d in enumerate(datas):
                pool.upload(k, d, **params)
            pool.step(args.warmup, sort_every=1)
            pool.sync()

        restart()
        t0 = time.perf_counter()
        pool.step(args.batch_steps, sort_every=1)  # back to back
        pool.sync()
        t_pool = time.perf_counter() - t0
        restart()
        lat = _timed_steps(pool, args.batch_steps)  # each step synchronised: latency
        restart()
        counts = []
        for _ in range(args.batch_steps):
            pool.step(1, sort_every=1)
            counts.append(pool.count_interactions())
        inter = float(np.mean(counts))
    finally:
        pool.destroy()
    pooled = total_n * args.batch_steps / t_pool
    log(f"pooled: {t_pool / args.batch_steps * 1e3:.3f} ms/step, {pooled / 1e6:.1f} M particle-steps/s")

    if args.no_plain:
        return {"n_total": total_n, "value": pooled, "unit": "particle-steps/s",
                "ms_per_step": t_pool / args.batch_steps * 1e3}
    # plain form 1: one galaxy after another, each on its own handle
    ksteps = args.batch_steps
    t_seq = 0.0
    for d in datas:
        sim = galaxy_hip.NBodyGPU(d, bucket=bucket, walk_wg=wg, radix_bits=radix)
        try:
            sim.update_params(**params)
            sim.step(args.warmup, sort_every=1)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(ksteps, sort_every=1)
            sim.sync()
            t_seq += (time.perf_counter() - t0) / ksteps
        finally:
            sim.destroy()
    plain_seq = total_n / t_seq
    log(f"plain sequential: {t_seq * 1e3:.3f} ms per round of 64, {plain_seq / 1e6:.1f} M/s")

    # plain form 2: 64 handles, each step queued on its own stream
    sims = []
    try:
        for d in datas:
            sim = galaxy_hip.NBodyGPU(d, bucket=bucket, walk_wg=wg, radix_bits=radix)
            sim.update_params(**params)
            sims.append(sim)
        for sim in sims:
            sim.step(args.warmup, sort_every=1)
        for sim in sims:
            sim.sync()
        t0 = time.perf_counter()
        for _ in range(ksteps):
            for sim in sims:
                sim.step(1, sort_every=1)
            for sim in sims:
                sim.sync()
        t_conc = (time.perf_counter() - t0) / ksteps
    finally:
        for sim in sims:
            sim.destroy()
    plain_conc = total_n / t_conc
    log(f"plain 64 streams: {t_conc * 1e3:.3f} ms per round, {plain_conc / 1e6:.1f} M/s")
    lat_ms = lat * 1e3
    best_plain = max(plain_seq, plain_conc)
    return {
        "workload": "batch64: 64 galaxies, N log-uniform in [4096, 262144] (seeded), "
                    "disk / Plummer a=1 50/50, theta=0.6, eps=0.01, dt=0.005, rebuild every step",
        "n_total": total_n, "galaxies": len(sizes),
        "value": pooled, "unit": "particle-steps/s",
        "ms_per_step": t_pool / args.batch_steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "steps": args.batch_steps,
        "interactions_per_s": inter / (t_pool / args.batch_steps),
        "plain_sequential": plain_seq, "plain_streams": plain_conc,
        "speedup_vs_plain": pooled / best_plain,
        "speedup_vs_plain_sequential": pooled / plain_seq,
        "autotune": {"walk_wg": wg, "bucket": bucket, "radix_bits": radix,
                     "trials": len(tuned["trials"]), "seconds": t_tune,
                     "space": {k: list(v) for k, v in autotune.DEFAULT_SPACE.items()}},
    }


def bench_serve(args, tuned):
    """Poisson job stream served by continuous batching (GalaxyServer) in
    real time, against the plain form: each job alone, one after another."""
    import concurrent.futures as cf

    import galaxy_hip
    from galaxy_hip import ic, serve
    c = dict(ic.CONFIGS["serve"])
    if args.serve_budget:
        c["hbm_budget"] = args.serve_budget
    jobs = ic.serve_jobs(args.serve_jobs, rate=c["rate"], lo=c["lo"], hi=c["hi"],
                         steps=args.serve_steps)
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(8) as ex:
        records = list(ex.map(ic.job_records, jobs))
    log(f"serve: {len(jobs)} jobs generated in {time.perf_counter() - t0:.1f} s, "
        f"N total {sum(j['n'] for j in jobs)}, last arrival {jobs[-1]['arrival']:.2f} s")
    params = dict(dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"])
    wg = tuned.get("walk_wg", 0)
    radix = tuned.get("radix_bits", 0)
    bucket = tuned.get("bucket", 16)
    t0 = time.perf_counter()
    with serve.GalaxyServer(hbm_budget=c["hbm_budget"], max_jobs=args.serve_slots,
                            resident_limit=args.serve_resident, bucket=bucket, walk_wg=wg,
                            radix_bits=radix, keep_results=False) as srv:
        log(f"serve: arena of {srv.pool.arena_info()['device_bytes'] / 1e9:.1f} GB for the "
            f"{c['hbm_budget'] / 1e9:.0f} GB budget, {args.serve_slots} slots "
            f"({time.perf_counter() - t0:.2f} s to allocate)")
        js = [serve.Job(j["id"], r, params, j["steps"], j["arrival"], bucket=bucket)
              for j, r in zip(jobs, records)]
        summ = srv.run(js, realtime=True, log=log)
    log(f"served: {summ['throughput'] / 1e6:.1f} M particle-steps/s, job-step p50 "
        f"{summ['step_p50_ms']:.2f} ms p99 {summ['step_p99_ms']:.2f} ms (server step p50 "
        f"{summ['server_step_p50_ms']:.2f} p99 {summ['server_step_p99_ms']:.2f}); queue wait "
        f"p50 {summ['queue_wait_p50_s']:.2f} s p99 {summ['queue_wait_p99_s']:.2f} s; budget "
        f"use peak {summ['budget_used_peak_bytes'] / 1e9:.2f} GB; held by "
        f"{summ['jobs_held_by']}")
    # plain: every job alone on its own handle, timed over a few steps
    t_plain = 0.0
    work = 0
    for j, r in zip(jobs, records):
        sim = galaxy_hip.NBodyGPU(r, bucket=bucket, walk_wg=wg, radix_bits=radix)
        try:
            sim.update_params(**params)
            sim.step(1, sort_every=1)
            sim.sync()
            t0 = time.perf_counter()
            sim.step(args.plain_steps, sort_every=1)
            sim.sync()
            t_plain += (time.perf_counter() - t0) / args.plain_steps * j["steps"]
            work += j["n"] * j["steps"]
        finally:
            sim.destroy()
    plain = work / t_plain
    log(f"served plain (one job at a time): {plain / 1e6:.1f} M particle-steps/s")
    # FP64 direct-sum force error of the served force path on three jobs of the
    # stream (the smallest, the median and the largest N) at their initial state
    by_n = sorted(range(len(jobs)), key=lambda k: jobs[k]["n"])
    fe = []
    for k in (by_n[0], by_n[len(by_n) // 2], by_n[-1]):
        acc, parts = gpu_forces(records[k], bucket, params["theta"], params["epsilon"])
        e = force_error(acc, parts, params["epsilon"], 4096, 0x6A1A5EED00000200 + k)
        fe.append(dict(job=jobs[k]["id"], n=jobs[k]["n"], preset=jobs[k]["preset"], **e))
        log(f"served job {jobs[k]['id']} ({jobs[k]['preset']}, N={jobs[k]['n']}): FP64 "
            f"direct-sum error median {e['median']:.3e} p99 {e['p99']:.3e}")
    summ.update({
        "workload": f"serve: {len(jobs)} jobs, Poisson arrivals 20 jobs/s, N log-uniform in "
                    f"[8192, 1048576], {args.serve_steps} steps per job, disk / Plummer / "
                    f"collision uniform, theta=0.5, eps=0.01, dt=0.005, "
                    f"{c['hbm_budget'] / 1e9:.0f} GB HBM budget (arena), {args.serve_slots} "
                    f"slots, bucket {bucket}",
        "admission_note": "a job waits for a free slot, for the resident-particle cap "
                          "(resident_limit, the latency bound) and for a free arena range "
                          "(the HBM budget); jobs_held_by counts the jobs each held at least once",
        "latency_note": "step_p50/p99 count each server step once per job it advanced; "
                        "server_step_p50/p99 count each server step once",
        "value": summ["throughput"], "unit": "particle-steps/s",
        "resident_limit": args.serve_resident,
        "plain_one_at_a_time": plain, "speedup_vs_plain": summ["throughput"] / plain,
        "saturated_speedup_vs_plain": summ["saturated_throughput"] / plain,
        "p99_over_p50": summ["step_p99_ms"] / max(summ["step_p50_ms"], 1e-9),
        "force_error_fp64": fe,
    })
    return summ


COPY_PROBE_MIB = 256


def probe_walk(args):
    """Child-process workload of the counter passes: a 256 MiB device copy
    (the calibration kernel of known size), then the headline galaxy sorted,
    refreshed and stepped twice."""
    copy_probe(COPY_PROBE_MIB, quiet=True)
    import galaxy_hip
    from galaxy_hip import ic
    sim = galaxy_hip.NBodyGPU(ic.make("collision_4m", n=args.n), bucket=32)
    sim.update_params(dt=0.005, G=1.0, epsilon=0.005, theta=0.5)
    sim.step(2, sort_every=1)
    sim.sync()
    sim.destroy()


def _pmc_pass(counter, args, workdir):
    """One rocprofv3 counter pass over probe_walk (counter: one name, or
    several separated by spaces); per counter, the mean value per walk
    dispatch and the largest copy dispatch (KiB for FETCH_SIZE / WRITE_SIZE).
    A single counter returns {"walk", "copy"}; several return one such dict
    per counter."""
    names = counter.split()
    import shutil
    import signal
    import sqlite3
    import subprocess
    exe = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    outdir = os.path.join(workdir, "_".join(names).lower())
    cmd = [exe, "--pmc", *names, "--kernel-include-regex", "walk_kernel|copyBuffer",
           "-d", outdir, "-o", "pmc", "--", sys.executable, os.path.abspath(__file__),
           "--probe-walk", "--n", str(args.n)]
    env = dict(os.environ, TMPDIR="/tmp")
    p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                         stderr=subprocess.DEVNULL, start_new_session=True)
    try:
        rc = p.wait(timeout=args.pmc_timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        raise RuntimeError(f"{counter} pass timed out")
    dbs = [os.path.join(r, f) for r, _, fs in os.walk(outdir) for f in fs if f.endswith(".db")]
    if rc != 0 and not dbs:
        raise RuntimeError(f"{counter} pass exited {rc}")
    out = {}
    for c in names:
        vals = {}
        for db in dbs:
            con = sqlite3.connect(db)
            for name, value in con.execute(
                    "select kernel_name, value from counters_collection where counter_name = ?",
                    (c,)):
                key = "walk" if "walk_kernel" in name else "copy" if "copyBuffer" in name else None
                if key:
                    vals.setdefault(key, []).append(float(value))
            con.close()
        if "walk" not in vals or "copy" not in vals:
            raise RuntimeError(f"{c} pass recorded no walk/copy dispatches")
        # the calibration copy is the largest copy dispatch (host uploads are
        # smaller blit copies of the same kernel)
        out[c] = {"walk": float(np.mean(vals["walk"])), "copy": float(max(vals["copy"]))}
    return out[names[0]] if len(names) == 1 else out


def pmc_traffic(args):
    """HBM bytes per walk launch from FETCH_SIZE + WRITE_SIZE, each scaled by
    what the counter reads for a device copy of COPY_PROBE_MIB MiB (which
    fetches and writes exactly that many bytes)."""
    import tempfile
    work = tempfile.mkdtemp(prefix="gs_pmc_", dir="/tmp")
    known_kib = COPY_PROBE_MIB * 1024.0
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        v = _pmc_pass(counter, args, work)
        scale = known_kib / v["copy"]
        res[counter] = {"walk_kib": v["walk"], "copy_kib": v["copy"], "scale": scale,
                        "walk_bytes": v["walk"] * scale * 1024.0}
    # L2 (TCC) hit rate of the walk: hits / (hits + misses) over all channels
    try:
        v = _pmc_pass("TCC_HIT_sum TCC_MISS_sum", args, work)
        hit, miss = v["TCC_HIT_sum"]["walk"], v["TCC_MISS_sum"]["walk"]
        res["L2"] = {"hits": hit, "misses": miss, "hit_rate": hit / max(hit + miss, 1.0)}
    except Exception as e:  # the traffic figures stand without it
        log(f"L2 counter pass failed ({type(e).__name__}: {e}); l2_hit_rate = null")
    return res


def copy_probe(mib, quiet=False):
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
    if not quiet:
        print(json.dumps({"copy_probe_bytes": nbytes, "copies": 3}))


def bench_configs(args):
    """The two other plain-form workloads: the 16,384-particle disk (theta 0.7,
    eps 0.02, bucket 16) and the 500,000-particle disk with gas, halo and drag
    (theta 0.7, eps 0.01, bucket 16, 100 steps)."""
    return {
        "cpu_disk_16k": bench_config("cpu_disk_16k", args.disk16k_steps, args.disk16k_cpu_steps,
                                     None, args),
        "disk_500k": bench_config("disk_500k", args.disk500k_steps, args.cpu_steps, 4096, args),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4194304)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="timed reference steps on the host cores (after one warm-up)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--disk16k-steps", type=int, default=200)
    ap.add_argument("--disk16k-cpu-steps", type=int, default=20)
    ap.add_argument("--disk500k-steps", type=int, default=100)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 16,384- and 500,000-particle plain workloads")
    ap.add_argument("--exact", action="store_true", help="IEEE sqrt/div force path")
    ap.add_argument("--lib", default=None, help="path of the libgalaxy_hip.so to load")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration)")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter passes (walk HBM traffic, L2 hit rate)")
    ap.add_argument("--pmc-timeout", type=int, default=120)
    ap.add_argument("--probe-walk", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-galaxy batched run")
    ap.add_argument("--no-serve", action="store_true", help="skip the served job stream")
    ap.add_argument("--only", choices=("headline", "configs", "batch", "serve"), default=None)
    ap.add_argument("--tuned", default=None, metavar="WG,BUCKET,RADIX",
                    help="skip the autotuner and use these batched-run parameters")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-plain", action="store_true",
                    help="batched run: time the pooled steps only (profiling)")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--plain-steps", type=int, default=5,
                    help="steps timed per job for the served plain form (scaled to the job's steps)")
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-resident", type=int, default=8 << 20,
                    help="resident particle limit of the server")
    ap.add_argument("--serve-slots", type=int, default=64,
                    help="arena slots of the server (jobs resident at once)")
    ap.add_argument("--serve-budget", type=float, default=None,
                    help="HBM budget of the server in bytes (default: the workload's 200 GB)")
    args = ap.parse_args()
    if args.lib:
        os.environ["GALAXY_HIP_LIB"] = os.path.abspath(args.lib)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.probe_walk:
        probe_walk(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.only == "batch":
        print(json.dumps(bench_batch(args)))
        return
    if args.only == "configs":
        print(json.dumps(bench_configs(args)))
        return
    if args.only == "serve":
        tuned = {}
        if args.tuned:
            wg, bucket, radix = (int(x) for x in args.tuned.split(","))
            tuned = {"walk_wg": wg, "bucket": bucket, "radix_bits": radix}
        print(json.dumps(bench_serve(args, tuned)))
        return
    traffic = None
    if not (args.no_pmc or args.only == "headline"):
        try:
            log("walk HBM traffic and L2 hit rate: three rocprofv3 counter passes")
            traffic = pmc_traffic(args)
        except Exception as e:
            log(f"counter passes failed ({type(e).__name__}: {e}); traffic = null")
    out = bench_headline(args)
    if traffic:
        fetch = traffic["FETCH_SIZE"]["walk_bytes"]
        write = traffic["WRITE_SIZE"]["walk_bytes"]
        out["roofline"]["traffic"] = fetch + write
        out["roofline"]["traffic_detail"] = {
            "fetch_bytes": fetch, "write_bytes": write,
            "fetch_scale": traffic["FETCH_SIZE"]["scale"],
            "write_scale": traffic["WRITE_SIZE"]["scale"],
            "algorithmic_bytes": out["roofline"]["walk_algorithmic_bytes"],
            "hbm_gbs": (fetch + write) / (out["roofline"]["walk_avg_ms"] * 1e-3) / 1e9,
            "source": f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per walk dispatch, scaled by a "
                      f"{COPY_PROBE_MIB} MiB device copy"}
        gbs = out["roofline"]["traffic_detail"]["hbm_gbs"]
        out["roofline"]["hbm"]["measured_gbs"] = gbs
        out["roofline"]["hbm"]["measured_frac"] = gbs / HBM_PEAK_GBS
        out["roofline"]["l2_hit_rate"] = traffic["L2"]["hit_rate"] if "L2" in traffic else None
        if "L2" in traffic:
            out["roofline"]["l2_detail"] = {
                "hits": traffic["L2"]["hits"], "misses": traffic["L2"]["misses"],
                "source": "rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum per walk dispatch"}
    if not (args.no_configs or args.only == "headline"):
        try:
            out["configs"] = bench_configs(args)
        except Exception as e:  # the headline line is still printed
            out["configs"] = {"error": f"{type(e).__name__}: {e}"}
        out["configs"]["collision_4m"] = {
            "n": out["config"]["n"], "gpu": {k: out[k] for k in ("value", "unit", "steps", "warmup",
                                                          "ms_per_step", "p50_ms", "p99_ms")},
            "cpu": out["cpu_baseline"], "force_error_fp64": out["force_error_fp64"],
            "note": "the headline above"}
    tuned = {}
    if not (args.no_batch or args.only == "headline"):
        try:
            out["batched"] = bench_batch(args)
            tuned = out["batched"]["autotune"]
        except Exception as e:  # the headline line is still printed
            out["batched"] = {"error": f"{type(e).__name__}: {e}"}
    if not (args.no_serve or args.only == "headline"):
        try:
            out["served"] = bench_serve(args, tuned)
        except Exception as e:
            out["served"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(out))


def bench_headline(args):
    import galaxy_hip
    from galaxy_hip import ic

    n, theta, eps, bucket = args.n, 0.5, 0.005, 32
    log(f"headline: collision_4m N={n}")
    data = ic.make("collision_4m", n=n)
    # The collision evolves (the interactions per step grow by ~10 % over the
    # first 40 steps), so every measurement below covers the same steps: the
    # initial particles, W untimed steps, then the K measured steps.
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, timing=True, exact_force=args.exact)
    sim.update_params(dt=0.005, G=1.0, epsilon=eps, theta=theta)

    def restart(s):
        s.upload(data)
        for _ in range(args.warmup):
            s.step(1, sort_every=1)
        s.sync()
        s.stage_times(reset=True)

    restart(sim)
    # throughput: K steps launched back to back (the simulation loop), one
    # synchronisation at each end
    t_start = time.perf_counter()
    sim.step(args.steps, sort_every=1)
    sim.sync()
    total = time.perf_counter() - t_start
    stages = sim.stage_times(reset=True)
    # latency: the same K steps, each followed by a synchronisation (p50/p99)
    restart(sim)
    lat = _timed_steps(sim, args.steps)
    # interactions of the same K steps, each counted right after its step
    # (that step's tree against the particles one drift later)
    restart(sim)
    counts = []
    for _ in range(args.steps):
        sim.step(1, sort_every=1)
        counts.append(sim.count_interactions())
    inter = float(np.mean(counts))
    info = sim.tree_info()
    sim.destroy()
    # stage breakdown: the timed steps replay two graphs (rebuild + refresh,
    # walk) with HIP events between them; the same K steps with direct
    # launches time every stage
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, timing=True, exact_force=args.exact,
                              graph=False)
    sim.update_params(dt=0.005, G=1.0, epsilon=eps, theta=theta)
    restart(sim)
    for _ in range(args.steps):
        sim.step(1, sort_every=1)
        sim.sync()
    direct = sim.stage_times(reset=True)
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
            "interactions_first_last": [int(counts[0]), int(counts[-1])],
            "walk_avg_ms": walk_avg_ms,
            "walk_algorithmic_bytes": (68 + 64 * nu) * n,
            # the walk against the HBM roofline too: its algorithmic bytes per
            # launch over its launch time (the rocprof-measured rate is added
            # when the counter passes run)
            "hbm": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                    "achieved": (68 + 64 * nu) * n / (walk_avg_ms * 1e-3) / 1e9,
                    "frac": (68 + 64 * nu) * n / (walk_avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "measured_gbs": None, "measured_frac": None},
        },
        "step_hbm": {
            "bytes_per_particle": 236 + 160 * nu, "nu": nu,
            "achieved_gbs": step_bytes / (ms_per_step * 1e-3) / 1e9, "peak_gbs": HBM_PEAK_GBS,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "stages_ms": {k: (v[0] / v[1] if v[1] else 0.0) for k, v in direct.items()},
        "stages_note": "stages_ms: the K timed steps again with direct launches, an event "
                       "pair per stage and a synchronisation per step; the timed steps replay "
                       "captured graphs",
        "timing_note": "every figure covers the same K steps from the initial particles after "
                       "W untimed ones (the collision evolves); value and ms_per_step: the K "
                       "steps launched back to back with one synchronisation at each end; "
                       "p50_ms / p99_ms: the K steps again, each followed by a stream "
                       "synchronisation; interactions_per_step: their mean, each step counted "
                       "on a third pass",
        "interactions_per_s": inter / (ms_per_step * 1e-3),
        "cpu_baseline": None,
    }
    log(f"headline: {out['value'] / 1e6:.1f} M particle-steps/s, walk {walk_avg_ms:.3f} ms")
    acc, parts = gpu_forces(data, bucket, theta, eps, args.exact)
    out["force_error_fp64"] = force_error(acc, parts, eps, 4096, 0x6A1A5EED00000100 + n)
    del acc, parts
    log(f"headline: FP64 direct-sum error on 4096 particles median "
        f"{out['force_error_fp64']['median']:.3e} p99 {out['force_error_fp64']['p99']:.3e}")
    if not args.no_cpu_baseline:
        log(f"cpu baseline: {args.cpu_steps} reference steps of all {n} particles on the host")
        c = dict(ic.CONFIGS["collision_4m"], n=n)
        # the first steps of the GPU's timed window
        cb = cpu_reference_steps(data, c, args.cpu_steps, warmup=args.warmup)
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"].update({k: cb[k] for k in ("step_s", "p50_s", "p99_s", "stages_s")})
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
        log(f"cpu baseline: {cb['value'] / 1e6:.3f} M particle-steps/s ({cb['step_s']:.2f} s/step "
            f"on {cb['cores']} threads)")
    return out


if __name__ == "__main__":
    main()
