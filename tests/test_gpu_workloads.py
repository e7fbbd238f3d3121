"""Parity at the batched and served workloads as benchmarked, and the default
force path's 1e-5 bar on every particle of the 500,000-particle disk.

* Batched (64 galaxies, N log-uniform in [4096, 262144], disk / Plummer): one
  64-slot pool with the tuned set-up the benchmark times (walk workgroup 64,
  bucket 24, 8-bit digits, theta 0.6, eps 0.01), two steps; every slot equals
  its galaxy stepped alone bit for bit, and in an EXACT pool two slots equal
  the oracle's steps bit for bit.
* Served (Poisson job stream, N log-uniform in [8192, 2^20], disk / Plummer /
  collision): eleven jobs of the stream's recipe, the two largest near 2^20
  particles, served from an arena with fewer slots than jobs (so jobs enter
  while others run) with the benchmark's bucket 24; each result equals the
  job run alone.
* The default path's forces on all 500,000 particles of the gas disk against
  the transliterated walk, to the bar of gpu_util.check_force_bar (the
  headline's 4,194,304 are checked in test_gpu_full.py)."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import ic, serve
from gpu_util import banner, check_force_bar, oracle_tree, rel_err
from oracle import oracle

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5
BATCH = dict(dt=0.005, G=1.0, epsilon=0.01, theta=0.6)
TUNED = dict(walk_wg=64, bucket=24, radix_bits=8)


def _alone(data, bucket, params, exact, nsteps, walk_wg=0, radix_bits=0):
    sim = galaxy_hip.NBodyGPU(data, bucket=bucket, exact_force=exact, walk_wg=walk_wg,
                              radix_bits=radix_bits)
    try:
        sim.update_params(**params)
        sim.step(nsteps, sort_every=1)
        return sim.particles(), sim.energy()
    finally:
        sim.destroy()


def _batch_pool(galaxies, exact):
    caps = [g.shape[0] for _, g in galaxies]
    pool = galaxy_hip.GalaxyPool(caps, buckets=[TUNED["bucket"]] * len(caps), exact_force=exact,
                                 walk_wg=TUNED["walk_wg"], radix_bits=TUNED["radix_bits"])
    for k, (_, g) in enumerate(galaxies):
        pool.upload(k, g, **BATCH)
    return pool


def test_batch64_pool_equals_alone():
    galaxies = ic.batch64()
    banner(f"batch64 pool: {len(galaxies)} galaxies, "
           f"N total {sum(g.shape[0] for _, g in galaxies)}")
    pool = _batch_pool(galaxies, exact=False)
    try:
        pool.step(2, sort_every=1)
        got = [(pool.particles(k), pool.energy(k)) for k in range(len(galaxies))]
    finally:
        pool.destroy()
    for k, (preset, g) in enumerate(galaxies):
        parts, energy = _alone(g, TUNED["bucket"], BATCH, False, 2, TUNED["walk_wg"],
                               TUNED["radix_bits"])
        assert np.array_equal(got[k][0].view(np.uint32), parts.view(np.uint32)), (k, preset)
        assert np.array_equal(got[k][1].view(np.uint32), energy.view(np.uint32)), (k, preset)


def test_batch64_exact_pool_slots_equal_oracle():
    galaxies = ic.batch64()
    sizes = [g.shape[0] for _, g in galaxies]
    picks = [int(np.argmax(sizes))]  # the largest galaxy
    picks.append(next(k for k, (p, g) in enumerate(galaxies)
                      if p != galaxies[picks[0]][0]))  # and one of the other preset
    banner(f"batch64 EXACT pool, slots {picks} vs oracle")
    pool = _batch_pool(galaxies, exact=True)
    try:
        pool.step(2, sort_every=1)
        got = {k: (pool.particles(k), pool.energy(k)) for k in picks}
    finally:
        pool.destroy()
    for k in picks:
        g = galaxies[k][1]
        n = g.shape[0]
        op = oracle.pack_params(n, dt=BATCH["dt"], G=BATCH["G"], epsilon=BATCH["epsilon"],
                                theta=BATCH["theta"])
        cur, went = g.reshape(-1).copy(), None
        for _ in range(2):
            ref = oracle.sort_and_build(cur, n, bucket=TUNED["bucket"])
            cur, went, _, _, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
        assert np.array_equal(got[k][0].reshape(-1).view(np.uint32), cur.view(np.uint32)), k
        assert np.array_equal(got[k][1].view(np.uint32), went.view(np.uint32)), k


def test_served_stream_jobs_equal_alone():
    stream = ic.serve_jobs(200)
    big = sorted(stream, key=lambda j: -j["n"])[:2]
    specs = stream[:9] + big
    assert any(j["preset"] == "collision" for j in specs)
    assert {j["preset"] for j in specs} == {"disk", "plummer", "collision"}
    params = dict(dt=0.005, G=1.0, epsilon=0.01, theta=0.5)
    bucket = TUNED["bucket"]
    recs = [ic.job_records(j) for j in specs]
    # 8 slots for 11 jobs: some must wait
    with serve.GalaxyServer(hbm_budget=4e9, max_jobs=8, resident_limit=3 << 20,
                            bucket=bucket) as srv:
        jobs = [serve.Job(j["id"], r, params, 2 + j["id"] % 3, 0.0, bucket=bucket)
                for j, r in zip(specs, recs)]
        summ = srv.run(jobs, realtime=False)
    assert summ["jobs"] == len(specs)
    assert any(j.admitted > min(k.finished for k in jobs) for j in jobs), \
        "no job entered while others ran"
    for job, spec, rec in zip(jobs, specs, recs):
        banner(f"  job {spec['id']} n={spec['n']} {spec['preset']} steps={job.steps}")
        parts, _ = _alone(rec, bucket, params, False, job.steps)
        assert np.array_equal(job.result.view(np.uint32), parts.view(np.uint32)), spec["id"]


def test_fast_forces_every_particle_disk_500k():
    c = ic.CONFIGS["disk_500k"]
    data = ic.make("disk_500k")
    n = data.shape[0]
    banner(f"disk_500k default-path forces on all {n} particles")
    ref = oracle_tree(data, bucket=c["bucket"])
    sim = galaxy_hip.NBodyGPU(data, bucket=c["bucket"])
    try:
        sim.update_params(dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"])
        sim.sort()
        sim.refresh()
        sim.force()
        acc, dens = sim.accel()
    finally:
        sim.destroy()
    op = oracle.pack_params(n, dt=c["dt"], G=1.0, epsilon=c["epsilon"], theta=c["theta"])
    _, _, racc, rdens, _ = oracle.walk(ref["sorted"], ref["nodes"], op)
    err = rel_err(acc[:, :3], racc[:, :3])
    perr = np.abs(acc[:, 3].astype(np.float64) - racc[:, 3]) / np.abs(racc[:, 3].astype(np.float64))
    worst, rel_big, abs_rms, ncancel = check_force_bar(acc, racc, FAST_TOL)
    banner(f"  relative force error max {worst:.3e} (p99 {np.quantile(err, 0.99):.3e}); "
           f"{ncancel} particles below 1% of the RMS force; max error {abs_rms:.3e} of it; "
           f"max rel potential err {perr.max():.3e}")
    assert perr.max() <= FAST_TOL
    assert np.array_equal(dens.view(np.uint32), rdens.view(np.uint32))
