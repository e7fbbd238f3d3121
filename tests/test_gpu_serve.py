"""The continuous-batching server on the GPU: jobs that arrive while others
run are admitted as slots and arena ranges free up, and each job's result is
the same alone and in the mix (bit-identical: a slot runs the single-galaxy
arithmetic on slot-local indices).  In the second test the HBM budget, not
the slots, holds jobs back."""
import numpy as np
import pytest

import galaxy_hip
from galaxy_hip import arena_charge, ic, serve
from gpu_util import banner

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=0.005, G=1.0, epsilon=0.01, theta=0.5)


def _jobs():
    spec = ic.serve_jobs(9, rate=20.0, lo=2048, hi=40000, steps=4)
    out = []
    for j in spec:
        j["steps"] = 3 + j["id"] % 4  # uneven lengths: jobs leave at different steps
        out.append((j, ic.job_records(j)))
    return out


def _check_alone(jobs, specs, exact):
    for job, (spec, rec) in zip(jobs, specs):
        banner(f"  job {spec['id']} n={spec['n']} {spec['preset']} steps={spec['steps']}")
        sim = galaxy_hip.NBodyGPU(rec, bucket=16, exact_force=exact)
        try:
            sim.update_params(**PARAMS)
            sim.step(spec["steps"], sort_every=1)
            want = sim.particles()
            red = sim.reduce()
        finally:
            sim.destroy()
        assert np.array_equal(job.result.view(np.uint32), want.view(np.uint32))
        for k in red:
            assert job.reduce[k] == pytest.approx(red[k], rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("exact", [False, True])
def test_served_jobs_equal_alone(exact):
    banner(f"served vs alone exact={exact}")
    specs = _jobs()
    # 6 slots for 9 jobs: some must wait
    with serve.GalaxyServer(hbm_budget=2e9, max_jobs=6, resident_limit=60000,
                            exact_force=exact) as srv:
        jobs = [serve.Job(j["id"], r, PARAMS, j["steps"], 0.0) for j, r in specs]
        summ = srv.run(jobs, realtime=False)
    assert summ["jobs"] == len(specs)
    assert any(j.admitted > min(k.finished for k in jobs) for j in jobs), \
        "no job entered while others ran"
    _check_alone(jobs, specs, exact)


def test_budget_holds_jobs_and_results_equal_alone():
    banner("served under a tight HBM budget")
    specs = _jobs()
    big = max(j["n"] for j, _ in specs)
    budget = int(2.5 * arena_charge(big))
    with serve.GalaxyServer(hbm_budget=budget, max_jobs=9) as srv:
        jobs = [serve.Job(j["id"], r, PARAMS, j["steps"], 0.0) for j, r in specs]
        summ = srv.run(jobs, realtime=False)
    banner(f"  budget {budget} B: peak use {summ['budget_used_peak_bytes']:.0f} B, held "
           f"{summ['jobs_held_by']}")
    assert summ["jobs"] == len(specs)
    assert summ["jobs_held_by"]["budget"] > 0 and summ["jobs_held_by"]["slots"] == 0
    assert summ["budget_used_peak_bytes"] <= budget
    _check_alone(jobs, specs, False)
