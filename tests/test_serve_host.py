"""Host logic of the continuous-batching server (admission by slots, latency
cap and HBM budget, retirement, budget accounting) with a stub arena that
charges each galaxy galaxy_hip.arena_charge(n) bytes against the budget:
no GPU needed."""
import numpy as np
import pytest

from galaxy_hip import ARENA_NODE_BYTES, ARENA_PARTICLE_BYTES, ArenaFull, GS_E_FULL, arena_charge
from galaxy_hip import serve


class StubArena:
    def __init__(self, budget, nslots):
        self.budget = budget
        self.n = [0] * nslots
        self.log = []

    def used(self):
        return sum(arena_charge(n) for n in self.n)

    def upload(self, slot, records, **params):
        n = records.shape[0]
        self.n[slot] = 0
        if self.used() + arena_charge(n) > self.budget:
            self.log.append(("full", slot, n))
            raise ArenaFull(GS_E_FULL, "arena full")
        self.n[slot] = n
        self.log.append(("upload", slot, n))

    def set_tree(self, slot, bucket, max_nodes=0):
        pass

    def step(self, nsteps=1, sort_every=1):
        self.log.append(("step", tuple(self.n)))

    def sync(self):
        pass

    def particles(self, slot):
        return np.zeros((self.n[slot], 8), np.float32)

    def reduce(self, slot):
        return {"mass": float(self.n[slot])}

    def clear(self, slot):
        self.n[slot] = 0
        self.log.append(("clear", slot))

    def arena_info(self):
        return {"device_bytes": self.budget, "used_bytes": self.used(), "largest_fit": 0,
                "budget_bytes": self.budget}

    def destroy(self):
        pass


def _job(i, n, steps=3):
    return serve.Job(i, np.zeros((n, 8), np.float32), dict(dt=0.01, G=1.0, epsilon=0.1), steps)


def _server(budget=10**12, slots=8, **kw):
    return serve.GalaxyServer(hbm_budget=budget, max_jobs=slots, pool_factory=StubArena, **kw)


def test_arena_charge_is_ranges():
    # n rounded up to 1024 particle records, 2n + 4096 node records
    assert arena_charge(0) == 0
    assert arena_charge(1) == 1024 * ARENA_PARTICLE_BYTES + 4098 * ARENA_NODE_BYTES
    assert arena_charge(1024) == 1024 * ARENA_PARTICLE_BYTES + 6144 * ARENA_NODE_BYTES
    assert arena_charge(1025) - arena_charge(1024) == \
        1024 * ARENA_PARTICLE_BYTES + 2 * ARENA_NODE_BYTES


def test_budget_queues_a_job_while_slots_are_free():
    """The budget, not slot availability nor the latency cap, holds the third
    job: two 100,000-particle galaxies fill the budget, six slots stay free,
    and the job enters once a retirement gives its range back."""
    budget = 2 * arena_charge(100_000) + arena_charge(50_000)
    srv = _server(budget, slots=8)
    a, b, c = _job(0, 100_000, steps=2), _job(1, 100_000, steps=4), _job(2, 100_000, steps=1)
    for j in (a, b, c):
        srv.submit(j)
    assert srv.admit() == 2
    assert c.slot == -1 and c.held_by == {"budget"}
    assert len(srv.free) == 6 and srv.resident_bytes == 2 * arena_charge(100_000)
    small = _job(3, 40_000, steps=1)  # a smaller one behind it still fits
    srv.submit(small)
    assert srv.admit() == 1 and small.slot >= 0 and c.slot == -1
    srv.step()
    srv.step()  # a retires after 2 steps
    assert a.finished is not None and srv.admit() == 1 and c.slot >= 0
    summ = srv.summary(1.0)
    assert summ["refusals"]["budget"] >= 2 and summ["refusals"]["slots"] == 0
    assert summ["refusals"]["latency"] == 0
    assert summ["budget_used_peak_bytes"] <= budget


def test_slots_and_latency_hold_jobs():
    srv = _server(slots=2)
    for i in range(3):
        srv.submit(_job(i, 10))
    assert srv.admit() == 2
    assert srv.waiting[0].held_by == {"slots"}
    srv2 = _server(slots=4, resident_limit=1500)
    for i in range(3):
        srv2.submit(_job(i, 800))
    assert srv2.admit() == 1  # 1600 > 1500: the others wait
    assert srv2.resident_n == 800
    assert all(j.held_by == {"latency"} for j in srv2.waiting)


def test_job_larger_than_the_budget_is_reported():
    srv = _server(arena_charge(1000), slots=2)
    with pytest.raises(RuntimeError, match="never be admitted"):
        srv.run([_job(0, 5000)], realtime=False)


def test_run_retires_every_job_after_its_steps():
    srv = _server(slots=2)
    jobs = [_job(i, 10 + i, steps=2 + i) for i in range(5)]
    summ = srv.run(jobs, realtime=False)
    assert summ["jobs"] == 5
    assert all(j.done == j.steps for j in jobs)
    assert all(j.reduce["mass"] == j.n for j in jobs)
    assert summ["particle_steps"] == sum(j.n * j.steps for j in jobs)
    assert srv.resident_n == 0 and srv.resident_bytes == 0 and sorted(srv.free) == [0, 1]
    assert summ["jobs_held_by"]["slots"] == 3
    # never more than two jobs resident, and each step steps exactly the residents
    steps = [e for e in srv.pool.log if e[0] == "step"]
    assert all(sum(1 for x in e[1] if x) <= 2 for e in steps)


def test_run_under_a_tight_budget_serves_everything():
    """A stream whose jobs do not all fit at once: every job is served, the
    budget is never exceeded, and the budget held some of them."""
    rng = np.random.default_rng(5)
    sizes = rng.integers(1000, 60000, 40)
    budget = 3 * arena_charge(60000)
    srv = _server(budget, slots=16)
    jobs = [_job(i, int(n), steps=1 + i % 4) for i, n in enumerate(sizes)]
    summ = srv.run(jobs, realtime=False)
    assert summ["jobs"] == 40
    assert summ["budget_used_peak_bytes"] <= budget
    assert summ["jobs_held_by"]["budget"] > 0
    assert all(e[0] != "step" or sum(arena_charge(n) for n in e[1]) <= budget
               for e in srv.pool.log)
