"""Host logic of the continuous-batching server (admission, slot choice,
retirement, budgets) with a stub pool: no GPU needed."""
import numpy as np
import pytest

from galaxy_hip import serve


class StubPool:
    def __init__(self, caps):
        self.caps = caps
        self.n = [0] * len(caps)
        self.log = []

    def upload(self, slot, records, **params):
        assert records.shape[0] <= self.caps[slot]
        self.n[slot] = records.shape[0]
        self.log.append(("upload", slot, records.shape[0]))

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

    def destroy(self):
        pass


def _job(i, n, steps=3):
    return serve.Job(i, np.zeros((n, 8), np.float32), dict(dt=0.01, G=1.0, epsilon=0.1), steps)


def test_slot_classes_cover_range():
    caps = serve.slot_classes(8192, 1 << 20, 2)
    assert caps[0] == 8192 and caps[-1] == 1 << 20
    assert len(caps) == 2 * 8
    assert serve.slot_classes(5000, 9000, lambda c: 1) == [8192, 16384]


def test_state_bytes_counts_nodes():
    assert serve.state_bytes(0) == 4096 * serve.NODE_BYTES
    assert serve.state_bytes(10) - serve.state_bytes(9) == \
        serve.PARTICLE_STATE_BYTES + 2 * serve.NODE_BYTES


def test_pool_over_budget_is_refused():
    with pytest.raises(ValueError):
        serve.GalaxyServer([1 << 20] * 4, hbm_budget=1e6, pool_factory=StubPool)


def test_admission_smallest_fit_and_skip_ahead():
    srv = serve.GalaxyServer([100, 1000, 1000], pool_factory=StubPool)
    big1, big2, small = _job(0, 900), _job(1, 800), _job(2, 50)
    for j in (big1, big2, small):
        srv.submit(j)
    assert srv.admit() == 3
    assert small.slot == 0  # the 100-slot, not a 1000-slot
    assert {big1.slot, big2.slot} == {1, 2}
    srv.submit(_job(3, 700))
    assert srv.admit() == 0  # no free slot: waits
    for _ in range(3):
        srv.step()
    assert len(srv.finished) == 3 and not srv.active
    assert srv.admit() == 1


def test_resident_limit_and_hbm_budget():
    srv = serve.GalaxyServer([1000] * 4, resident_limit=1500, pool_factory=StubPool)
    for i in range(4):
        srv.submit(_job(i, 800))
    assert srv.admit() == 1  # 1600 > 1500: the second waits
    assert srv.resident_n == 800
    budget = serve.state_bytes(1000) * 4
    srv2 = serve.GalaxyServer([1000] * 4, hbm_budget=budget, pool_factory=StubPool)
    srv2.resident_bytes = budget - serve.state_bytes(10) + 1
    srv2.submit(_job(0, 10))
    assert srv2.admit() == 0  # the job's footprint would cross the budget


def test_run_retires_every_job_after_its_steps():
    srv = serve.GalaxyServer([64, 64], pool_factory=StubPool)
    jobs = [_job(i, 10 + i, steps=2 + i) for i in range(5)]
    summ = srv.run(jobs, realtime=False)
    assert summ["jobs"] == 5
    assert all(j.done == j.steps for j in jobs)
    assert all(j.reduce["mass"] == j.n for j in jobs)
    assert summ["particle_steps"] == sum(j.n * j.steps for j in jobs)
    assert srv.resident_n == 0 and sorted(srv.free) == [0, 1]
    # never more than two jobs resident, and each step steps exactly the residents
    steps = [e for e in srv.pool.log if e[0] == "step"]
    assert all(sum(1 for x in e[1] if x) <= 2 for e in steps)
