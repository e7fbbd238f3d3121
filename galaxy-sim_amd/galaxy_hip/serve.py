"""Continuous-batching server for independent galaxy simulations on one GPU.

Jobs (a galaxy's particle records, its updateParams dictionary and a step
count) are admitted into slots of one GalaxyPool and stepped together: every
server step is one replay of the pool's captured step graph (rebuild +
refresh + walk for every resident galaxy), so a job that arrives while others
run joins the next step, and a finished job frees its slot between steps.
Particle and tree state stay resident in HBM for the life of a job.

Admission, per waiting job in arrival order (a job that does not fit does not
block smaller ones behind it):
  * a free slot whose capacity holds the job (smallest such slot);
  * HBM: the job's state footprint (state_bytes(n)) plus that of every
    resident job stays within the budget, and the pool itself was sized
    within the budget;
  * latency: the resident particle total stays within `resident_limit`, so a
    step's latency stays close to its steady value instead of growing with
    the queue (steps scale with the resident particle count).

The reference runs one galaxy per NBodyGPU and has no server; the per-job
arithmetic is the single-galaxy step, and a job's result is bit-identical to
running it alone (slot-local indexing, same wavefront composition).
"""
import collections
import time

import numpy as np

from . import GalaxyPool

# device bytes per particle of resident state: 2 x (posm + velp) ping-pong,
# 2 x (key + index) sort ping-pong, raw code, energy, acceleration, density,
# interaction count; plus a node pool of 2n + 4096 32-byte records and their
# sizes
PARTICLE_STATE_BYTES = 2 * 32 + 2 * 8 + 4 + 4 + 16 + 4 + 4
NODE_BYTES = 32 + 4


def state_bytes(n):
    return n * PARTICLE_STATE_BYTES + (2 * n + 4096) * NODE_BYTES


def slot_classes(lo, hi, per_class):
    """Power-of-two slot capacities covering [lo, hi]: `per_class` slots of
    each capacity 2^k (an int or a function of the capacity)."""
    caps = []
    c = 1 << int(np.ceil(np.log2(lo)))
    while True:
        k = per_class(c) if callable(per_class) else per_class
        caps += [c] * int(k)
        if c >= hi:
            break
        c <<= 1
    return caps


class Job:
    __slots__ = ("id", "records", "params", "steps", "arrival", "bucket", "slot", "done",
                 "admitted", "finished", "result", "reduce", "n")

    def __init__(self, id, records, params, steps, arrival=0.0, bucket=16):
        self.id = id
        self.records = np.ascontiguousarray(records, np.float32).reshape(-1, 8)
        self.n = self.records.shape[0]
        self.params = dict(params)
        self.steps = int(steps)
        self.arrival = arrival
        self.bucket = bucket
        self.slot = -1
        self.done = 0
        self.admitted = None
        self.finished = None
        self.result = None
        self.reduce = None


class GalaxyServer:
    def __init__(self, slot_caps, hbm_budget=200e9, resident_limit=None, bucket=16,
                 walk_wg=0, radix_bits=0, exact_force=False, keep_results=True,
                 pool_factory=None):
        """slot_caps: slot capacities of the pool (see slot_classes).
        pool_factory(caps) replaces the GalaxyPool (host-side tests)."""
        caps = sorted(int(c) for c in slot_caps)
        pool_bytes = sum(state_bytes(c) for c in caps)
        if pool_bytes > hbm_budget:
            raise ValueError(f"slot pool needs {pool_bytes / 1e9:.1f} GB, over the "
                             f"{hbm_budget / 1e9:.1f} GB budget")
        if pool_factory is None:
            self.pool = GalaxyPool(caps, buckets=[bucket] * len(caps), walk_wg=walk_wg,
                                   radix_bits=radix_bits, exact_force=exact_force)
        else:
            self.pool = pool_factory(caps)
        self.caps = np.array(caps, np.int64)
        self.free = list(range(len(caps)))  # ascending capacity
        self.hbm_budget = hbm_budget
        self.pool_bytes = pool_bytes
        self.resident_limit = resident_limit or int(self.caps.sum())
        self.keep_results = keep_results
        self.waiting = collections.deque()
        self.active = {}
        self.finished = []
        self.resident_n = 0
        self.resident_bytes = 0
        self.step_lat = []
        self.step_particles = []
        self.step_jobs = []
        self.step_queued = []  # jobs waiting for admission when the step ran
        self.step_end = []
        self.particle_steps = 0

    def close(self):
        self.pool.destroy()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- admission ---------------------------------------------------------
    def submit(self, job):
        if job.n > self.caps.max():
            raise ValueError(f"job {job.id}: n={job.n} exceeds the largest slot")
        self.waiting.append(job)

    def _fits(self, job):
        if self.resident_n + job.n > self.resident_limit and self.active:
            return -1
        if self.resident_bytes + state_bytes(job.n) > self.hbm_budget:
            return -1
        for i, s in enumerate(self.free):
            if self.caps[s] >= job.n:
                return i
        return -1

    def admit(self, now=None):
        """Move waiting jobs into free slots; returns the number admitted."""
        admitted = 0
        keep = collections.deque()
        while self.waiting:
            job = self.waiting.popleft()
            i = self._fits(job)
            if i < 0:
                keep.append(job)
                continue
            slot = self.free.pop(i)
            self.pool.upload(slot, job.records, **job.params)
            self.pool.set_tree(slot, job.bucket)
            job.slot = slot
            job.admitted = time.perf_counter() if now is None else now
            self.active[slot] = job
            self.resident_n += job.n
            self.resident_bytes += state_bytes(job.n)
            admitted += 1
        self.waiting = keep
        return admitted

    # -- stepping ----------------------------------------------------------
    def step(self):
        """One step of every resident job; retires the jobs that reach their
        step count.  Returns the list of jobs finished by this step."""
        if not self.active:
            return []
        t0 = time.perf_counter()
        self.pool.step(1, sort_every=1)
        self.pool.sync()
        self.step_lat.append(time.perf_counter() - t0)
        self.step_particles.append(self.resident_n)
        self.step_jobs.append(len(self.active))
        self.step_queued.append(len(self.waiting))
        self.step_end.append(time.perf_counter())
        self.particle_steps += self.resident_n
        done = []
        for slot, job in list(self.active.items()):
            job.done += 1
            if job.done >= job.steps:
                done.append(job)
        for job in done:
            self._retire(job)
        return done

    def _retire(self, job):
        slot = job.slot
        job.reduce = self.pool.reduce(slot)  # FP64 energy / momentum of the last step
        if self.keep_results:
            job.result = self.pool.particles(slot)
        self.pool.clear(slot)
        del self.active[slot]
        self.free.append(slot)
        self.free.sort(key=lambda s: (self.caps[s], s))
        self.resident_n -= job.n
        self.resident_bytes -= state_bytes(job.n)
        job.finished = time.perf_counter()
        self.finished.append(job)

    def run(self, jobs, realtime=True, log=None):
        """Serve `jobs` (sorted by .arrival, seconds from the start).  With
        realtime, a job becomes visible when the wall clock passes its
        arrival; otherwise all are visible at once.  Returns a summary."""
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        t_start = time.perf_counter()
        first_busy = None
        while pending or self.waiting or self.active:
            now = time.perf_counter() - t_start
            while pending and (not realtime or pending[0].arrival <= now):
                j = pending.popleft()
                j.arrival = t_start + j.arrival  # absolute
                self.submit(j)
            self.admit()
            if self.active:
                if first_busy is None:
                    first_busy = time.perf_counter()
                self.step()
                if log and len(self.step_lat) % 100 == 0:
                    log(f"server step {len(self.step_lat)}: resident {len(self.active)} jobs / "
                        f"{self.resident_n} particles, waiting {len(self.waiting)}, "
                        f"finished {len(self.finished)}")
            elif pending:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - now)))
        t_end = time.perf_counter()
        return self.summary(t_end - (first_busy or t_end))

    def _saturated_rate(self):
        """Particle-steps per wall second over the steps that ran while jobs
        were waiting (each step's interval counts its admission work too)."""
        q = np.array(self.step_queued[1:])
        if not q.size or not (q > 0).any():
            return 0.0
        sel = q > 0
        dt = np.diff(np.array(self.step_end))[sel]
        return float(np.array(self.step_particles[1:])[sel].sum() / dt.sum())

    def summary(self, busy_s):
        """Step latency is reported per job step (each server step counted once
        for every job it advanced: the latency a job sees per step) and per
        server step."""
        lat = np.array(self.step_lat) * 1e3
        jlat = np.repeat(lat, np.array(self.step_jobs, np.int64)) if lat.size else lat
        jl = np.array([j.finished - j.arrival for j in self.finished]) if self.finished else []
        wait = np.array([j.admitted - j.arrival for j in self.finished]) if self.finished else []
        return {
            "jobs": len(self.finished),
            "steps": len(self.step_lat),
            "particle_steps": int(self.particle_steps),
            "busy_s": busy_s,
            "throughput": self.particle_steps / busy_s if busy_s > 0 else 0.0,
            "step_p50_ms": float(np.percentile(jlat, 50)) if jlat.size else 0.0,
            "step_p99_ms": float(np.percentile(jlat, 99)) if jlat.size else 0.0,
            "server_step_p50_ms": float(np.percentile(lat, 50)) if lat.size else 0.0,
            "server_step_p99_ms": float(np.percentile(lat, 99)) if lat.size else 0.0,
            "mean_resident_particles": float(np.mean(self.step_particles))
            if self.step_particles else 0.0,
            # while jobs wait for admission the server runs at its resident
            # limit: the sustained rate, free of the stream's ramp-up and drain
            "saturated_throughput": self._saturated_rate(),
            "job_latency_p50_s": float(np.percentile(jl, 50)) if len(jl) else 0.0,
            "job_latency_p99_s": float(np.percentile(jl, 99)) if len(jl) else 0.0,
            "queue_wait_p50_s": float(np.percentile(wait, 50)) if len(wait) else 0.0,
            "queue_wait_p99_s": float(np.percentile(wait, 99)) if len(wait) else 0.0,
            "pool_bytes": self.pool_bytes,
            "hbm_budget": self.hbm_budget,
            "slots": int(self.caps.size),
        }
