"""Continuous-batching server for independent galaxy simulations on one GPU.

Jobs (a galaxy's particle records, its updateParams dictionary and a step
count) are admitted into the slots of one GalaxyArena and stepped together:
every server step is one replay of the arena's captured step graph (rebuild
+ refresh + walk for every resident galaxy), so a job that arrives while
others run joins the next step, and a finished job frees its slot and its
arena range between steps.  Particle and tree state stay resident in HBM for
the life of a job.

The HBM budget is the arena: its device state is sized from the budget once,
and a job is placed in a free range of it when admitted (arena_charge(n)
bytes: n rounded up to 1024 particle records, 2n + 4096 node records) and
gives the range back when it retires.  Admission, per waiting job in arrival
order (a job that does not fit does not block smaller ones behind it):
  * slots: a free slot (the number of galaxies resident at once);
  * latency: the resident particle total stays within `resident_limit`, so a
    step's latency stays close to its steady value instead of growing with
    the queue (steps scale with the resident particle count);
  * budget: a free range of the arena holds the job (the arena refuses it
    otherwise, and the job waits until retirements give ranges back).
The summary reports how many jobs each constraint held, and the budget use.

The reference runs one galaxy per NBodyGPU and has no server; the per-job
arithmetic is the single-galaxy step, and a job's result is bit-identical to
running it alone (slot-local indexing, same wavefront composition).
"""
import collections
import time

import numpy as np

from . import ArenaFull, GalaxyArena, arena_charge

REASONS = ("slots", "latency", "budget")


class Job:
    __slots__ = ("id", "records", "params", "steps", "arrival", "bucket", "slot", "done",
                 "admitted", "finished", "result", "reduce", "n", "held_by")

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
        self.held_by = set()  # the constraints that kept it waiting at some admission


class GalaxyServer:
    def __init__(self, hbm_budget=200e9, max_jobs=64, resident_limit=None, bucket=16,
                 walk_wg=0, radix_bits=0, exact_force=False, keep_results=True,
                 pool_factory=None):
        """hbm_budget: device bytes of the arena; max_jobs: its slots (jobs
        resident at once, at most 256); resident_limit: resident particles
        (None: no limit).  pool_factory(budget, slots) replaces the
        GalaxyArena (host-side tests)."""
        if pool_factory is None:
            self.pool = GalaxyArena(hbm_budget, max_jobs, walk_wg=walk_wg, radix_bits=radix_bits,
                                    exact_force=exact_force)
        else:
            self.pool = pool_factory(hbm_budget, max_jobs)
        self.hbm_budget = int(hbm_budget)
        self.max_jobs = int(max_jobs)
        self.free = list(range(self.max_jobs))
        self.resident_limit = resident_limit
        self.keep_results = keep_results
        self.waiting = collections.deque()
        self.active = {}
        self.finished = []
        self.resident_n = 0
        self.resident_bytes = 0  # sum of arena_charge over resident jobs
        self.refusals = dict.fromkeys(REASONS, 0)  # admission attempts refused, per reason
        self.step_lat = []
        self.step_particles = []
        self.step_jobs = []
        self.step_queued = []  # jobs waiting for admission when the step ran
        self.step_bytes = []   # arena bytes taken by the resident jobs when the step ran
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
        self.waiting.append(job)

    def _hold(self, job, reason):
        self.refusals[reason] += 1
        job.held_by.add(reason)

    def admit(self, now=None):
        """Move waiting jobs into free slots and arena ranges; returns the
        number admitted."""
        admitted = 0
        keep = collections.deque()
        while self.waiting:
            job = self.waiting.popleft()
            if not self.free:
                self._hold(job, "slots")
                keep.append(job)
                continue
            if (self.resident_limit is not None and self.active
                    and self.resident_n + job.n > self.resident_limit):
                self._hold(job, "latency")
                keep.append(job)
                continue
            slot = self.free[0]
            try:
                self.pool.upload(slot, job.records, **job.params)
            except ArenaFull:
                self._hold(job, "budget")
                keep.append(job)
                continue
            self.free.pop(0)
            self.pool.set_tree(slot, job.bucket)
            job.slot = slot
            job.admitted = time.perf_counter() if now is None else now
            self.active[slot] = job
            self.resident_n += job.n
            self.resident_bytes += arena_charge(job.n)
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
        self.step_bytes.append(self.resident_bytes)
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
        self.pool.clear(slot)  # the slot's arena ranges go back
        del self.active[slot]
        self.free.append(slot)
        self.free.sort()
        self.resident_n -= job.n
        self.resident_bytes -= arena_charge(job.n)
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
                        f"{self.resident_n} particles / {self.resident_bytes / 1e9:.2f} GB, "
                        f"waiting {len(self.waiting)}, finished {len(self.finished)}")
            elif self.waiting:
                raise RuntimeError(f"{len(self.waiting)} jobs can never be admitted (each "
                                   "needs more than the whole budget)")
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
        sb = np.array(self.step_bytes, np.float64)
        info = self.pool.arena_info() if hasattr(self.pool, "arena_info") else {}
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
            "hbm_budget": self.hbm_budget,
            "arena_device_bytes": info.get("device_bytes"),
            "budget_used_peak_bytes": float(sb.max()) if sb.size else 0.0,
            "budget_used_mean_bytes": float(sb.mean()) if sb.size else 0.0,
            "budget_used_peak_frac": float(sb.max()) / self.hbm_budget if sb.size else 0.0,
            "slots": self.max_jobs,
            "resident_limit": self.resident_limit,
            # jobs that some admission found held by each constraint, and the
            # refused admission attempts per constraint
            "jobs_held_by": {r: sum(r in j.held_by for j in self.finished) for r in REASONS},
            "refusals": dict(self.refusals),
        }
