"""Serving policy on the CPU with a fake backend: admission by HBM footprint, slot
pools, LRU grid eviction, continuous batching, failure of oversized jobs."""
import numpy as np

from attract_jax_amd.serve import JobSpec, Scheduler, SlotPool


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeBackend:
    """Chunks take `cost` seconds per pose-step / 1e6; grids cost grid_bytes[key]."""

    def __init__(self, clock, grid_bytes):
        self.clock = clock
        self.gb = grid_bytes
        self.launches = []
        self.dropped = []

    def grid_bytes(self, key):
        return self.gb[key]

    def make_group(self, g):
        g.handle = {}

    def grow_group(self, g, cap):
        g.handle["grown"] = cap

    def drop_group(self, g):
        self.dropped.append(g.key)

    def load_job(self, g, job):
        job.handle = {}

    def launch(self, g, jobs, steps, dirty):
        work = sum(j.npose * min(steps, j.remaining) for j in jobs)
        self.launches.append((g.key, [j.id for j in jobs], steps))
        return self.clock() + 1e-6 * work

    def finished(self, token):
        return self.clock() >= token

    def fetch_job(self, g, job):
        return {"npose": job.npose}

    def release_job(self, g, job):
        pass


def spec(key, n, steps, arrival=0.0):
    return JobSpec(key, np.zeros((3, 3)), np.ones(3, int), np.zeros((n, 4)), np.zeros((n, 3)), steps,
                   arrival=arrival)


def drive(s, clock, dt=1e-3, limit=10000):
    for _ in range(limit):
        s.tick()
        if s.idle():
            return
        clock.t += dt
    raise AssertionError("scheduler did not drain")


def test_slot_pool():
    p = SlotPool(100)
    a = p.alloc(30)
    b = p.alloc(50)
    assert (a, b) == (0, 30) and p.alloc(30) == -1
    p.release(a, 30)
    assert p.alloc(20) == 0 and p.alloc(10) == 20
    p.release(0, 20)
    p.release(20, 10)
    p.release(30, 50)
    assert p.free == [(0, 100)]
    p.grow(150)
    assert p.alloc(150) == 0


def test_all_jobs_complete_and_memory_returns():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 1000, 1: 2000})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=10, initial_slots=64, clock=clock)
    jobs = [s.submit(spec(i % 2, 50 + i, 25 + 3 * i)) for i in range(6)]
    drive(s, clock)
    assert all(j.state == "done" and j.steps_done == j.spec.nsteps for j in jobs)
    assert all(g.pool.used == 0 for g in s.groups.values())
    # jobs of the same receptor were advanced by shared launches
    assert any(len(ids) > 1 for _, ids, _ in be.launches)


def test_continuous_batching_joins_running_group():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 1000})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=5, initial_slots=1000, clock=clock)
    a = s.submit(spec(0, 1000, 200))
    for _ in range(30):  # let a run a few chunks
        s.tick()
        clock.t += 1e-3
    b = s.submit(spec(0, 100, 10))
    drive(s, clock)
    assert b.t_admit < a.t_done and b.t_done < a.t_done  # b overtook a: no head-of-line blocking
    assert any(set(ids) == {a.id, b.id} for _, ids, _ in be.launches)


def test_admission_waits_for_memory_and_lru_eviction():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 600, 1: 600, 2: 600})
    per_slot = 23 * 4 + 8
    budget = 600 * 2 + 2 * 64 * per_slot
    s = Scheduler(be, hbm_budget=budget, chunk_steps=10, initial_slots=64, clock=clock)
    j0 = s.submit(spec(0, 10, 50))
    j1 = s.submit(spec(1, 10, 50))
    j2 = s.submit(spec(2, 10, 10))   # needs one of the two grids to go
    s.tick()
    assert j0.state == "admitted" and j1.state == "admitted" and j2.state == "pending"
    drive(s, clock)
    assert j2.state == "done" and j2.t_admit >= min(j0.t_done, j1.t_done)
    assert len(be.dropped) >= 1
    assert s.used <= budget


def test_oversized_job_fails_cleanly():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 10 ** 6})
    s = Scheduler(be, hbm_budget=1000, chunk_steps=10, initial_slots=8, clock=clock)
    j = s.submit(spec(0, 10, 10))
    ok = s.submit(spec(0, 1, 1))
    s.tick()
    assert j.state == "failed" and "budget" in j.error
    assert ok.state == "failed"


def test_slot_pool_growth_between_chunks():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 100})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=10, initial_slots=16, clock=clock)
    a = s.submit(spec(0, 10, 30))
    b = s.submit(spec(0, 40, 30))  # does not fit the initial 16 slots
    drive(s, clock)
    assert a.state == b.state == "done"
    assert s.groups[0].pool.capacity >= 50


def test_shortest_remaining_work_first():
    """With room for 100 poses in flight, a short request that arrives while a long one runs
    is advanced alone until it is done; the long one then resumes."""
    clock = FakeClock()
    be = FakeBackend(clock, {0: 1000})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=5, initial_slots=2000, clock=clock, run_poses=100)
    a = s.submit(spec(0, 1000, 200))
    for _ in range(30):
        s.tick()
        clock.t += 1e-3
    b = s.submit(spec(0, 100, 10))
    drive(s, clock)
    assert a.state == b.state == "done" and b.t_done < a.t_done
    alone = [i for i, (_, ids, _) in enumerate(be.launches) if ids == [b.id]]
    assert len(alone) == 2  # 10 steps in chunks of 5, without a
    assert all(ids == [a.id] for _, ids, _ in be.launches[alone[-1] + 1:])
    assert a.steps_done == 200


def test_long_job_is_not_starved():
    """Short jobs arriving back to back do not hold a long one back beyond max_wait."""
    clock = FakeClock()
    be = FakeBackend(clock, {0: 1000})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=5, initial_slots=4000, clock=clock, run_poses=100,
                  max_wait=0.05)
    a = s.submit(spec(0, 1000, 50))
    shorts = []
    for k in range(400):
        if k % 2 == 0:
            shorts.append(s.submit(spec(0, 100, 10)))
        s.tick()
        clock.t += 1e-3
    assert a.steps_done > 0  # advanced while short jobs kept coming
    drive(s, clock)
    assert a.state == "done" and all(j.state == "done" for j in shorts)
