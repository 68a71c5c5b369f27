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

    def reset_group(self, g, cap):
        g.handle["reset"] = cap

    def drop_group(self, g):
        self.dropped.append(g.key)

    def load_job(self, g, job, slot):
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


def test_idle_group_too_small_is_made_again_at_the_job_size():
    """Grid 100 B, 100 B per slot, 16 initial slots, budget 5000 B: after a 10-pose job the
    idle 16-slot group cannot grow by 40 (5700 B) but a fresh 40-slot pool fits (4100 B); the
    grid stays resident."""
    clock = FakeClock()
    be = FakeBackend(clock, {0: 100})
    s = Scheduler(be, hbm_budget=5000, chunk_steps=10, initial_slots=16, clock=clock)
    a = s.submit(spec(0, 10, 20))
    drive(s, clock)
    b = s.submit(spec(0, 40, 20))
    drive(s, clock, limit=1000)
    assert a.state == b.state == "done"
    assert s.groups[0].pool.capacity == 40 and s.used == 4100
    assert be.dropped == [] and s.groups[0].handle["reset"] == 40


def test_pool_grows_by_what_the_job_lacks_when_doubling_does_not_fit():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 100})
    s = Scheduler(be, hbm_budget=3200, chunk_steps=10, initial_slots=16, clock=clock)
    a = s.submit(spec(0, 10, 200))
    s.tick()
    b = s.submit(spec(0, 20, 20))  # 6 free at the end: doubling (32) would not fit, +14 does
    for _ in range(5):
        clock.t += 1.0
        s.tick()
    assert b.state in ("admitted", "done") and s.groups[0].pool.capacity == 30
    drive(s, clock)
    assert a.state == b.state == "done"


def test_pool_doubles_when_it_fits():
    clock = FakeClock()
    be = FakeBackend(clock, {0: 100})
    s = Scheduler(be, hbm_budget=10 ** 6, chunk_steps=10, initial_slots=16, clock=clock)
    a = s.submit(spec(0, 10, 200))
    s.tick()
    b = s.submit(spec(0, 20, 20))  # 6 slots free at the end: doubles to 32 slots
    for _ in range(5):
        clock.t += 1.0
        s.tick()
    assert b.state in ("admitted", "done") and s.groups[0].pool.capacity == 32
    drive(s, clock)
    assert a.state == b.state == "done"


class TickClock:
    """A clock that advances by dt at every reading (run_stream reads it in its loop)."""

    def __init__(self, dt=1e-4):
        self.t, self.dt = 0.0, dt

    def __call__(self):
        self.t += self.dt
        return self.t


def test_run_stream_twice_over_the_same_specs():
    """A spec's arrival stays relative to the start of the stream: a second run over the same
    list submits and completes every job again (the scheduler records its own clock)."""
    from attract_jax_amd.serve import run_stream
    specs = [spec(i % 2, 20 + i, 15, arrival=0.002 * i) for i in range(6)]
    before = [sp.arrival for sp in specs]
    for _ in range(2):
        clock = TickClock()
        be = FakeBackend(clock, {0: 1000, 1: 1000})
        s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=5, initial_slots=64, clock=clock)
        lat, makespan, ps = run_stream(s, specs, poll=0.0, timeout=50.0)
        assert len(lat) == len(specs) and all(j.state == "done" for j in s.jobs)
        assert ps == sum(len(sp.quat) * sp.nsteps for sp in specs)
    assert [sp.arrival for sp in specs] == before


def test_run_stream_raises_when_nothing_advances():
    from attract_jax_amd.serve import run_stream

    class Stuck(FakeBackend):
        def finished(self, token):
            return False

    clock = TickClock(dt=0.01)
    s = Scheduler(Stuck(clock, {0: 1000}), hbm_budget=10 ** 9, chunk_steps=5, initial_slots=64, clock=clock)
    import pytest
    with pytest.raises(RuntimeError, match="no job advanced"):
        run_stream(s, [spec(0, 10, 10)], poll=0.0, timeout=100.0, stall=1.0)


def test_job_the_backend_cannot_load_fails_without_taking_slots():
    class Picky(FakeBackend):
        def load_job(self, g, job, slot):
            if len(job.spec.ligand_coor) == 0:
                raise ValueError("a ligand needs at least one bead")
            job.handle = {}

    clock = FakeClock()
    s = Scheduler(Picky(clock, {0: 100}), hbm_budget=10 ** 6, chunk_steps=10, initial_slots=16, clock=clock)
    bad = s.submit(JobSpec(0, np.zeros((0, 3)), np.ones(0, int), np.zeros((5, 4)), np.zeros((5, 3)), 10))
    good = s.submit(spec(0, 16, 10))
    drive(s, clock)
    assert bad.state == "failed" and "bead" in bad.error
    assert good.state == "done" and s.groups[0].pool.capacity == 16


class TwoLayouts(FakeBackend):
    """Offers a large preferred layout and a small fallback for every receptor."""

    def grid_options(self, key, job):
        return [("big", 4000), ("small", 500)]

    def make_group(self, g):
        g.handle = {"layout": g.layout}


def test_layout_falls_back_when_the_preferred_one_does_not_fit():
    clock = FakeClock()
    be = TwoLayouts(clock, {})
    s = Scheduler(be, hbm_budget=5000 + 64 * 100 * 2, chunk_steps=10, initial_slots=64, clock=clock)
    a = s.submit(spec(0, 10, 200))
    s.tick()
    b = s.submit(spec(1, 10, 10))  # receptor 0 holds 4000 B while it runs: 1 gets the small layout
    s.tick()
    assert s.groups[0].layout == "big" and s.groups[1].layout == "small"
    drive(s, clock)
    assert a.state == b.state == "done"


def test_evictions_and_admission_waits_are_counted():
    clock = FakeClock()
    be = FakeBackend(clock, {k: 1000 for k in range(4)})
    per = 64 * 100
    s = Scheduler(be, hbm_budget=2 * (1000 + per), chunk_steps=10, initial_slots=64, clock=clock)
    jobs = [s.submit(spec(k, 10, 30)) for k in range(4)]
    drive(s, clock)
    assert all(j.state == "done" for j in jobs)
    assert s.evictions >= 2
    waits = s.admission_waits()
    assert len(waits) == 4 and waits.max() > 0


class BuildingBackend(FakeBackend):
    """Grids take `build_ticks` calls of ready() to build; keys in `broken` fail their build."""

    def __init__(self, clock, grid_bytes, build_ticks=5, broken=()):
        super().__init__(clock, grid_bytes)
        self.build_ticks = build_ticks
        self.broken = set(broken)

    def make_group(self, g):
        g.handle = {"left": self.build_ticks}

    def ready(self, g):
        if g.key in self.broken:
            raise RuntimeError("no memory for the grid")
        g.handle["left"] -= 1
        return g.handle["left"] < 0


def test_groups_launch_once_their_grid_is_built():
    """A group launches only after its grid is ready, and meanwhile another receptor's jobs
    run; jobs of a receptor whose build fails are failed with the reason, the group is
    dropped and later jobs run."""
    clock = FakeClock()
    be = BuildingBackend(clock, {"a": 100, "b": 100, "bad": 100}, build_ticks=5, broken={"bad"})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=10, initial_slots=16, clock=clock)
    s.submit(spec("a", 8, 30))
    s.tick()
    for _ in range(3):
        clock.t += 1e-3
        s.tick()
    assert be.launches == []  # a is still building
    s.submit(spec("bad", 8, 30))
    s.submit(spec("b", 8, 30))
    drive(s, clock)
    failed = [j for j in s.done if j.state == "failed"]
    assert [j.spec.receptor_key for j in failed] == ["bad"] and "grid build failed" in failed[0].error
    assert "bad" in be.dropped and "bad" not in s.groups
    assert sorted(j.spec.receptor_key for j in s.done if j.state == "done") == ["a", "b"]
    assert be.launches[0][0] == "a"


def test_chunk_work_runs_small_launches_longer():
    """With chunk_work, a launch of few poses takes more steps (up to chunk_work pose-steps, at
    least chunk_steps, at most what its jobs have left); a launch of many poses keeps chunk_steps."""
    clock = FakeClock()
    be = FakeBackend(clock, {"r": 100, "s": 100})
    s = Scheduler(be, hbm_budget=10 ** 9, chunk_steps=10, initial_slots=16, clock=clock, chunk_work=2000)
    s.submit(spec("r", 20, 150))    # 20 poses: 100 steps per launch (2000 / 20), then the last 50
    s.submit(spec("s", 400, 30))    # 400 poses: 2000 / 400 = 5 < chunk_steps, so 10 steps per launch
    drive(s, clock)
    steps_r = [st for key, _, st in be.launches if key == "r"]
    steps_s = [st for key, _, st in be.launches if key == "s"]
    assert steps_r == [100, 50]
    assert steps_s == [10, 10, 10]
    assert all(j.state == "done" and j.steps_done == j.spec.nsteps for j in s.done)
