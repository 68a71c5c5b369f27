"""Serving: a continuous-batching request loop for docking jobs on one GPU.

A job = (receptor, ligand, starting poses, ADAM steps).  Jobs arrive at any time;
the scheduler
  * admits them in arrival order while their HBM footprint (pose state, plus the
    receptor grid if it is not resident) fits the budget, evicting idle resident
    grids least-recently-used first;
  * keeps receptor grids resident on the device and gives every resident receptor a
    pool of pose slots; jobs of the same receptor share the pool, so one persistent
    launch (atj_minimize) advances all of them together;
  * runs each receptor group in chunks of `chunk_steps` ADAM steps on its own HIP
    stream, so groups overlap on the device and newly admitted jobs join the next
    chunk of their group (no waiting for the running jobs to finish); with `chunk_work`,
    a launch of few poses takes more steps, up to chunk_work pose-steps, so that a small
    job is not stretched over many short launches;
  * advances, at each chunk boundary, the jobs with the least remaining work first
    (pose-steps times ligand beads), as many as fill the GPU (`run_poses` poses in
    flight): a short request overtakes long ones instead of sharing the device with
    them, and a job that has waited `max_wait` seconds since admission is advanced
    whatever its size;
  * retires a job when its steps are done: its results are copied to the host and
    its slots and memory are freed.
The policy (this file's Scheduler and SlotPool) is independent of the device: the
GPU backend drives the HIP kernels, the fake backend in the tests drives a clock.
"""
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np

POSE_STATE_BYTES = (7 + 7 + 7 + 1 + 1) * 4  # pose, m, v, count, energy per pose slot
PLAN_BYTES = 8                              # pose_job + order entries per slot


@dataclass
class JobSpec:
    """A docking request.  receptor_key names the receptor (grid); the backend knows how to
    build / look up its grid.  quat (n,4) / trans (n,3) are the starting poses."""
    receptor_key: object
    ligand_coor: np.ndarray
    ligand_types: np.ndarray
    quat: np.ndarray
    trans: np.ndarray
    nsteps: int
    lr_rot: float = 1e-3
    lr_trans: float = 1e-3
    arrival: float = 0.0


@dataclass
class Job:
    id: int
    spec: JobSpec
    npose: int
    slot: int = -1
    steps_done: int = 0
    state: str = "pending"     # pending -> admitted -> done | failed
    t_arrival: float = None    # scheduler clock at submission
    t_admit: float = None
    t_done: float = None
    result: dict = None
    error: str = None
    handle: object = None      # backend data (device ligand, ...)

    @property
    def remaining(self):
        return self.spec.nsteps - self.steps_done

    @property
    def remaining_work(self):
        """Pose-steps left, weighted by ligand beads (the cost of a pose-step grows with them)."""
        return self.npose * self.remaining * max(1, len(self.spec.ligand_coor))


class SlotPool:
    """First-fit allocator of contiguous pose-slot ranges in [0, capacity)."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.free = [(0, self.capacity)]  # sorted (start, length)

    def alloc(self, n):
        for i, (s, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, l - n)
                return s
        return -1

    def release(self, start, n):
        self.free.append((start, n))
        self.free.sort()
        merged = []
        for s, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((s, l))
        self.free = merged

    def tail_free(self):
        """Length of the free range that ends at the capacity (0 if the last slot is used)."""
        if self.free and self.free[-1][0] + self.free[-1][1] == self.capacity:
            return self.free[-1][1]
        return 0

    def grow(self, new_capacity):
        if new_capacity <= self.capacity:
            return
        self.release(self.capacity, new_capacity - self.capacity)
        self.capacity = int(new_capacity)

    @property
    def used(self):
        return self.capacity - sum(l for _, l in self.free)


@dataclass
class Group:
    """A resident receptor and the jobs currently running against it."""
    key: object
    grid_bytes: int
    pool: SlotPool
    layout: object = None        # the grid layout the backend chose for it
    jobs: list = field(default_factory=list)
    inflight: object = None      # backend token of the running chunk
    chunk_jobs: list = field(default_factory=list)
    chunk_steps: int = 0
    last_used: float = 0.0
    handle: object = None        # backend data (device grid, slot arrays, stream)
    dirty: bool = True           # job membership changed since the last plan


class Scheduler:
    def __init__(self, backend, hbm_budget, chunk_steps=25, initial_slots=1 << 16, clock=time.perf_counter,
                 run_poses=1 << 16, max_wait=0.5, chunk_work=0):
        self.backend = backend
        self.budget = int(hbm_budget)
        self.chunk_steps = int(chunk_steps)
        self.chunk_work = int(chunk_work)
        self.run_poses = int(run_poses)
        self.max_wait = float(max_wait)
        self.initial_slots = int(initial_slots)
        self.clock = clock
        self.pending = deque()
        self.groups = OrderedDict()   # LRU order: oldest first
        self.jobs = []
        self.done = []
        self.used = 0
        self.evictions = 0

    # ------------------------------------------------------------------ bookkeeping
    def submit(self, spec, now=None):
        job = Job(len(self.jobs), spec, len(spec.quat))
        job.t_arrival = self.clock() if now is None else now
        self.jobs.append(job)
        self.pending.append(job)
        return job

    def _group_bytes(self, capacity):
        return capacity * (POSE_STATE_BYTES + PLAN_BYTES)

    def _drop(self, g):
        self.backend.drop_group(g)
        self.used -= g.grid_bytes + self._group_bytes(g.pool.capacity)
        del self.groups[g.key]
        self.evictions += 1

    def _grid_options(self, key, job):
        """[(layout, grid bytes)] the backend offers for a new group, preferred first."""
        if hasattr(self.backend, "grid_options"):
            return self.backend.grid_options(key, job)
        return [(None, self.backend.grid_bytes(key))]

    def _evict_for(self, need, keep=None):
        """Free idle resident grids (LRU first) until `need` more bytes fit."""
        for key in list(self.groups):
            if self.used + need <= self.budget:
                return True
            g = self.groups[key]
            if g is keep or g.jobs or g.inflight is not None:
                continue
            self._drop(g)
        return self.used + need <= self.budget

    def _fail(self, job, msg):
        job.state, job.error = "failed", msg
        job.t_done = self.clock()
        self.done.append(job)

    def _admit(self, now):
        """Admit pending jobs in arrival order while they fit.  A job that would not fit even
        with the device to itself fails at once; one that fits later waits (and holds back
        the jobs behind it).  A busy group's slot pool doubles, or grows by what the job lacks
        at its end when doubling does not fit; an idle group too small for the job gets a new
        pool of the job's size (its grid stays resident)."""
        while self.pending:
            job = self.pending[0]
            key = job.spec.receptor_key
            g = self.groups.get(key)
            options = self._grid_options(key, job) if g is None else [(g.layout, g.grid_bytes)]
            if min(b for _, b in options) + self._group_bytes(max(self.initial_slots, job.npose)) > self.budget:
                self.pending.popleft()
                self._fail(job, "job does not fit the HBM budget")
                continue
            start = -1
            if g is not None:
                start = g.pool.alloc(job.npose)
                if start < 0 and not g.jobs and g.inflight is None:
                    cap = max(self.initial_slots, job.npose)
                    delta = self._group_bytes(cap) - self._group_bytes(g.pool.capacity)
                    if not self._evict_for(delta, keep=g):
                        return  # wait until running jobs free memory
                    self.backend.reset_group(g, cap)
                    g.pool = SlotPool(cap)
                    self.used += delta
                    start = g.pool.alloc(job.npose)
                elif start < 0:
                    if g.inflight is not None:
                        return  # slot arrays are resized only between chunks
                    lack = job.npose - g.pool.tail_free()
                    extra = max(lack, g.pool.capacity)
                    if not self._evict_for(self._group_bytes(extra), keep=g):
                        extra = lack
                        if not self._evict_for(self._group_bytes(extra), keep=g):
                            return  # wait until running jobs free memory
                    self.backend.grow_group(g, g.pool.capacity + extra)
                    g.pool.grow(g.pool.capacity + extra)
                    self.used += self._group_bytes(extra)
                    start = g.pool.alloc(job.npose)
            if g is None:
                cap = max(self.initial_slots, job.npose)
                for layout, gbytes in options:  # the first layout that fits, evicting idle grids for it
                    need = gbytes + self._group_bytes(cap)
                    if need <= self.budget and self._evict_for(need):
                        break
                else:
                    return  # wait until running jobs free memory
                g = Group(key, gbytes, SlotPool(cap), layout)
                self.backend.make_group(g)
                self.used += need
                self.groups[key] = g
                start = g.pool.alloc(job.npose)
            self.pending.popleft()
            try:
                self.backend.load_job(g, job, start)
            except ValueError as e:  # a ligand the grid or pair table cannot score
                g.pool.release(start, job.npose)
                self._fail(job, str(e))
                continue
            job.slot = start
            job.state = "admitted"
            job.t_admit = now
            g.jobs.append(job)
            g.dirty = True
            self.groups.move_to_end(g.key)

    def _ready(self, g):
        """Whether a group's grid is built (a backend without ready() builds at once).  A group
        whose build failed has its jobs failed and is dropped."""
        ready = getattr(self.backend, "ready", None)
        if ready is None:
            return True
        try:
            return ready(g)
        except Exception as e:  # the build raised: its jobs cannot run
            for j in list(g.jobs):
                g.jobs.remove(j)
                self._fail(j, f"grid build failed: {e}")
            self._drop(g)
            return False

    def _run_set(self, now, groups):
        """Ids of the jobs to advance: jobs waiting longer than max_wait since admission,
        then the rest in order of remaining work, until their poses reach run_poses."""
        runnable = [j for g in groups for j in g.jobs if j.remaining > 0]
        runnable.sort(key=lambda j: (now - j.t_admit <= self.max_wait, j.remaining_work, j.id))
        chosen, poses = set(), 0
        for j in runnable:
            if poses >= self.run_poses:
                break
            chosen.add(j.id)
            poses += j.npose
        return chosen

    def _launch(self, now):
        # groups whose grid is built; jobs of groups still building wait outside the run set
        ready = [g for g in list(self.groups.values()) if g.jobs and self._ready(g)]
        chosen = None
        for g in ready:
            if g.inflight is not None:
                continue
            if chosen is None:
                chosen = self._run_set(now, ready)
            jobs = [j for j in g.jobs if j.id in chosen]
            if not jobs:
                continue
            steps = self.chunk_steps
            if self.chunk_work:  # launches of few poses take more steps, up to chunk_work pose-steps
                steps = max(steps, -(-self.chunk_work // sum(j.npose for j in jobs)))
            steps = min(steps, max(j.remaining for j in jobs))
            dirty = g.dirty or [j.id for j in jobs] != [j.id for j in g.chunk_jobs]
            g.chunk_jobs = jobs
            g.chunk_steps = steps
            g.inflight = self.backend.launch(g, jobs, steps, dirty)
            g.dirty = False
            g.last_used = now
            self.groups.move_to_end(g.key)

    def _reap(self, now):
        for g in list(self.groups.values()):
            if g.inflight is None or not self.backend.finished(g.inflight):
                continue
            g.inflight = None
            finished = []
            for j in g.chunk_jobs:
                j.steps_done += min(g.chunk_steps, j.remaining)
                if j.remaining <= 0:
                    finished.append(j)
            for j in finished:
                j.result = self.backend.fetch_job(g, j)
                j.t_done = self.clock()
                j.state = "done"
                g.jobs.remove(j)
                g.pool.release(j.slot, j.npose)
                self.backend.release_job(g, j)
                self.done.append(j)
                g.dirty = True

    def tick(self, now=None):
        now = self.clock() if now is None else now
        self._reap(now)
        self._admit(now)
        self._launch(now)

    def idle(self):
        return not self.pending and all(not g.jobs and g.inflight is None for g in self.groups.values())

    def latencies(self):
        return np.array([j.t_done - j.t_arrival for j in self.done if j.state == "done"])

    def admission_waits(self):
        return np.array([j.t_admit - j.t_arrival for j in self.done if j.state == "done"])


# ---------------------------------------------------------------------------- GPU backend

class GpuBackend:
    """Drives the HIP kernels: resident grids, per-receptor slot arrays, one stream per group.

    receptors: key -> workloads.ReceptorSpec (the grid is built on the device when a group for
    the receptor is made) or a dict {'grid_bytes', 'build'} whose build()
    returns a ScoringSystem.  The grid layout of a ReceptorSpec is chosen per group: corner-packed
    (8x the plain footprint, about twice the minimizer's speed) for jobs of at least
    corner_min_work pose-steps when it fits, plain otherwise; layout= forces one.

    A grid build never makes the scheduler wait: make_group launches its first half
    (potentials, list counts) on the build stream, and ready() finishes it (the lists, sized by
    the count) once the count has reached the host; a group launches once its grid is ready."""

    def __init__(self, device, receptors, layout=None, waves_per_block=16, corner_min_work=2e6, launch_options=0):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.receptors = receptors
        self.layout = layout
        self.corner_min_work = float(corner_min_work)
        self.waves = waves_per_block
        self.options = launch_options
        # One stream per group, so that a host wait on a group's stream (result copy, plan
        # upload) waits for that group's work only; streams of dropped groups are reused.  Grids
        # are built on one build stream, which no group launches on: the caching allocator keeps
        # memory per stream, so grids built there share one pool of cached memory however many
        # group streams there are.
        self._free_streams = []
        self._build_stream = torch.cuda.Stream(self.device)
        # launches of fewer poses than 8 per wavefront the device holds take their own claim size
        self._small_launch = 8 * waves_per_block * torch.cuda.get_device_properties(self.device).multi_processor_count
        self._built = {}
        self.build_seconds = []       # time from each grid build's start to its grid being ready
        self.layouts = []             # layout of each group made

    def grid_bytes(self, key):
        return self.grid_options(key, None)[0][1]

    def grid_options(self, key, job):
        from . import _lib
        r = self.receptors[key]
        if isinstance(r, dict):
            return [(self.layout, r["grid_bytes"])]
        lists = r.list_bytes()
        corner = (_lib.ATJ_LAYOUT_CORNER, r.corner_bytes() + lists)
        plain = (_lib.ATJ_LAYOUT_PLAIN, r.plain_bytes() + lists)
        if self.layout is not None:
            return [corner if self.layout == _lib.ATJ_LAYOUT_CORNER else plain]
        work = 0 if job is None else job.npose * job.spec.nsteps
        return [corner, plain] if work >= self.corner_min_work else [plain]

    def _build_begin(self, key, layout):
        """Start building the receptor's scoring system on the build stream: a finished
        system, or (receptor, pair table, GridBuild) to finish in _build_finish."""
        torch = self.torch
        system = self._built.get(key)
        if system is not None:
            return system
        r = self.receptors[key]
        with torch.cuda.stream(self._build_stream):
            if isinstance(r, dict):
                return r["build"]()
            from .device import DeviceGrid, ForceField, Receptor
            from .workloads import NTYPES
            ff = ForceField.create(r.rc, r.ac, r.ivor, device=self.device)
            rec = Receptor.create(r.coor, r.types - 1, self.device)
            return rec, ff, DeviceGrid.build_begin(rec, ff, np.arange(NTYPES), layout=layout, device=self.device,
                                                   **r.grid_kw)

    def _build_finish(self, key, pending):
        from .score import ScoringSystem
        with self.torch.cuda.stream(self._build_stream):
            rec, ff, gb = pending
            system = ScoringSystem(gb.finish(), ff, rec)
        self._built[key] = system
        return system

    def make_group(self, g):
        from .minimize import PoseState
        torch = self.torch
        cap = g.pool.capacity
        stream = self._free_streams.pop() if self._free_streams else torch.cuda.Stream(self.device)
        # the group's arrays belong to its stream: the caching allocator then reuses their
        # memory only in that stream's order once they are dropped (grow_group, drop_group)
        z = torch.zeros
        with torch.cuda.stream(stream):
            state = PoseState(z((cap, 7), device=self.device), z((cap, 7), device=self.device),
                              z((cap, 7), device=self.device), z(cap, dtype=torch.int32, device=self.device),
                              z(cap, device=self.device))
            counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        g.handle = dict(system=None, state=state, stream=stream, counter=counter, plan=None,
                        build=self._build_begin(g.key, g.layout), t_build=time.perf_counter())
        self.ready(g)

    def ready(self, g, wait=False):
        """True once the group's grid is built (wait: finish it now, waiting for the device).
        A failed build raises here."""
        from . import _lib
        from .score import ScoringSystem
        h = g.handle
        if h["system"] is None:
            b = h["build"]
            if isinstance(b, ScoringSystem):
                system = b
            else:
                if not wait and not b[2].ready():
                    return False
                system = self._build_finish(g.key, b)
            system.launch = _lib.AtjLaunch(self.waves, 0, self.options)
            done = self.torch.cuda.Event()
            done.record(self._build_stream)
            h["stream"].wait_event(done)  # the group's launches follow the grid's build
            h["system"], h["build"] = system, None
            self.build_seconds.append(time.perf_counter() - h["t_build"])
            self.layouts.append(system.grid.layout)
        return True

    def grow_group(self, g, new_capacity):
        torch = self.torch
        st = g.handle["state"]
        cap = st.N

        def grow(t):
            shape = (new_capacity,) + tuple(t.shape[1:])
            n = torch.zeros(shape, dtype=t.dtype, device=self.device)
            n[:cap] = t
            return n
        from .minimize import PoseState
        with torch.cuda.stream(g.handle["stream"]):
            g.handle["state"] = PoseState(grow(st.pose), grow(st.m), grow(st.v), grow(st.count), grow(st.energy))

    def reset_group(self, g, capacity):
        """Fresh slot arrays of `capacity` slots for an idle group (its grid stays)."""
        from .minimize import PoseState
        torch = self.torch
        z = torch.zeros
        with torch.cuda.stream(g.handle["stream"]):
            g.handle["state"] = PoseState(z((capacity, 7), device=self.device), z((capacity, 7), device=self.device),
                                          z((capacity, 7), device=self.device),
                                          z(capacity, dtype=torch.int32, device=self.device),
                                          z(capacity, device=self.device))
        g.handle["plan"] = None

    def drop_group(self, g):
        """Release an idle group (no launch or copy of it is pending, so its grid is free for
        the next build at once).  The memory stays with the caching allocator for the next
        grids: emptying the cache would wait for every running launch and make the next grid
        a fresh device allocation."""
        try:  # a group whose job failed to load may still be building
            self.ready(g, wait=True)
        except Exception:  # a failed build leaves nothing to keep
            pass
        self._built.pop(g.key, None)
        self._free_streams.append(g.handle["stream"])
        g.handle = None

    def load_job(self, g, job, slot):
        """Device ligand and starting poses of a job in slots [slot, slot + npose).  ValueError
        (nothing loaded) for a ligand without beads or with types beyond the grid or table."""
        from .device import Ligand
        torch = self.torch
        h = g.handle
        sp = job.spec
        lig = Ligand.create(sp.ligand_coor, sp.ligand_types - 1, sp.ligand_types - 1, self.device, stream=h["stream"])
        r = self.receptors[g.key]
        if h["system"] is None and not isinstance(r, dict):  # the grid is still building
            from .workloads import NTYPES
            if lig.max_grid_type is not None and lig.max_grid_type >= NTYPES:
                raise ValueError(f"ligand grid type {lig.max_grid_type} beyond the {NTYPES} potential grids")
            if lig.max_ff_type is not None and lig.max_ff_type >= np.shape(r.rc)[1]:
                raise ValueError(f"ligand type {lig.max_ff_type} beyond the {np.shape(r.rc)[1]} columns of the "
                                 f"pair table")
        else:
            self.ready(g, wait=True)
            lig.check_types(h["system"].grid, h["system"].ff)
        sl = slice(slot, slot + job.npose)
        with torch.cuda.stream(h["stream"]):
            pose = torch.cat([torch.as_tensor(sp.quat, dtype=torch.float32),
                              torch.as_tensor(sp.trans, dtype=torch.float32)], 1).pin_memory()
            h["state"].pose[sl].copy_(pose, non_blocking=True)
            h["state"].m[sl].zero_()
            h["state"].v[sl].zero_()
            h["state"].count[sl].zero_()
        job.handle = dict(ligand=lig, pinned=pose)

    def _plan(self, g, jobs):
        torch = self.torch
        system = g.handle["system"]
        ligs = [j.handle["ligand"] for j in jobs]
        import ctypes
        from . import _lib
        arr = (_lib.AtjJob * len(jobs))()
        for i, j in enumerate(jobs):
            opt = _lib.AtjOptimizer(_lib.ATJ_OPT_ADAM, j.spec.lr_rot, j.spec.lr_trans, 0.9, 0.999, 1e-8, 0.0, 0.0)
            arr[i] = _lib.AtjJob(ligs[i].struct(), j.slot, j.npose, j.spec.nsteps, 0, opt)
        from .device import _upload  # pinned and non-blocking: the launch never waits on the host
        jobs_t = _upload(np.frombuffer(bytes(arr), np.uint8).copy(), self.device)
        cap = g.handle["state"].N
        pose_job = np.zeros(cap, np.int32)
        order = []
        for i, j in sorted(enumerate(jobs), key=lambda ij: -ij[1].handle["ligand"].nbeads):
            pose_job[j.slot:j.slot + j.npose] = i
            order.append(np.arange(j.slot, j.slot + j.npose, dtype=np.int32))
        order = np.concatenate(order)
        return dict(jobs=jobs_t, njobs=len(jobs), pose_job=_upload(pose_job, self.device),
                    order=_upload(order, self.device), N=len(order))

    def launch(self, g, jobs, steps, dirty):
        import ctypes
        from . import _lib
        from .device import _ptr
        torch = self.torch
        h = g.handle
        with torch.cuda.stream(h["stream"]):
            if dirty or h["plan"] is None:
                h["plan"] = self._plan(g, jobs)
            p = h["plan"]
            system = h["system"]
            lib = _lib.load()
            launch = system.launch
            if p["N"] < self._small_launch:
                # fewer than 8 poses per wavefront of the device: the launch chooses its own claim
                # size (down to one pose per claim, so that the poses spread over every CU)
                keep = launch.options & ~(_lib.ATJ_LAUNCH_CLAIM_MASK | _lib.ATJ_LAUNCH_TAIL_MASK |
                                          _lib.ATJ_LAUNCH_TAIL_ROUNDS_MASK)
                launch = _lib.AtjLaunch(launch.waves_per_block, launch.blocks, keep)
            _lib.check(lib.atj_minimize(ctypes.byref(system._gs), ctypes.byref(system._fs),
                                        ctypes.byref(system._rs), _ptr(p["jobs"]), p["njobs"], _ptr(p["pose_job"]),
                                        _ptr(p["order"]), p["N"], ctypes.byref(h["state"].struct()), int(steps),
                                        system.flags, ctypes.byref(launch), _ptr(h["counter"]),
                                        ctypes.c_void_p(0), ctypes.c_void_p(h["stream"].cuda_stream)))
            ev = torch.cuda.Event()
            ev.record(h["stream"])
        return ev

    def finished(self, token):
        return token.query()

    def fetch_job(self, g, job):
        h = g.handle
        sl = slice(job.slot, job.slot + job.npose)
        with self.torch.cuda.stream(h["stream"]):
            pose = h["state"].pose[sl].cpu()
            energy = h["state"].energy[sl].cpu()
        return dict(pose=pose.numpy(), energy=energy.numpy())

    def release_job(self, g, job):
        job.handle = None


def run_stream(scheduler, specs, poll=0.0001, timeout=600.0, stall=60.0):
    """Submit specs at their arrival times (seconds from now) and run until all are done.
    Returns (latencies (s), makespan (s), pose-steps).  Raises RuntimeError when jobs wait
    or run but none advances for `stall` seconds, TimeoutError after `timeout`."""
    t0 = scheduler.clock()
    order = sorted(specs, key=lambda s: s.arrival)
    k = 0
    mark, t_mark = None, t0
    while True:
        now = scheduler.clock()
        while k < len(order) and now - t0 >= order[k].arrival:
            sp = order[k]
            scheduler.submit(sp, now=t0 + sp.arrival)
            k += 1
        scheduler.tick(now)
        if k == len(order) and scheduler.idle():
            break
        progress = (len(scheduler.done), sum(j.steps_done for j in scheduler.jobs))
        if progress != mark or scheduler.idle():
            mark, t_mark = progress, now
        elif now - t_mark > stall:
            waiting = [j.id for j in scheduler.pending]
            raise RuntimeError(f"no job advanced for {stall:.0f} s (pending jobs {waiting[:8]}, "
                               f"{len(scheduler.jobs) - len(scheduler.done)} unfinished)")
        if now - t0 > timeout:
            raise TimeoutError("serving run exceeded its time limit")
        time.sleep(poll)
    makespan = scheduler.clock() - t0
    pose_steps = sum(j.npose * j.spec.nsteps for j in scheduler.done if j.state == "done")
    return scheduler.latencies(), makespan, pose_steps
