"""Batched rigid-body minimization (test4.py:430-455, test5.py:446-458) on the device.

`minimize` packs the poses of any number of ligands (uneven bead counts) into ONE
persistent launch that runs every step (atj_minimize).  `minimize_plain` is the
per-job comparison form: one ligand at a time, one energrad launch plus one
optimizer launch per step.  `PlainGraph` captures the per-job form's steps once in a
HIP graph and replays them, so the per-step launches cost no host work.

Optimizer semantics, per pose and per parameter (quaternion x, y, z, w and
translation): optax.adam defaults (test4.py:446-453): m = (1-b1) g + b1 m,
v = (1-b2) g^2 + b2 v; optax.adabelief (test5.py:451-457): the second moment is
taken of g - m_prev and eps_root is added to it; both update -lr m_hat /
(sqrt(v_hat) + eps) with bias-corrected moments, and the quaternion is renormalised
after every step (Rotation.from_quat(as_quat())).  gravity > 0 adds gravity |t|^2 to
the energy (test5.py:418-420).  The reference uses one learning rate; lr_rot /
lr_trans may differ.  The energy reported per pose is the one evaluated in the last
step, before its update (test4.py:450,455).  Zeroing count (and the moments)
re-initialises the optimizer, as test5.py:452 does at every stage.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import _ptr, _stream


@dataclass
class PoseState:
    pose: torch.Tensor    # (N, 7): quaternion (x, y, z, w), translation
    m: torch.Tensor       # (N, 7) ADAM first moment
    v: torch.Tensor       # (N, 7) ADAM second moment
    count: torch.Tensor   # (N,) int32 steps done
    energy: torch.Tensor  # (N,) energy of the last evaluated step

    @classmethod
    def create(cls, quat, trans, device):
        f = lambda a: (a.detach().to(device=device, dtype=torch.float32) if isinstance(a, torch.Tensor)
                       else torch.as_tensor(np.asarray(a, np.float32), device=device))
        pose = torch.cat([f(quat).reshape(-1, 4), f(trans).reshape(-1, 3)], dim=1).contiguous()
        N = pose.shape[0]
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device=device)
        return cls(pose, z(N, 7), z(N, 7), torch.zeros(N, dtype=torch.int32, device=device), z(N))

    def struct(self):
        return _lib.AtjPoseState(self.pose.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                 self.count.data_ptr(), self.energy.data_ptr())

    @property
    def quat(self):
        return self.pose[:, :4]

    @property
    def trans(self):
        return self.pose[:, 4:]

    @property
    def N(self):
        return self.pose.shape[0]


class BatchPlan:
    """Pose slots of several ligands packed into one launch: job table, per-pose job
    index and a processing order (largest ligands first, so the tail of the
    persistent launch is made of short poses)."""

    def __init__(self, system, ligands, pose_counts, nsteps, lr_rot, lr_trans, beta1=0.9, beta2=0.999,
                 eps=1e-8, *, optimizer=_lib.ATJ_OPT_ADAM, eps_root=0.0, gravity=0.0):
        self.system = system
        self.ligands = list(ligands)
        self.pose_counts = [int(n) for n in pose_counts]
        d = system.device
        self.set_optimizer(nsteps, lr_rot, lr_trans, beta1, beta2, eps, optimizer=optimizer, eps_root=eps_root,
                           gravity=gravity)
        self.njobs = len(self.ligands)
        job_of = np.repeat(np.arange(self.njobs, dtype=np.int32), self.pose_counts)
        self.pose_job = torch.from_numpy(job_of).to(d)
        nb = np.array([l.nbeads for l in self.ligands])
        order = np.argsort(-nb[job_of], kind="stable").astype(np.int32)
        self.order = torch.from_numpy(order).to(d)
        self.N = int(sum(self.pose_counts))
        self.offsets = np.concatenate(([0], np.cumsum(self.pose_counts)))

    def set_optimizer(self, nsteps, lr_rot, lr_trans, beta1=0.9, beta2=0.999, eps=1e-8, *,
                      optimizer=_lib.ATJ_OPT_ADAM, eps_root=0.0, gravity=0.0):
        """(Re)write the job table: step counts and optimizer settings (per job or shared)."""
        self.nsteps = int(np.max(nsteps))  # per-job step counts may differ
        self.jobs = self.system.jobs_tensor(self.ligands, nsteps, lr_rot, lr_trans, beta1, beta2, eps,
                                            self.pose_counts, optimizer=optimizer, eps_root=eps_root,
                                            gravity=gravity)


def minimize(system, plan, state, nsteps_launch, flags=None, stats=False, stream=None):
    """Run up to nsteps_launch ADAM steps of every pose of `plan` in one launch."""
    lib = _lib.load()
    fl = (system.flags if flags is None else flags) & ~_lib.ATJ_POSE_QT
    st = stream if stream is not None else _stream(system.device)
    _lib.check(lib.atj_minimize(ctypes.byref(system._gs), ctypes.byref(system._fs), ctypes.byref(system._rs),
                                _ptr(plan.jobs), plan.njobs, _ptr(plan.pose_job), _ptr(plan.order), plan.N,
                                ctypes.byref(state.struct()), int(nsteps_launch), fl, ctypes.byref(system.launch),
                                _ptr(system.counter), _ptr(system.stats) if stats else ctypes.c_void_p(0), st))
    return state


def minimize_plain(system, ligand, state, nsteps, lr_rot, lr_trans, beta1=0.9, beta2=0.999, eps=1e-8,
                   flags=None, stream=None, jobs=None, *, optimizer=_lib.ATJ_OPT_ADAM, eps_root=0.0, gravity=0.0,
                   launch=None):
    """Per-job form: for each step, one energrad launch (quaternion poses) then one
    optimizer launch; pose state lives in HBM between launches.  The energrad launches use
    `launch`, by default the system's waves per block and receptor path with the claim size
    left to each launch: claim options tuned for a persistent minimization (many poses per
    wavefront) would leave most of the device idle in these one-step launches."""
    lib = _lib.load()
    d = system.device
    st = stream if stream is not None else _stream(d)
    N = state.N
    if jobs is None:
        jobs = system.jobs_tensor([ligand])
    dEdM = torch.empty((N, 16), dtype=torch.float32, device=d)
    fl = ((system.flags if flags is None else flags) | _lib.ATJ_POSE_QT)
    if launch is None:
        sl = system.launch
        launch = _lib.AtjLaunch(sl.waves_per_block, sl.blocks, sl.options & _lib.ATJ_LAUNCH_REC_GLOBAL)
    ss = state.struct()
    opt = _lib.AtjOptimizer(int(optimizer), float(lr_rot), float(lr_trans), float(beta1), float(beta2), float(eps),
                            float(eps_root), float(gravity))
    for _ in range(int(nsteps)):
        _lib.check(lib.atj_energrad(ctypes.byref(system._gs), ctypes.byref(system._fs), ctypes.byref(system._rs),
                                    _ptr(jobs), 1, ctypes.c_void_p(0), ctypes.c_void_p(0), _ptr(state.pose), N,
                                    _ptr(state.energy), _ptr(dEdM), fl, ctypes.byref(launch),
                                    _ptr(system.counter), ctypes.c_void_p(0), st))
        _lib.check(lib.atj_optimizer_step(_ptr(dEdM), _ptr(state.energy), N, ctypes.byref(ss), ctypes.byref(opt), st))
    return state


class PlainGraph:
    """The per-job form's launches (one energrad and one optimizer launch per step, as
    minimize_plain issues them) captured once in a HIP graph (torch.cuda.CUDAGraph over the
    C-ABI calls) and replayed on the current stream.  A replay runs the `nsteps` steps on the
    arrays of `state` as they are at the time of the replay: reset them (state.pose.copy_(...),
    m, v, count zeroed) to start again.  The captured launches are those minimize_plain makes,
    so a replay gives the same bits as minimize_plain from the same state.  Every launch resets
    the system's work counter with a fill kernel; nothing in the graph is a memset or a copy.
    The graph holds the system's counter: replays of graphs of one system must not overlap."""

    def __init__(self, system, ligand, state, nsteps, lr_rot, lr_trans, beta1=0.9, beta2=0.999, eps=1e-8,
                 flags=None, jobs=None, *, optimizer=_lib.ATJ_OPT_ADAM, eps_root=0.0, gravity=0.0, launch=None):
        d = system.device
        if jobs is None:
            jobs = system.jobs_tensor([ligand])  # uploaded before the capture: no copy in the graph
        # the graph's launches hold device pointers of the system (grid, receptor, pair table,
        # counter), the ligand, the job table and the state: they are kept alive with it
        self.system, self.ligand, self.jobs, self.state, self.nsteps = system, ligand, jobs, state, int(nsteps)
        kw = dict(flags=flags, jobs=jobs, optimizer=optimizer, eps_root=eps_root, gravity=gravity, launch=launch)
        args = (lr_rot, lr_trans, beta1, beta2, eps)
        # one step on copies of the state first, so that every kernel is loaded and every launch
        # attribute set before the capture
        tmp = PoseState(state.pose.clone(), state.m.clone(), state.v.clone(), state.count.clone(),
                        state.energy.clone())
        minimize_plain(system, ligand, tmp, 1, *args, **kw)
        torch.cuda.synchronize(d)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, capture_error_mode="relaxed"):
            minimize_plain(system, ligand, state, nsteps, *args, **kw)
        torch.cuda.synchronize(d)

    def replay(self):
        self.graph.replay()
        return self.state
