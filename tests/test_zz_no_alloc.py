"""The allocation-free minimizer entry point (atj_minimize_noalloc, work counter in the
library's scratch array) against atj_minimize with a caller workspace, through the C-ABI:
bit-identical poses, moments, step counts and energies on two calls in a row, the second
with fewer poses, uneven ligands of several jobs in one launch."""
import ctypes

import numpy as np
import pytest
import torch

from attract_jax_amd import _lib
from attract_jax_amd.device import _ptr
from attract_jax_amd.minimize import BatchPlan, PoseState, minimize
from attract_jax_amd.workloads import build_system, c3

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _noalloc(system, plan, state, nsteps):
    lib = _lib.load()
    _lib.check(lib.atj_minimize_noalloc(ctypes.byref(system._gs), ctypes.byref(system._fs), ctypes.byref(system._rs),
                                        _ptr(plan.jobs), plan.njobs, _ptr(plan.pose_job), _ptr(plan.order), plan.N,
                                        ctypes.byref(state.struct()), int(nsteps), system.flags,
                                        ctypes.byref(system.launch), ctypes.c_void_p(0),
                                        ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)))


def test_noalloc_minimize_equals_workspace_form():
    w = c3(nligands=6, poses_per_ligand=96, nsteps=15)
    system, ligs = build_system(w, DEV)
    for npl in (96, 40):  # second call with fewer poses: the scratch keeps the last call's counter
        counts = [npl] * len(ligs)
        sl = np.concatenate([np.arange(k * 96, k * 96 + npl) for k in range(len(ligs))])
        plan = BatchPlan(system, ligs, counts, 15, w.lr_rot, w.lr_trans)
        a = PoseState.create(w.quat[sl], w.trans[sl], DEV)
        b = PoseState.create(w.quat[sl], w.trans[sl], DEV)
        minimize(system, plan, a, 15)
        _noalloc(system, plan, b, 15)
        torch.cuda.synchronize()
        for x, y in zip((a.pose, a.m, a.v, a.count, a.energy), (b.pose, b.m, b.v, b.count, b.energy)):
            assert torch.equal(x, y)
        assert int(b.count.min()) == 15 and torch.isfinite(b.energy).all()
