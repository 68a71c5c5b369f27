"""Launch-shape autotuner of the persistent pose kernels.

Times a short persistent minimization for each candidate (waves per block, receptor in
LDS or read through L1) and keeps the fastest; results are cached per process by
(receptor size, grid dims, ligand-size histogram bucket, candidate set)."""
import time

import torch

from . import _lib
from .minimize import BatchPlan, PoseState, minimize

_L1 = _lib.ATJ_LAUNCH_REC_GLOBAL
CANDIDATES = ((4, 0), (8, 0), (16, 0), (16, _L1), (8, _L1))
_cache = {}


def autotune_minimize(system, plan, w, device, candidates=CANDIDATES, steps=25, sample=131072, log=None):
    """Steady-state timing: `steps` ADAM steps (at most the plan's) so that per-step work,
    not per-pose set-up, decides; `sample` poses so that each wavefront gets about 32 of
    them at 16 waves per CU and the persistent kernel's tail stays small."""
    steps = max(1, min(steps, plan.nsteps))
    nb = sorted(l.nbeads for l in plan.ligands)
    key = (system.receptor.nrec, tuple(system.grid.dim), nb[len(nb) // 2] // 8, tuple(candidates))
    if key in _cache:
        return _cache[key]
    # a pose sample spread over all ligands
    per = max(1, sample // plan.njobs)
    counts = [min(per, c) for c in plan.pose_counts]
    sub = BatchPlan(system, plan.ligands, counts, steps, 1e-3, 1e-3)
    import numpy as np
    idx = np.concatenate([np.arange(plan.offsets[k], plan.offsets[k] + counts[k]) for k in range(plan.njobs)])
    q = torch.as_tensor(w.quat[idx], dtype=torch.float32)
    t = torch.as_tensor(w.trans[idx], dtype=torch.float32)
    old = system.launch
    best, best_t = None, float("inf")
    for wpb, opt in candidates:
        system.launch = _lib.AtjLaunch(wpb, 0, opt)
        times = []
        for rep in range(3):
            st = PoseState.create(q, t, device)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            minimize(system, sub, st, steps)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        tt = min(times[1:])
        if log:
            log(f"[autotune] waves/block {wpb} options {opt}: {tt * 1e3:.2f} ms for {sub.N} poses x {steps} steps")
        if tt < best_t:
            best, best_t = (wpb, opt), tt
    system.launch = old
    _cache[key] = best
    return best
