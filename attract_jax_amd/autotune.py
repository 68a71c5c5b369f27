"""Launch-shape autotuner of the persistent pose kernels.

Times a short persistent minimization for each candidate -- grid layout (corner-packed or
plain), waves per block, receptor in LDS or read through L1, poses per work claim and a tail of
smaller claims -- and keeps the fastest.  This replaces the reference's hand-set batch knobs (test3.py:7-12 BATCH /
NB_CHUNK, test4.py:99-103 nb_chunk_thresholds), which sized XLA chunks for the CPU.  Results
are cached per process by (receptor size, grid dims, ligand-size histogram bucket,
candidate set)."""
import time

import numpy as np
import torch

from . import _lib
from .minimize import BatchPlan, PoseState, minimize

_L1 = _lib.ATJ_LAUNCH_REC_GLOBAL
SHAPES = ((4, 0), (8, 0), (16, 0), (16, _L1), (8, _L1))
# claim options: the launch's own choice (8 poses per claim for large launches), claims of 4 or
# 2, and claims of 8 whose last one or two rounds are claimed 4 at a time
CLAIMS = (0, _lib.ATJ_LAUNCH_CLAIM(4), _lib.ATJ_LAUNCH_CLAIM(2), _lib.ATJ_LAUNCH_CLAIM(8) | _lib.ATJ_LAUNCH_TAIL(4, 1),
          _lib.ATJ_LAUNCH_CLAIM(8) | _lib.ATJ_LAUNCH_TAIL(4, 2))
_cache = {}


def autotune_minimize(systems, plan, w, device, shapes=SHAPES, claims=CLAIMS, steps=25, sample=131072, log=None):
    """systems: {layout: ScoringSystem} (the same receptor and lists in each layout).
    Returns (layout, waves_per_block, launch options).  Steady-state timing: `steps` ADAM steps
    (at most the plan's) so that per-step work, not per-pose set-up, decides; `sample` poses
    spread over all ligands so that each wavefront gets about 32 of them at 16 waves per CU."""
    steps = max(1, min(steps, plan.nsteps))
    any_sys = next(iter(systems.values()))
    nb = sorted(l.nbeads for l in plan.ligands)
    key = (any_sys.receptor.nrec, tuple(any_sys.grid.dim), nb[len(nb) // 2] // 8, tuple(sorted(systems)),
           tuple(shapes), tuple(claims))
    if key in _cache:
        return _cache[key]
    per = max(1, sample // plan.njobs)
    counts = [min(per, c) for c in plan.pose_counts]
    sub = BatchPlan(any_sys, plan.ligands, counts, steps, 1e-3, 1e-3)
    idx = np.concatenate([np.arange(plan.offsets[k], plan.offsets[k] + counts[k]) for k in range(plan.njobs)])
    q = torch.as_tensor(w.quat[idx], dtype=torch.float32)
    t = torch.as_tensor(w.trans[idx], dtype=torch.float32)
    best, best_t = None, float("inf")
    for layout, system in systems.items():
        old = system.launch
        for wpb, opt in shapes:
            for claim in claims:
                o = opt | claim
                system.launch = _lib.AtjLaunch(wpb, 0, o)
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
                    log(f"[autotune] layout {'corner' if layout == _lib.ATJ_LAYOUT_CORNER else 'plain'} waves/block "
                        f"{wpb} options {o:#x}: {tt * 1e3:.2f} ms for {sub.N} poses x {steps} steps")
                if tt < best_t:
                    best, best_t = (layout, wpb, o), tt
        system.launch = old
    _cache[key] = best
    return best
