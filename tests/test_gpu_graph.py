"""The per-job form captured in a HIP graph (PlainGraph): every replay gives the bits of the eager
per-step launches and of the persistent minimizer, with the program's own eager work (an energrad
launch, an optimizer launch on other arrays, a reduction that allocates) and synchronizations
between the replays."""
import pytest
import torch

from attract_jax_amd.minimize import BatchPlan, PlainGraph, PoseState, minimize, minimize_plain
from attract_jax_amd.pose import mat_from_qt
from attract_jax_amd.workloads import build_system, c1, c3

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _reset(state, init):
    state.pose.copy_(init)
    state.m.zero_()
    state.v.zero_()
    state.count.zero_()
    state.energy.zero_()


@pytest.mark.parametrize("work", ["c1", "c3"])
def test_graph_replays_equal_eager_steps(work):
    if work == "c1":
        w = c1(npose=512)
        nsteps, k = 30, 0
    else:  # a 48-bead ligand: poses take whole-wave bead phases
        w = c3(nligands=4, poses_per_ligand=300, nsteps=25)
        nsteps = 25
        k = max(range(len(w.ligands)), key=lambda i: len(w.ligands[i].coor))
    system, ligs = build_system(w, DEV)
    n = 512 if work == "c1" else 300
    sl = slice(0, n) if work == "c1" else slice(k * n, (k + 1) * n)
    print(f"graph test {work}: ligand of {ligs[k].nbeads} beads, {n} poses, {nsteps} steps", flush=True)
    eager = PoseState.create(w.quat[sl], w.trans[sl], DEV)
    minimize_plain(system, ligs[k], eager, nsteps, w.lr_rot, w.lr_trans)
    # the persistent minimizer gives the same bits (the graph is checked against both)
    plan = BatchPlan(system, [ligs[k]], [n], nsteps, w.lr_rot, w.lr_trans)
    batched = PoseState.create(w.quat[sl], w.trans[sl], DEV)
    minimize(system, plan, batched, nsteps)
    torch.cuda.synchronize()
    assert torch.equal(eager.pose, batched.pose) and torch.equal(eager.energy, batched.energy)

    state = PoseState.create(w.quat[sl], w.trans[sl], DEV)
    init = state.pose.clone()
    g = PlainGraph(system, ligs[k], state, nsteps, w.lr_rot, w.lr_trans)
    # the capture itself ran nothing on the state
    assert torch.equal(state.pose, init) and int(state.count.max()) == 0
    other = PoseState.create(w.quat[sl], w.trans[sl], DEV)
    mat = torch.tensor(mat_from_qt(w.quat[sl], w.trans[sl]), dtype=torch.float32, device=DEV)
    for rep in range(4):
        _reset(state, init)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(state.pose, eager.pose), f"replay {rep}: poses differ"
        assert torch.equal(state.energy, eager.energy), f"replay {rep}: energies differ"
        assert torch.equal(state.m, eager.m) and torch.equal(state.v, eager.v)
        assert int(state.count.min()) == nsteps and int(state.count.max()) == nsteps
        # eager work between replays: energrad, a plain step on other arrays, a multi-block sum
        E, G = system.energrad(mat, ligs[k])
        minimize_plain(system, ligs[k], other, 1, w.lr_rot, w.lr_trans)
        assert torch.isfinite(E.sum()) and torch.isfinite(G.abs().sum())
        torch.cuda.synchronize()
