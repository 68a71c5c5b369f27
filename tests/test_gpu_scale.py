"""The full batched-minimization workload (262,144 poses of 64 ligands of 5-60 beads against
the 2000-bead receptor, 200 ADAM steps in one persistent launch) checked through properties
that do not depend on its size: every pose takes its 200 steps and ends finite; minimization
lowers the energy of most poses; the poses of one ligand equal, bit for bit, those of a launch
of that ligand alone; and sampled poses of small, medium and large ligands agree with the
oracle's FP32 twin."""
import numpy as np
import pytest
import torch

from attract_jax_amd.minimize import BatchPlan, PoseState, minimize
from attract_jax_amd.pose import mat_from_qt
from attract_jax_amd.workloads import build_system, c3
from oracle import oracle as O
from oracle.twin import Twin
from tests.systems import pose_rmsd, twin_conditioning

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_full_batched_minimization_properties():
    w = c3()
    system, ligs = build_system(w, DEV, keep_plain=True)
    plan = BatchPlan(system, ligs, w.pose_counts, w.nsteps, w.lr_rot, w.lr_trans)
    st = PoseState.create(w.quat, w.trans, DEV)
    minimize(system, plan, st, w.nsteps)
    torch.cuda.synchronize()
    assert plan.N == 262144
    assert torch.all(st.count == w.nsteps)
    assert torch.isfinite(st.pose).all() and torch.isfinite(st.energy).all()

    off = plan.offsets
    e0 = np.concatenate([system.energrad(mat_from_qt(w.quat[off[k]:off[k + 1]], w.trans[off[k]:off[k + 1]])
                                         .astype(np.float32), ligs[k], grad=False)[0].cpu().numpy()
                         for k in range(len(ligs))])
    e1 = st.energy.cpu().numpy()
    assert (e1 <= e0 + 1e-3 * np.maximum(np.abs(e0), 1.0)).mean() > 0.95

    sizes = np.array([l.nbeads for l in ligs])
    picks = [int(np.nonzero(sizes <= 21)[0][0]), int(np.nonzero((sizes > 21) & (sizes <= 32))[0][0]),
             int(np.nonzero(sizes > 32)[0][0])]
    P = st.pose.cpu().numpy()
    grid = system.grid.to_reference_grid()
    f = system.ff
    ffa = [t.cpu().numpy() for t in (f.rc, f.ac, f.ivor, f.emin, f.rmin2)]
    for k in picks:
        sl = slice(int(off[k]), int(off[k + 1]))
        alone_plan = BatchPlan(system, [ligs[k]], [w.pose_counts[k]], w.nsteps, w.lr_rot, w.lr_trans)
        alone = PoseState.create(w.quat[sl], w.trans[sl], DEV)
        minimize(system, alone_plan, alone, w.nsteps)
        assert torch.equal(alone.pose, st.pose[sl]) and torch.equal(alone.energy, st.energy[sl]), k

        lig = w.ligands[k]
        twin = Twin(grid, ffa, w.coor_rec, w.rec_types - 1, lig.coor, lig.types - 1, lig.types - 1)
        idx = np.arange(int(off[k]), int(off[k]) + 4096, 64)
        Pt, Et, good, rc = twin_conditioning(twin, np.concatenate([w.quat[idx], w.trans[idx]], 1), w.nsteps,
                                             w.lr_rot, w.lr_trans, lig.coor)
        r = pose_rmsd(P[idx], Pt, lig.coor)
        Eg = st.energy.cpu().numpy()[idx]
        erel = np.abs(Eg - Et) / np.maximum(np.abs(Et), 1.0)
        print(f"C3 ligand {k} ({lig.coor.shape[0]} beads): {good.sum()} of {len(idx)} poses well conditioned; "
              f"RMSD vs twin max {r[good].max():.2e} median {np.median(r[good]):.2e} A, energy rel max "
              f"{erel[good].max():.2e}; all: RMSD median {np.median(r):.2e}")
        assert good.mean() >= 0.8, (k, good.mean())
        assert r[good].max() < 0.05, (k, r[good].max())
        tight = good & (r < 1e-4)
        print(f"  energy rel where RMSD < 1e-4 A ({tight.sum()} poses): median {np.median(erel[tight]):.1e}, "
              f"99th pct {np.percentile(erel[tight], 99):.1e}")
        assert tight.mean() >= 0.5, (k, tight.mean())
        assert np.percentile(erel[tight], 99) < 1e-4 and erel[good].max() < 1e-3, (k, erel[good].max())


def test_full_single_shot_scoring_and_adjoint_properties():
    """The 131,072-pose energrad: a sample matches the twin (voxels bit for bit, energies within
    1e-4) and scoring that sample alone gives the same bits.  The 65,536-pose pair-table adjoint:
    a rerun gives the same bits (wavefront tables, static pose order, fixed-order sums), and it
    is linear in the pose weights (g(w1 + w2) = g(w1) + g(w2) to rounding)."""
    from attract_jax_amd.autodiff import param_adjoint
    from attract_jax_amd.workloads import c2, c4
    w = c2()
    system, ligs = build_system(w, DEV, keep_plain=True)
    mat = mat_from_qt(w.quat, w.trans).astype(np.float32)
    assert len(mat) == 131072
    E, G = system.energrad(mat, ligs[0])
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(len(mat), 1024, replace=False))
    Es, Gs = system.energrad(mat[idx], ligs[0])
    assert torch.equal(Es, E[idx]) and torch.equal(Gs, G[idx])
    grid = system.grid.to_reference_grid()
    f = system.ff
    ffa = [t.cpu().numpy() for t in (f.rc, f.ac, f.ivor, f.emin, f.rmin2)]
    lig = w.ligands[0]
    twin = Twin(grid, ffa, w.coor_rec, w.rec_types - 1, lig.coor, lig.types - 1, lig.types - 1)
    Et, Gt = twin.energrad(mat[idx])
    En = Es.cpu().numpy()
    assert (np.abs(En - Et) / np.maximum(np.abs(Et), 1.0)).max() < 1e-4
    vox = system.bead_terms(mat[idx], ligs[0])[1].cpu().numpy()
    assert np.array_equal(vox, twin.bead_terms(mat[idx])["voxel"])

    w4 = c4()
    m4 = torch.from_numpy(mat_from_qt(w4.quat, w4.trans).astype(np.float32)).to(DEV).reshape(-1, 16)
    assert m4.shape[0] == 65536
    w1 = torch.from_numpy(rng.uniform(0.5, 1.5, 65536).astype(np.float32)).to(DEV)
    w2 = torch.from_numpy(rng.uniform(-1.0, 1.0, 65536).astype(np.float32)).to(DEV)
    g1 = param_adjoint(system, ligs[0], m4, system.ff, w1, system.flags)
    g1b = param_adjoint(system, ligs[0], m4, system.ff, w1, system.flags)
    assert g1.abs().max().item() > 0
    assert torch.equal(g1, g1b)
    g2 = param_adjoint(system, ligs[0], m4, system.ff, w2, system.flags)
    g12 = param_adjoint(system, ligs[0], m4, system.ff, w1 + w2, system.flags)
    assert ((g12 - g1 - g2).abs().max() / (g1.abs() + g2.abs()).max()).item() < 1e-5


def test_full_size_grid_build_samples():
    """The C2 grid (96^3 inner, 64^3 outer, 32 types, 2000 receptor beads) built on the GPU:
    sampled grid points' potentials match the FP64 oracle's sum over receptor beads, and their
    neighbour lists hold exactly the beads within neighbourdis (up to FP32 rounding at the edge)."""
    from attract_jax_amd.workloads import c2
    w = c2(npose=8)
    system, _ = build_system(w, DEV, keep_plain=True)
    g = system.grid.to_reference_grid()
    f = system.ff
    ff = O.FF(*[t.cpu().numpy().astype(np.float64) for t in (f.rc, f.ac, f.ivor, f.emin, f.rmin2)])
    rng = np.random.default_rng(11)
    idx = np.stack([rng.integers(0, d, 300) for d in g.dim], 1)
    pts = g.origin + g.gridspacing * idx
    for t in (0, 13, 31):
        ref = O.grid_point_potential(pts, w.coor_rec, w.rec_types - 1, t, ff, g.plateaudis)
        got = g.inner_potential_grid[t][idx[:, 0], idx[:, 1], idx[:, 2]]
        assert (np.abs(got - ref) / np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)).max() < 1e-4, t
    counts, starts, lists = O.csr_lists(g)
    flat = np.ravel_multi_index(idx.T, g.dim)
    nonempty = 0
    for v, p in zip(flat, pts):
        ref = O.grid_point_neighbours(p, w.coor_rec, g.neighbourdis)
        got = np.sort(lists[starts[v]:starts[v] + counts[v]])
        nonempty += len(ref) > 0
        if not np.array_equal(got, ref):
            dd = np.linalg.norm(w.coor_rec[np.setxor1d(got, ref).astype(int)] - p, axis=1)
            assert np.all(np.abs(dd - g.neighbourdis) < 1e-4)
    assert nonempty > 20
