"""The test5.py protocol on the GPU: adabelief + gravity in the fused minimizer and in the
plain per-step form (bit-identical), the staged schedule against the oracle's CPU twin
stage by stage, conformer libraries (a pose's result does not depend on the other
conformers' poses) and dock_test5 end to end from raw inputs (type-99 beads, 1-based
conformer indices, neighbour lists in original receptor indices)."""
import numpy as np
import pytest
import torch

from attract_jax_amd import _lib
from attract_jax_amd import workloads as W
from attract_jax_amd.device import DeviceGrid, ForceField, Ligand, Receptor
from attract_jax_amd.minimize import BatchPlan, PoseState, minimize, minimize_plain
from attract_jax_amd.pose import mat_from_qt
from attract_jax_amd.protocol import ConformerSet, dock_test5, prepare, run_schedule
from oracle import oracle as O
from oracle.twin import Twin

from .systems import _nudged, pose_rmsd, twin_conditioning
from .test_gpu_parity import Sys

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
ADAB = dict(optimizer=_lib.ATJ_OPT_ADABELIEF, eps_root=1e-16, gravity=0.1)
SHORT = ((4, 0.5, True), (6, 0.2, True), (8, 0.05, False))


@pytest.fixture(scope="module")
def s1():
    return Sys(W.c1())


def _rmsd_to_twin(lig_coor, P, Pt):
    xg = O.transform(O.mat_from_qt(P[:, :4].astype(np.float64), P[:, 4:].astype(np.float64)), lig_coor)
    xt = O.transform(O.mat_from_qt(Pt[:, :4].astype(np.float64), Pt[:, 4:].astype(np.float64)), lig_coor)
    return O.rmsd(xg, xt)


def _twin_schedule(tw, pose7, schedule, gravity=0.1):
    P = np.array(pose7, np.float32)
    E = None
    for epochs, lr, grav in schedule:
        P, E = tw.minimize(P, epochs, lr, lr, eps=1e-16, kind=1, eps_root=1e-16, gravity=gravity if grav else 0.0)
    return P, E


def test_adabelief_gravity_vs_twin_and_plain(s1):
    w = s1.w
    n, nst, lr = 300, 20, 0.05
    plan = BatchPlan(s1.system, s1.ligs, [n], nst, lr, lr, eps=1e-16, **ADAB)
    a = PoseState.create(w.quat[:n], w.trans[:n], DEV)
    minimize(s1.system, plan, a, nst)
    b = PoseState.create(w.quat[:n], w.trans[:n], DEV)
    minimize_plain(s1.system, s1.ligs[0], b, nst, lr, lr, eps=1e-16, **ADAB)
    assert torch.equal(a.pose, b.pose) and torch.equal(a.energy, b.energy)
    Pt, Et, good, _ = twin_conditioning(s1.twins[0], np.concatenate([w.quat[:n], w.trans[:n]], 1), nst, lr, lr,
                                        w.ligands[0].coor, eps=1e-16, kind=1, eps_root=1e-16, gravity=0.1)
    r = _rmsd_to_twin(w.ligands[0].coor, a.pose.cpu().numpy(), Pt)
    E = a.energy.cpu().numpy()
    erel = np.abs(E - Et) / np.maximum(np.abs(Et), 1.0)
    # every well-conditioned pose (the twin from starts 4 ulps away ends within 0.005 A)
    assert good.mean() >= 0.9
    tight = good & (r < 1e-4)
    assert r[good].max() < 0.05 and tight.mean() >= 0.5
    assert np.percentile(erel[tight], 99) < 1e-4 and erel[good].max() < 1e-3
    assert np.median(r) < 0.01
    # the reported energy contains the restraint: far poses (no receptor terms) give 0.1 |t|^2
    far = np.concatenate([w.quat[:2], np.array([[300.0, 0, 0], [0, -400.0, 0]])], 1)
    c = PoseState.create(far[:, :4], far[:, 4:], DEV)
    minimize(s1.system, BatchPlan(s1.system, s1.ligs, [2], 1, lr, lr, eps=1e-16, **ADAB), c, 1)
    assert np.allclose(c.energy.cpu().numpy(), [0.1 * 300.0 ** 2, 0.1 * 400.0 ** 2], rtol=1e-6)


def test_schedule_vs_twin_stage_by_stage(s1):
    w = s1.w
    n = 256
    plan = BatchPlan(s1.system, s1.ligs, [n], 1, 0.1, 0.1)
    st = PoseState.create(w.quat[:n], w.trans[:n], DEV)
    run_schedule(s1.system, plan, st, SHORT)
    assert np.all(st.count.cpu().numpy() == SHORT[-1][0])
    start = np.concatenate([w.quat[:n], w.trans[:n]], 1).astype(np.float32)
    Pt, Et = _twin_schedule(s1.twins[0], start, SHORT)
    good = np.ones(n, bool)
    for cols, direction in ((slice(4, 7), np.inf), (slice(4, 7), -np.inf), (slice(0, 4), np.inf),
                            (slice(0, 4), -np.inf)):
        Pm, _ = _twin_schedule(s1.twins[0], _nudged(start, cols, 4, direction), SHORT)
        good &= pose_rmsd(Pt, Pm, w.ligands[0].coor) < 0.005
    r = _rmsd_to_twin(w.ligands[0].coor, st.pose.cpu().numpy(), Pt)
    E = st.energy.cpu().numpy()
    erel = np.abs(E - Et) / np.maximum(np.abs(Et), 1.0)
    tight = good & (r < 1e-4)
    assert good.mean() >= 0.9 and tight.mean() >= 0.5
    assert r[good].max() < 0.05
    assert np.percentile(erel[tight], 99) < 1e-4 and erel[good].max() < 1e-3
    assert np.median(r) < 0.01


def test_conformer_library_jobs(s1):
    w = s1.w
    rng = np.random.default_rng(17)
    lig = w.ligands[0]
    confs = np.stack([lig.coor + rng.normal(scale=0.3, size=lig.coor.shape) for _ in range(3)])
    cs = ConformerSet(confs, lig.types - 1, lig.types - 1, DEV)
    n = 90
    ci = rng.integers(0, 3, n)
    plan, perm = cs.plan(s1.system, ci, 12, 0.01, 0.1)
    st = PoseState.create(w.quat[:n][perm], w.trans[:n][perm], DEV)
    minimize(s1.system, plan, st, 12)
    P = np.empty((n, 7), np.float32)
    P[perm] = st.pose.cpu().numpy()
    for c in range(3):
        sel = np.nonzero(ci == c)[0]
        alone = PoseState.create(w.quat[sel], w.trans[sel], DEV)
        lc = Ligand.create(confs[c], lig.types - 1, lig.types - 1, DEV)
        minimize(s1.system, BatchPlan(s1.system, [lc], [len(sel)], 12, 0.01, 0.1), alone, 12)
        assert np.array_equal(alone.pose.cpu().numpy(), P[sel]), c


def test_dock_test5_end_to_end():
    rng = np.random.default_rng(2024)
    nrec = 400
    radius = 25.0 * (nrec / 2000.0) ** (1 / 3)  # the bead density of the C2 receptor
    coor = W.random_receptor(nrec, radius, 2.25, rng)
    types = rng.integers(1, 33, nrec)
    dummy = rng.random(nrec) < 0.08
    types[dummy] = 99  # dummy beads, dropped by the protocol
    rc32, ac32, ivor32, _, _ = W.random_pair_table(rng)
    par = [np.ones((99, 99)), np.ones((99, 99)), np.ones((99, 99))]
    for a, b in zip(par, (rc32, ac32, ivor32)):
        a[:32, :32] = b
    lig_types = rng.integers(1, 33, 9)
    lig0 = W.random_chain(9, rng)
    lib = np.stack([lig0 + rng.normal(scale=0.25, size=lig0.shape) for _ in range(2)])
    # the grid, built for the filtered receptor, with neighbour lists in original indices
    alphabet = np.zeros(99, bool)
    alphabet[np.unique(lig_types) - 1] = True
    P = prepare(coor, types, lib, lig_types, par, alphabet)
    ff = ForceField.create(*P.ff, device=DEV)
    rec = Receptor.create(P.coor_rec, P.rec_atomtypes, DEV)
    d = int(np.ceil(2 * (radius + 10.0) / 0.9)) + 1
    o = -0.9 * (d - 1) / 2
    dg = DeviceGrid.build(rec, ff, np.arange(ff.nlt), origin=(o, o, o), spacing=0.9, dim=(d, d, d),
                          dim2=(32, 32, 32), gridextension=16, plateaudis=6.0, neighbourdis=6.8, keep_plain=True,
                          device=DEV)
    g = dg.to_reference_grid(alphabet=alphabet)
    old_index = np.nonzero(~dummy)[0]
    g = g._replace(neighbours=old_index[np.asarray(g.neighbours)].astype(np.uint32))
    nq = 120
    q, t = W.posemod.random_poses(nq, rng, radius, radius + 6)
    mats = mat_from_qt(q, t, fill=1.0)
    conf = rng.integers(1, 3, nq)  # 1-based, as the --conformers file
    poses, E = dock_test5(mats, coor, types, lib, lig_types, g, par, conformers=conf, schedule=SHORT, device=DEV)
    assert poses.shape == (nq, 4, 4) and np.all(poses[:, :, 3] == 1.0) and np.all(np.isfinite(E))
    # twin on the prepared system, stage by stage, per conformer
    gref = dg.to_reference_grid(alphabet=alphabet)
    ffa = [np.asarray(a, np.float32) for a in P.ff]
    for c in (1, 2):
        sel = np.nonzero(conf == c)[0]
        tw = Twin(gref, ffa, P.coor_rec, P.rec_atomtypes, P.coor_lig[c - 1], P.lig_atomtype_pos, P.lig_atomtypes)
        Pt, Et = _twin_schedule(tw, np.concatenate([q[sel], t[sel]], 1), SHORT)
        R = poses[sel, :3, :3]
        xg = np.einsum("nj,pjk->pnk", P.coor_lig[c - 1].astype(np.float64), R) + poses[sel, 3, None, :3]
        xt = O.transform(O.mat_from_qt(Pt[:, :4].astype(np.float64), Pt[:, 4:].astype(np.float64)),
                         P.coor_lig[c - 1])
        r = O.rmsd(xg, xt)
        assert np.median(r) < 0.01 and (r < 0.05).mean() >= 0.9, c
