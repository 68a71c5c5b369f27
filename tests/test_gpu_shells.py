"""Shell-ordered neighbour lists (DeviceGrid.shell_lists, atj_grid_shell_lists): every grid
point's list reordered into distance shells around the point, and a bead reading only the
leading entries its shell can reach within plateaudis.  The reordering is checked against
a NumPy restatement of the grouping (float64 distances; entries within 1e-4 A of a shell
boundary may fall either side), and scoring / minimization against the whole-list path:
the same pairs contribute (pair counts identical, bead voxels and grid terms bit-equal),
energies agree to the rounding of a different summation order, and both match the
oracle's CPU twin."""
import numpy as np
import pytest
import torch

from attract_jax_amd.minimize import BatchPlan, PoseState, minimize
from attract_jax_amd.score import point_terms
from attract_jax_amd.workloads import build_system, c1
from oracle import oracle as O

from .test_gpu_parity import DEV, Sys, _grel, _rel

pytestmark = pytest.mark.gpu
MARGIN = 1e-3


@pytest.fixture(scope="module")
def s1():
    return Sys(c1(), nshell=0)  # whole lists in builder order


@pytest.fixture(scope="module", params=[4, 6])
def shelled(request, s1):
    system, ligs = build_system(s1.w, DEV, nshell=request.param)
    return request.param, system, ligs


def _groups(d, r0, h, n):
    return np.where(d < r0, 0, np.minimum(n + 1, np.ceil((d - r0) / h))).astype(np.int64)


def test_shell_order_matches_numpy(s1, shelled):
    n, system, _ = shelled
    g = system.grid
    assert g.nshell == n
    c0, st0, fl0 = s1.system.grid.host_lists()
    c1_, st1, fl1 = g.host_lists()
    assert np.array_equal(c0, c1_) and np.array_equal(st0, st1)
    assert np.array_equal(fl0[:32], fl1[:32])
    counts = c0.reshape(-1)
    e = g.nbshell.cpu().numpy().astype(np.int64)
    assert np.array_equal(e[:, 0], st0)
    packed = e[:, 1:] & 0xFFFFFFFF
    nbshell = np.stack([(packed[:, k // 2] >> (16 * (k % 2))) & 0xFFFF for k in range(n)], 1)
    h = g.spacing * np.sqrt(3.0) / (2 * n)
    r0 = g.plateaudis + MARGIN
    rec = s1.w.coor_rec.astype(np.float64)
    rng = np.random.default_rng(3)
    pts = np.nonzero(counts > 0)[0]
    for v in rng.choice(pts, 500, replace=False):
        p = g.origin.astype(np.float64) + g.spacing * np.array(np.unravel_index(v, g.dim))
        a = fl0[st0[v]:st0[v] + counts[v]].astype(np.int64)
        b = fl1[st0[v]:st0[v] + counts[v]].astype(np.int64)
        assert np.array_equal(np.sort(a), np.sort(b)), v  # a permutation of the point's list
        d = np.linalg.norm(rec[b] - p, axis=1)
        grp = _groups(d, r0, h, n)
        edge = np.zeros(len(d), bool)
        for j in range(n + 1):  # distances within 1e-4 of a group boundary
            edge |= np.abs(d - (r0 + j * h)) < 1e-4
        want = np.sort(grp[~edge], kind="stable")
        assert np.array_equal(grp[~edge], want), v  # groups in increasing order
        for j in np.unique(grp[~edge]):  # receptor order within a group
            ids = b[(grp == j) & ~edge]
            assert np.all(np.diff(ids) > 0), (v, j)
        for k in range(n):
            cnt = nbshell[v, k]
            assert cnt <= counts[v]
            need = d < g.plateaudis + (k + 1) * h - 1e-4
            assert np.all(np.nonzero(need)[0] < cnt), (v, k)  # every reachable entry is in the prefix
            assert np.all(grp[cnt:][~edge[cnt:]] > k + 1), (v, k)
        assert np.all(np.diff(nbshell[v]) >= 0)


def test_scoring_same_pairs_as_whole_lists(s1, shelled):
    n, system, ligs = shelled
    mat = s1.mats()
    s1.system.stats.zero_()
    system.stats.zero_()
    E1, G1 = s1.system.energrad(mat, s1.ligs[0], stats=True)
    E2, G2 = system.energrad(mat, ligs[0], stats=True)
    st1, st2 = s1.system.stats.cpu().numpy(), system.stats.cpu().numpy()
    assert st1[0] == st2[0] and st1[1] == st2[1] and st1[3] == st2[3]  # grid beads, inner beads, pairs
    assert st2[2] < st1[2]  # fewer list entries read
    E1, G1, E2, G2 = (t.cpu().numpy() for t in (E1, G1, E2, G2))
    assert _rel(E2, E1).max() < 1e-5
    assert _grel(G2, G1).max() < 1e-5
    Et, Gt = s1.twins[0].energrad(mat)
    assert _rel(E2, Et).max() < 1e-4 and _grel(G2, Gt).max() < 1e-4
    grid4a, voxa, nbea, _ = s1.system.bead_terms(mat, s1.ligs[0])
    grid4b, voxb, nbeb, _ = system.bead_terms(mat, ligs[0])
    assert torch.equal(voxa, voxb) and torch.equal(grid4a, grid4b)
    assert _rel(nbeb.cpu().numpy(), nbea.cpu().numpy()).max() < 1e-5


def test_points_at_cell_centres_and_corners(s1, shelled):
    """Beads at grid points (shell 0) and at the far corners of their cells (last shell)."""
    n, system, _ = shelled
    g = system.grid
    rng = np.random.default_rng(5)
    counts = g.host_lists()[0]
    idx = np.argwhere(counts[1:-1, 1:-1, 1:-1] > 0)[rng.choice(int((counts[1:-1, 1:-1, 1:-1] > 0).sum()), 200)] + 1
    base = g.origin.astype(np.float64) + g.spacing * idx
    sgn = rng.choice([-1.0, 1.0], size=idx.shape)
    corner = base + sgn * (0.5 * g.spacing - 1e-4)
    xyz = np.concatenate([base, corner, base + rng.uniform(-0.5, 0.5, idx.shape) * g.spacing]).astype(np.float32)
    gt = rng.integers(0, 32, len(xyz))
    lt = rng.integers(0, 32, len(xyz))
    a = point_terms(xyz, gt, lt, s1.system.grid, ff=s1.system.ff, receptor=s1.system.receptor)
    b = point_terms(xyz, gt, lt, g, ff=system.ff, receptor=system.receptor)
    assert torch.equal(a[1], b[1]) and torch.equal(a[0], b[0])
    assert _rel(b[2].cpu().numpy(), a[2].cpu().numpy()).max() < 1e-5
    ga, gb = a[3].cpu().numpy(), b[3].cpu().numpy()
    assert (np.abs(gb - ga) / np.maximum(np.abs(ga).max(axis=1, keepdims=True), 1.0)).max() < 1e-5


def test_minimizer_vs_twin(s1, shelled):
    n, system, ligs = shelled
    w = s1.w
    plan = BatchPlan(system, ligs, [w.npose], w.nsteps, w.lr_rot, w.lr_trans)
    st = PoseState.create(w.quat, w.trans, DEV)
    minimize(system, plan, st, w.nsteps)
    P = st.pose.cpu().numpy()
    Pt, Et = s1.twins[0].minimize(np.concatenate([w.quat, w.trans], 1), w.nsteps, w.lr_rot, w.lr_trans)
    lig = w.ligands[0]
    xg = O.transform(O.mat_from_qt(P[:, :4].astype(np.float64), P[:, 4:].astype(np.float64)), lig.coor)
    xt = O.transform(O.mat_from_qt(Pt[:, :4].astype(np.float64), Pt[:, 4:].astype(np.float64)), lig.coor)
    r = O.rmsd(xg, xt)
    assert np.median(r) < 0.01 and (r < 0.05).mean() >= 0.95


def test_whole_lists_again(s1):
    system, ligs = build_system(s1.w, DEV, nshell=6)
    system.grid.shell_lists(system.receptor, 0)
    system.refresh()
    assert system.grid.nshell == 0 and system.grid.nbshell is None
    mat = s1.mats(slice(0, 256))
    s1.system.stats.zero_()
    system.stats.zero_()
    E1, G1 = s1.system.energrad(mat, s1.ligs[0], stats=True)
    E2, G2 = system.energrad(mat, ligs[0], stats=True)
    assert np.array_equal(s1.system.stats.cpu().numpy(), system.stats.cpu().numpy())
    assert _rel(E2.cpu().numpy(), E1.cpu().numpy()).max() < 1e-5


def test_shells_of_other_receptor_coordinates_not_used(s1):
    """A system on a grid whose shells were computed for other receptor coordinates reads
    whole lists (the shell order is still a valid list order)."""
    from attract_jax_amd.device import Receptor
    from attract_jax_amd.score import ScoringSystem
    system, ligs = build_system(s1.w, DEV, nshell=6)
    moved = Receptor(system.receptor.beads.clone(), system.receptor.nrec)
    moved.beads[:, :3] += 0.3
    other = ScoringSystem(system.grid, system.ff, moved)
    assert other._gs.nshell == 0 and system._gs.nshell == 6
    ref = ScoringSystem(s1.system.grid, s1.system.ff, moved, nshell=0)
    mat = s1.mats(slice(0, 256))
    other.stats.zero_()
    ref.stats.zero_()
    E1, G1 = ref.energrad(mat, s1.ligs[0], stats=True)
    E2, G2 = other.energrad(mat, ligs[0], stats=True)
    assert np.array_equal(ref.stats.cpu().numpy(), other.stats.cpu().numpy())
    assert _rel(E2.cpu().numpy(), E1.cpu().numpy()).max() < 1e-5
