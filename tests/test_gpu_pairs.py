"""Pair records at the branches of ATTRACT's pair energy (test-grid-energrad.py:95-111):
E = vlj + (ivor - 1) emin for dsq < rmin2, ivor vlj beyond.  Points are placed just inside
and just outside rmin of an attractive (ivor = +1) and a repulsive (ivor = -1) type pair, and
near a repulsive pair with ac = 0 (rmin2 = infinity, emin = 0, so the inner branch always
holds).  They are scored through atj_point_terms, and poses carrying those ligand types
through atj_energrad, against the FP64 oracle on the same neighbour lists.  Swapping the
branches of a repulsive pair changes its energy by 2 vlj (about 2 |emin| at rmin), far above
the 1e-4 relative tolerance."""
import numpy as np
import pytest
import torch

from attract_jax_amd.pose import mat_from_qt
from attract_jax_amd.score import point_terms
from attract_jax_amd.workloads import build_system, c1
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"

# (receptor type, ligand type, ivor, ac = 0), type indices 0-based
CASES = [(0, 1, -1.0, False), (2, 3, -1.0, True), (4, 5, 1.0, False)]


@pytest.fixture(scope="module")
def branch_system():
    w = c1(npose=256)
    for a, b, iv, zero_ac in CASES:
        for i, j in ((a, b), (b, a)):  # the table stays symmetric
            w.ivor[i, j] = iv
            if zero_ac:
                w.ac[i, j] = 0.0
    lig = w.ligands[0]
    lig.types = np.array([CASES[0][1], CASES[1][1], CASES[2][1], lig.types[3] - 1]) + 1
    with np.errstate(divide="ignore"):
        system, ligs = build_system(w, DEV, keep_plain=True)
    grid = system.grid.to_reference_grid()
    assert grid.inner_potential_grid is not None  # the oracle's grid term and plateau need them
    f = system.ff
    ff = O.FF(*[t.cpu().numpy().astype(np.float64) for t in (f.rc, f.ac, f.ivor, f.emin, f.rmin2)])
    return w, system, ligs, grid, ff


def _points(w, grid, ff, rng):
    """Points around receptor beads of each case's receptor type, at distances straddling rmin
    (or spread over the plateau for the ac = 0 pair), inside the inner grid and at least 2.5 A
    from every other receptor bead."""
    xyz, lt = [], []
    lo = grid.origin + grid.gridspacing
    hi = grid.origin + grid.gridspacing * (np.asarray(grid.dim) - 2)
    for a, b, _, zero_ac in CASES:
        beads = np.nonzero(w.rec_types - 1 == a)[0][:40]
        assert len(beads) > 0
        r = np.sqrt(ff.rmin2[a, b]) if not zero_ac else 3.0
        fac = [0.97, 0.995, 0.9995, 1.0005, 1.005, 1.03] if not zero_ac else [0.9, 1.1, 1.3]
        for k in beads:
            for f in np.repeat(fac, 4):  # four random directions per distance
                d = rng.normal(size=3)
                p = w.coor_rec[k] + f * r * d / np.linalg.norm(d)
                others = np.delete(np.linalg.norm(w.coor_rec - p, axis=1), k)
                if f * r < grid.plateaudis and others.min() > 2.5 and np.all(p > lo) and np.all(p < hi):
                    xyz.append(p)
                    lt.append(b)
    return np.array(xyz, np.float32), np.array(lt)


def test_point_terms_at_the_branches(branch_system):
    w, system, ligs, grid, ff = branch_system
    xyz, lt = _points(w, grid, ff, np.random.default_rng(21))
    assert len(xyz) >= 40
    for b in {c[1] for c in CASES}:
        assert (lt == b).sum() >= 8
    _, _, nbe, nbg = point_terms(xyz, np.zeros(len(xyz), np.int64), lt, system.grid, ff=system.ff,
                                 receptor=system.receptor, terms=("nb",))
    Eref, Gref, npair = O.nb_energrad(xyz[None].astype(np.float64), lt, w.coor_rec, w.rec_types - 1, ff, grid)
    assert npair >= len(xyz)
    E, G = nbe.cpu().numpy(), nbg.cpu().numpy()
    assert (np.abs(E - Eref[0]) / np.maximum(np.abs(Eref[0]), 1.0)).max() < 1e-4
    gscale = np.maximum(np.abs(Gref[0]).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(G - Gref[0]) / gscale).max() < 1e-4
    # the points do tell the branches apart: with every ivor read as +1 the oracle differs
    bad = O.FF(ff.rc, ff.ac, np.ones_like(ff.ivor), ff.emin, ff.rmin2)
    Ebad, _, _ = O.nb_energrad(xyz[None].astype(np.float64), lt, w.coor_rec, w.rec_types - 1, bad, grid)
    assert (np.abs(E - Ebad[0]) / np.maximum(np.abs(Ebad[0]), 1.0)).max() > 1e-2


def test_energrad_with_the_branch_types(branch_system):
    w, system, ligs, grid, ff = branch_system
    mat = mat_from_qt(w.quat, w.trans)
    mat[:, 3, :3] *= 0.55  # pull the poses into the receptor's plateau range
    mat32 = mat.astype(np.float32)
    E, G = system.energrad(mat32, ligs[0])
    l = w.ligands[0]
    Eref, Gref = O.energrad(mat32.astype(np.float64), l.coor, l.types - 1, l.types - 1, w.coor_rec,
                            w.rec_types - 1, ff, grid)
    E, G = E.cpu().numpy(), G.cpu().numpy().reshape(-1, 4, 4)
    assert np.count_nonzero(E) > len(E) // 2
    assert (np.abs(E - Eref) / np.maximum(np.abs(Eref), 1.0)).max() < 1e-4
    gscale = np.maximum(np.abs(Gref).max(axis=(1, 2), keepdims=True), 1.0)
    assert (np.abs(G - Gref.reshape(-1, 4, 4)) / gscale).max() < 1e-4
