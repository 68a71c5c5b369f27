"""The oracle against the reference's own printed results on its mono-leu system (CPU):
test-nb-energrad.out, test-nb-energrad-plateau.out, test-rmsd.lrmsd, and ATTRACT's
mono-leu/minim-sorted-nbgrid.score; plus the grid box these depend on."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.mono_leu import NEIGHBOURDIS, SPACING, MonoLeu, grad_rel, rel_to_scale


@pytest.fixture(scope="module")
def ml():
    return MonoLeu()


def test_inferred_box_reproduces_the_printed_list_statistics(ml):
    """Lists within neighbourdis of every grid point of the inferred box: the printed total
    (test-grid-energrad-retile.out:3) to one entry, the printed longest list exactly (:2)."""
    from scipy.spatial import cKDTree
    assert tuple(ml.dim) == (89, 74, 59)
    ii = np.stack(np.meshgrid(*[np.arange(d) for d in ml.dim], indexing="ij"), -1).reshape(-1, 3)
    c = cKDTree(ml.coor).query_ball_point(ml.origin + SPACING * ii, NEIGHBOURDIS, return_length=True)
    assert int(c.max()) == int(ml.z["list_max"])
    assert abs(int(c.sum()) - int(ml.z["list_total"])) <= 1
    # a box one voxel smaller or larger on every side misses the total by far more
    for grow in (-1, 1):
        c2 = cKDTree(ml.coor).query_ball_point(ml.origin - grow * SPACING + SPACING * np.stack(
            np.meshgrid(*[np.arange(d + 2 * grow) for d in ml.dim], indexing="ij"), -1).reshape(-1, 3),
            NEIGHBOURDIS, return_length=True)
        assert abs(int(c2.sum()) - int(ml.z["list_total"])) > 100


@pytest.mark.parametrize("plateau", [False, True])
def test_oracle_reproduces_test_nb_energrad_out(ml, plateau):
    """Energies, -dE/dM translation rows and torque blocks of the ten printed poses, within
    1e-4 relative (test-nb-energrad.py:253-261; plateau variant test-nb-energrad-plateau.py)."""
    E_ref, T_ref, Q_ref = ml.jax(plateau)
    E, G, S = ml.oracle(ml.poses[:10], plateau)
    assert (np.abs(E - E_ref) / np.maximum(np.abs(E_ref), 1.0)).max() < 1e-4
    assert grad_rel(-G[:, 3, :3], T_ref).max() < 1e-5
    assert grad_rel(-G[:, :3, :3], Q_ref).max() < 1e-5
    # pose 6 has a bead beyond the box's upper z bound: with it the energy is off by 4.2 (0.63 with
    # the plateau subtracted)
    E_all, _, _ = O.nb_energrad_complete(ml.poses[6:7], ml.coor, ml.types, ml.coor, ml.types, ml.ff,
                                         ml.origin - 10.0, SPACING, ml.dim + 40, 5.0, plateau=plateau)
    assert abs(E_all[0] - E_ref[6]) > 0.5
    assert rel_to_scale(E, E_ref, S).max() < 2e-5


def test_oracle_reproduces_test_rmsd_lrmsd(ml):
    """test-rmsd.py:13-19 over the fixture's poses: every value at its printed 3 decimals."""
    lig, ref = ml.backbone()
    r = O.rmsd(O.transform(ml.poses, lig), ref[None])
    assert np.abs(np.round(r, 3) - ml.z["lrmsd"]).max() == 0.0


def test_oracle_against_attract_nbgrid_scores(ml):
    """ATTRACT's FP64 scores of the same structures (mono-leu/minim-sorted-nbgrid.score, first
    1000 .dat entries, mono-leu/dock.sh:22-27).  ATTRACT looks a bead up while its nearest grid
    point is in the grid, half a voxel further than the reference's `< dim - 1` test: with that
    box every energy agrees within 2e-5 of the pair-term scale and every translational gradient
    within 1e-4; with the reference's box about 7% of the structures differ."""
    M = ml.dat_mats()
    E_att, G_att = ml.z["att_E"], ml.z["att_G"]
    E, G, S = O.nb_energrad_complete(M, ml.coor, ml.types, ml.coor, ml.types, ml.ff, ml.origin - SPACING / 2,
                                     SPACING, ml.dim + 1, 5.0)
    assert rel_to_scale(E, E_att, S).max() < 2e-5
    assert grad_rel(-G[:, 3, :3], G_att[:, 3:]).max() < 1e-4
    E_ref_box, _, _ = O.nb_energrad_complete(M, ml.coor, ml.types, ml.coor, ml.types, ml.ff, ml.origin, SPACING,
                                             ml.dim, 5.0)
    off = rel_to_scale(E_ref_box, E_att, S) > 1e-3
    assert 0.03 < off.mean() < 0.15
