"""The HIP path against the reference's own printed results on its mono-leu system, through
the C-ABI: test-nb-energrad.out and test-nb-energrad-plateau.out (neighbour-list energrad,
test-nb-energrad.py / test-nb-energrad-plateau.py), test-rmsd.lrmsd (test-rmsd.py), and
ATTRACT's mono-leu/minim-sorted-nbgrid.score.  The grid is built on the device with the box
ATTRACT chose (tests/mono_leu.py) from the receptor beads, as make-grid did (mono-leu/dock.sh:4)."""
import numpy as np
import pytest
import torch

from attract_jax_amd import _lib, score, structio
from attract_jax_amd.device import DeviceGrid, ForceField, Ligand, Receptor
from tests.mono_leu import NEIGHBOURDIS, PLATEAUDIS, SPACING, MonoLeu, grad_rel, rel_to_scale

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


class Dev:
    def __init__(self, ml, grow=0):
        """Device grid of the inferred box (grow = 1: half a voxel further on every side, the
        lookup ATTRACT itself performs)."""
        self.ml = ml
        self.rec = Receptor.create(ml.coor, ml.types, DEV)
        self.lig = Ligand.create(ml.coor, np.zeros(ml.nbead, np.int64), ml.types, DEV)
        self.ff = ForceField.create(ml.ff.rc, ml.ff.ac, ml.ff.ivor, device=DEV)
        origin = ml.origin - grow * SPACING / 2
        dim = ml.dim + grow
        self.grid = DeviceGrid.build(self.rec, self.ff, [0], origin=origin, spacing=SPACING, dim=dim,
                                     plateaudis=PLATEAUDIS, neighbourdis=NEIGHBOURDIS, potentials=False, device=DEV)
        torch.cuda.synchronize()
        self.system = {}

    def energrad(self, mat, plateau):
        if plateau not in self.system:
            flags = _lib.ATJ_TERM_NB | (_lib.ATJ_NB_PLATEAU if plateau else 0)
            self.system[plateau] = score.ScoringSystem(self.grid, self.ff, self.rec, flags=flags)
        E, G = self.system[plateau].energrad(np.asarray(mat, np.float32), self.lig)
        torch.cuda.synchronize()
        return E.cpu().numpy().astype(np.float64), G.cpu().numpy().astype(np.float64)


@pytest.fixture(scope="module")
def ml():
    return MonoLeu()


@pytest.fixture(scope="module")
def dev(ml):
    assert torch.cuda.is_available()
    return Dev(ml)


def test_device_lists_match_the_printed_statistics(ml, dev):
    counts, _, _ = dev.grid.host_lists()
    assert int(counts.max()) == int(ml.z["list_max"])
    assert abs(int(counts.sum()) - int(ml.z["list_total"])) <= 2


@pytest.mark.parametrize("plateau", [False, True])
def test_energrad_matches_test_nb_energrad_out(ml, dev, plateau):
    """The ten printed poses: energies within 1e-4 of the printed ones relative to the scale
    of their pair terms (max(|E|, sum |pair E|)), -dE/dM translation rows and torque blocks
    within 1e-4 of their largest component; and against the FP64 oracle the same."""
    E_ref, T_ref, Q_ref = ml.jax(plateau)
    E, G = dev.energrad(ml.poses[:10], plateau)
    Eo, Go, S = ml.oracle(ml.poses[:10], plateau)
    print("energy rel (scale)", rel_to_scale(E, E_ref, S).max(), "plain", (np.abs(E - E_ref) / np.abs(E_ref)).max())
    assert rel_to_scale(E, E_ref, S).max() < 1e-4
    assert grad_rel(-G[:, 3, :3], T_ref).max() < 1e-4
    assert grad_rel(-G[:, :3, :3], Q_ref).max() < 1e-4
    assert rel_to_scale(E, Eo, S).max() < 1e-4
    assert grad_rel(G, Go).max() < 1e-4


def test_reference_named_value_and_grad_on_mono_leu(ml, dev):
    """value_and_grad_main (test-nb-energrad.py:234-251) with the reference's arguments and a
    reference-format Grid (CSR lists) of the device grid."""
    grid = dev.grid.to_reference_grid()
    (tot, E), G = score.value_and_grad_main(ml.poses[:10], ml.coor, ml.types, ml.coor, ml.types,
                                            np.zeros(ml.nbead, np.int64), ml.ff, grid)
    E = E.cpu().numpy()
    G = G.cpu().numpy()
    E_ref, T_ref, Q_ref = ml.jax(False)
    _, _, S = ml.oracle(ml.poses[:10], False)
    assert rel_to_scale(E, E_ref, S).max() < 1e-4
    assert grad_rel(-G[:, 3, :3], T_ref).max() < 1e-4
    assert grad_rel(-G[:, :3, :3], Q_ref).max() < 1e-4
    assert abs(float(tot) - E.sum()) <= 1e-4 * np.abs(E).sum()


def test_ligand_rmsd_matches_test_rmsd_lrmsd(ml):
    """atj_ligand_rmsd over the fixture's 2000 poses vs test-rmsd.lrmsd (3 decimals): every
    value within half a unit of the last printed digit plus FP32 rounding, and rounded to 3
    decimals equal for all but values that sit on a rounding boundary."""
    lig, ref = ml.backbone()
    r = structio.ligand_rmsd(ml.poses, lig, ref, device=DEV).cpu().numpy().astype(np.float64)
    lr = ml.z["lrmsd"]
    assert np.abs(r - lr).max() <= 5e-4 + 2e-6 * np.abs(lr).max()
    assert (np.round(r, 3) == lr).mean() >= 0.995


def test_energies_against_attract_nbgrid_scores(ml, dev):
    """1000 structures of mono-leu/minim-sorted.dat through dat_to_mat: with the reference's
    lookup box the HIP energies equal the FP64 oracle's (1e-4 of the pair-term scale); with the
    box ATTRACT looks beads up in (half a voxel further) they equal ATTRACT's own FP64 scores,
    and the translational gradients ATTRACT's, within 1e-4."""
    M = ml.dat_mats()
    E, G = dev.energrad(M, False)
    Eo, Go, S = ml.oracle(M, False)
    assert rel_to_scale(E, Eo, S).max() < 1e-4
    assert grad_rel(G, Go).max() < 1e-4
    att = Dev(ml, grow=1)
    Ea, Ga = att.energrad(M, False)
    re_ = rel_to_scale(Ea, ml.z["att_E"], S)
    rg = grad_rel(-Ga[:, 3, :3], ml.z["att_G"][:, 3:])
    print("against ATTRACT: energy max", re_.max(), "median", np.median(re_), "gradient max", rg.max())
    assert re_.max() < 1e-4
    assert rg.max() < 1e-4
