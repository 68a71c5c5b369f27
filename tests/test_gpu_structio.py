"""Device ligand RMSD (atj_ligand_rmsd, test-rmsd.py:13-17) against the oracle's FP64
RMSD and closed forms; a ligand longer than one wavefront."""
import numpy as np
import pytest

from attract_jax_amd import structio as S
from attract_jax_amd.pose import mat_from_qt, random_quaternions
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("nb", [4, 64, 150])
def test_ligand_rmsd_vs_oracle(nb):
    rng = np.random.default_rng(nb)
    lig = rng.normal(scale=5.0, size=(nb, 3))
    ref = lig + rng.normal(scale=2.0, size=(nb, 3))
    N = 1000
    M = mat_from_qt(random_quaternions(N, rng), rng.uniform(-20, 20, (N, 3)), fill=1.0)
    r = S.ligand_rmsd(M, lig, ref, DEV).cpu().numpy()
    ro = O.rmsd(O.transform(M.astype(np.float32).astype(np.float64), lig.astype(np.float32)),
                ref.astype(np.float32).astype(np.float64)[None])
    assert np.abs(r - ro).max() / np.abs(ro).max() < 1e-5


def test_ligand_rmsd_translation_closed_form():
    rng = np.random.default_rng(1)
    lig = rng.normal(size=(30, 3))
    t = np.array([[3.0, 4.0, 0.0], [0.0, 0.0, -2.0], [0.0, 0.0, 0.0]])
    M = mat_from_qt(np.tile([0.0, 0.0, 0.0, 1.0], (3, 1)), t, fill=1.0)
    r = S.ligand_rmsd(M, lig, lig, DEV).cpu().numpy()
    assert np.allclose(r, [5.0, 2.0, 0.0], atol=1e-5)
