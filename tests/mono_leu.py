"""The reference's mono-leu system and its printed results (tests/golden/mono_leu.npz).

The scripts that printed test-nb-energrad.out / test-nb-energrad-plateau.out read
mono-leu/mono-leuc.grid, which the reference does not hold (.gitignore:1-3).  Their neighbour
energies depend on that grid only through its inner box: its lists are complete (made with
neighbourdis 7 A and plateaudis 5 A, mono-leu/dock.sh:4, so every receptor bead within 5 A of
any point of a voxel's cell lies within 5 + 0.78 A of its grid point), and a bead is looked up
only while 0 <= (x' - origin) / spacing < dim - 1 (test-nb-energrad.py:182-190).

The box ATTRACT's make-grid chose is inferred as origin = min(receptor) - (plateaudis -
spacing / 2), dim = ceil((max - min + 2 (plateaudis - spacing / 2)) / spacing) = (89, 74, 59),
spacing 0.9 A.  Three independent facts fix it:
  * the grid's list total and longest list, printed by test-grid-energrad-retile.py:95-101
    (2,510,079 entries, 72 longest): lists of this box give 2,510,078 and 72, the one missing
    entry being a pair 7.000001 A from its grid point;
  * test-nb-energrad.out: pose 6 has a contributing bead at z' = 24.21 A beyond the box's upper
    z bound of 23.93 A, every other contributing bead of the ten poses is inside, and all ten
    energies then agree with the FP64 restatement within 1e-4;
  * ATTRACT's own scores with the same grid geometry (mono-leu/minim-sorted-nbgrid.score) agree
    best with a z cut at 24.38 A, the box bound plus half a voxel (ATTRACT looks a bead up while
    its nearest grid point is in the grid).
"""
import math
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SPACING = 0.9
PLATEAUDIS = 5.0
NEIGHBOURDIS = 7.0


def attract_inner_box(coor_rec, spacing=SPACING, plateaudis=PLATEAUDIS):
    """(origin (3,), dim (3,)) of the inner grid make-grid builds around coor_rec."""
    lo = np.asarray(coor_rec, np.float64).min(axis=0)
    hi = np.asarray(coor_rec, np.float64).max(axis=0)
    m = plateaudis - spacing / 2
    dim = np.array([int(math.ceil(v)) for v in (hi - lo + 2 * m) / spacing])
    return lo - m, dim


class MonoLeu:
    def __init__(self):
        z = np.load(os.path.join(HERE, "golden", "mono_leu.npz"))
        self.z = z
        self.coor = z["rec_xyz"].astype(np.float64)  # receptor and ligand (test-nb-energrad.py:43-44)
        self.nbead = len(self.coor)
        self.types = np.zeros(self.nbead, np.int64)  # one type (15): compacted alphabet index 0
        one = lambda v: np.full((1, 1), float(v))
        self.ff = O.ff_from_par(one(z["par_rbc"]), one(z["par_abc"]), one(z["par_ipon"]))
        self.origin, self.dim = attract_inner_box(self.coor)
        self.poses = z["poses"]

    def jax(self, plateau):
        p = "nbp_" if plateau else "nb_"
        return self.z[p + "E"], self.z[p + "T"], self.z[p + "Q"]

    def oracle(self, mat, plateau):
        """FP64 neighbour-list energies, dE/dM and per-pose sum of |pair energy|."""
        return O.nb_energrad_complete(mat, self.coor, self.types, self.coor, self.types, self.ff, self.origin,
                                      SPACING, self.dim, PLATEAUDIS, plateau=plateau)

    def dat_mats(self):
        """Pose matrices of the .dat structures the ATTRACT scores belong to."""
        from attract_jax_amd import structio
        d = structio.DatFile(pivots={1: np.zeros(3), 2: np.zeros(3)}, dofs=self.z["dat_dofs"])
        return structio.dat_to_mat(d)

    def backbone(self):
        """(ligand backbone coordinates, reference backbone coordinates) of test-rmsd.py:8-12."""
        b = self.z["backbone"]
        return self.coor[b], self.z["best_xyz"][b]


def rel_to_scale(E, E_ref, S):
    """|E - E_ref| relative to max(|E_ref|, sum |pair terms|, 1): the scale of an energy that is a
    sum of terms of both signs (a sum near zero carries the rounding of its terms)."""
    return np.abs(np.asarray(E, np.float64) - E_ref) / np.maximum(np.maximum(np.abs(E_ref), S), 1.0)


def grad_rel(neg_grad, ref):
    """Per-pose max |difference| over the block's largest reference component."""
    ax = tuple(range(1, np.ndim(ref)))
    return np.abs(np.asarray(neg_grad, np.float64) - ref).max(axis=ax) / np.abs(ref).max(axis=ax)
