"""The full attract-jax docking protocol on the device: staged adabelief minimization with
the gravity restraint (test5.py), conformer libraries (test1.py:46-63, test4.py:54-72,
test5.py:57-86) and the script-level preprocessing of receptor, ligand and grid
(test5.py:88-146).

`run_schedule` runs the stage list; every stage re-initialises the optimizer (fresh
moments and step count, test5.py:451-452) and is one persistent minimization launch.
`ConformerSet` turns a conformer library plus a per-pose conformer index into one job
per conformer inside a single launch.  `dock_test5` is test5.py as a function, and
`python -m attract_jax_amd.protocol` takes test5.py's arguments.
"""
import argparse
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import ForceField, Ligand, Receptor
from .minimize import BatchPlan, PoseState, minimize
from .pose import mat_from_qt, qt_from_mat

# (epochs, learning rate, gravity) of each stage, as test5.py:450 unpacks SCHEMA
# (test5.py:14-22; the comment there names the first two columns the other way round).
SCHEMA_TEST5 = (
    (5, 10.0, True),
    (10, 5.0, True),
    (10, 1.0, True),
    (50, 0.5, True),
    (100, 0.1, False),
    (100, 0.05, False),
    (100, 0.02, False),
)
GRAVITY_TEST5 = 0.1  # test5.py:420
ADABELIEF_EPS = 1e-16  # optax.adabelief defaults eps = eps_root = 1e-16
ADAM_EPS = 1e-8


def run_schedule(system, plan, state, schedule=SCHEMA_TEST5, optimizer=_lib.ATJ_OPT_ADABELIEF,
                 gravity=GRAVITY_TEST5, beta1=0.9, beta2=0.999, stream=None):
    """Minimize every pose of `plan` through the stages of `schedule` ((epochs, lr, gravity
    on/off) per stage; one learning rate for rotation and translation, as test5.py).
    state.energy ends with the energies of the last evaluation (test5.py:461)."""
    adab = optimizer == _lib.ATJ_OPT_ADABELIEF
    eps = ADABELIEF_EPS if adab else ADAM_EPS
    eps_root = ADABELIEF_EPS if adab else 0.0
    for epochs, lr, grav in schedule:
        plan.set_optimizer(int(epochs), float(lr), float(lr), beta1, beta2, eps, optimizer=optimizer,
                           eps_root=eps_root, gravity=gravity if grav else 0.0)
        state.m.zero_()
        state.v.zero_()
        state.count.zero_()
        minimize(system, plan, state, int(epochs), stream=stream)
    return state


class ConformerSet:
    """A conformer library (nconf, nlig, 3) of one ligand: one device Ligand per conformer.
    Poses choosing conformer c run as job c of one launch (test5.py:57-86: coordinates of
    pose p are coor_lig[conformers[p]])."""

    def __init__(self, coords, lig_atomtype_pos, lig_atomtypes, device=None):
        coords = np.asarray(coords, np.float32)
        if coords.ndim == 2:
            coords = coords[None]
        if coords.ndim != 3 or coords.shape[-1] != 3:
            raise ValueError("conformer coordinates must be (nconf, nlig, 3)")
        self.ligands = [Ligand.create(c, lig_atomtype_pos, lig_atomtypes, device) for c in coords]

    def __len__(self):
        return len(self.ligands)

    def plan(self, system, conformer_index, nsteps=0, lr_rot=1e-3, lr_trans=1e-3, **opt):
        """(BatchPlan, perm): poses grouped by conformer; pose slot i of the plan holds input
        pose perm[i].  conformer_index is 0-based, one entry per pose."""
        ci = np.asarray(conformer_index, np.int64).reshape(-1)
        if len(ci) and (ci.min() < 0 or ci.max() >= len(self.ligands)):
            raise ValueError("conformer index out of range")
        perm = np.argsort(ci, kind="stable")
        counts = np.bincount(ci, minlength=len(self.ligands))
        used = np.nonzero(counts)[0]
        plan = BatchPlan(system, [self.ligands[c] for c in used], counts[used], nsteps, lr_rot, lr_trans, **opt)
        return plan, perm


@dataclass
class Prepared:
    """Receptor, ligand and force field after test5.py's preprocessing."""
    coor_rec: np.ndarray        # (nrec', 3), type-99 beads removed
    rec_atomtypes: np.ndarray   # 0-based rows of the compacted pair table
    coor_lig: np.ndarray        # (nconf, nlig', 3)
    lig_atomtypes: np.ndarray   # 0-based columns of the compacted pair table
    lig_atomtype_pos: np.ndarray  # potential-grid index of every ligand bead
    ff: tuple                   # (rc, ac, ivor, emin, rmin2), each (nrec_types, nlig_types)
    rec_mapping: np.ndarray     # old receptor index -> new index (test5.py:94)


def prepare(coor_rec, rec_types, coor_lig, lig_types, par, grid_alphabet):
    """test5.py:88-146: drop type-99 beads, compact the type alphabets with np.unique,
    slice the 99 x 99 parameter tables (par = (rc, ac, ivor), attract-par.npz) to them,
    derive emin and rmin2 (test5.py:132-133) and locate every ligand type in the grid
    alphabet (test5.py:144-146)."""
    coor_rec = np.asarray(coor_rec)
    coor_lig = np.asarray(coor_lig)
    if coor_lig.ndim == 2:
        coor_lig = coor_lig[None]
    rt = np.asarray(rec_types).astype(np.uint8)
    lt = np.asarray(lig_types).astype(np.uint8)
    if rt.ndim != 1 or len(rt) != len(coor_rec):
        raise ValueError("one receptor type per receptor bead")
    if lt.ndim != 1 or len(lt) != coor_lig.shape[-2]:
        raise ValueError("one ligand type per ligand bead")
    rmask = rt != 99
    rec_mapping = np.cumsum(rmask) - 1
    lmask = lt != 99
    lig_alphabet, lig_atomtypes = np.unique(lt[lmask], return_inverse=True)
    rec_alphabet, rec_atomtypes = np.unique(rt[rmask], return_inverse=True)
    rc, ac, ivor = (np.asarray(a, np.float64)[rec_alphabet - 1][:, lig_alphabet - 1] for a in par)
    emin = -27.0 * ac ** 4 / (256.0 * rc ** 3)
    rmin2 = 4.0 * rc / (3.0 * ac)
    alpos = (np.nonzero(np.asarray(grid_alphabet))[0] + 1).tolist()
    try:
        lig_alphabet_pos = np.array([alpos.index(a) for a in lig_alphabet])
    except ValueError as e:
        raise ValueError("a ligand type has no potential grid") from e
    return Prepared(coor_rec[rmask], rec_atomtypes, coor_lig[:, lmask], lig_atomtypes,
                    lig_alphabet_pos[lig_atomtypes], (rc, ac, ivor, emin, rmin2), rec_mapping)


def remap_grid_neighbours(grid, rec_mapping):
    """Receptor indices of the grid's neighbour lists after type-99 removal
    (test5.py:120-122 for read_grid2's dense table; test-grid-energrad.py:82 for CSR)."""
    if hasattr(grid, "nr_neighbours"):
        ng = np.array(grid.neighbour_grid)
        flat = ng.reshape(-1)
        m = flat < 0xFFFF
        flat[m] = rec_mapping[flat[m]]
        return grid._replace(neighbour_grid=ng)
    nb = np.array(grid.neighbours)
    return grid._replace(neighbours=rec_mapping[nb].astype(nb.dtype))


def dock_test5(poses, coor_rec, rec_types, coor_lig, lig_types, grid, par, conformers=None,
               schedule=SCHEMA_TEST5, device="cuda:0", launch=None):
    """test5.py as a function: poses (N,4,4) row-vector matrices -> minimized poses (N,4,4)
    (rotation block, translation row, ones elsewhere as test5.py:463-465) and energies (N,).
    conformers: 1-based conformer index per pose (test5.py:70-80), or None."""
    from .score import ScoringSystem
    P = prepare(coor_rec, rec_types, coor_lig, lig_types, par, grid.alphabet)
    mat = np.asarray(poses, np.float64)
    N = len(mat)
    if conformers is None:
        ci = np.zeros(N, np.int64)
    else:
        ci = np.asarray(conformers).astype(np.int64) - 1
        if len(ci) != N or ci.min(initial=0) < 0 or ci.max(initial=0) >= len(P.coor_lig):
            raise ValueError("conformer indices must be 1..nconf, one per pose")
    grid = remap_grid_neighbours(grid, P.rec_mapping)
    system = ScoringSystem(grid, ForceField.create(*P.ff, device=device),
                           Receptor.create(P.coor_rec, P.rec_atomtypes, device), launch=launch)
    cs = ConformerSet(P.coor_lig, P.lig_atomtype_pos, P.lig_atomtypes, device)
    plan, perm = cs.plan(system, ci)
    q, t = qt_from_mat(mat[perm])
    state = PoseState.create(q, t, device)
    run_schedule(system, plan, state, schedule)
    pose = state.pose.cpu().numpy().astype(np.float64)
    out = np.ones_like(mat)
    out[perm] = mat_from_qt(pose[:, :4], pose[:, 4:], fill=1.0)
    E = np.zeros(N)
    E[perm] = state.energy.cpu().numpy()
    return out, E


def main(argv=None):
    """Command line of test5.py:41-50 (same positional arguments and options)."""
    from .grid import read_grid2
    ap = argparse.ArgumentParser(description="staged adabelief docking minimization (test5.py)")
    ap.add_argument("poses", help="docking poses as 4x4 matrices (.npy)")
    ap.add_argument("receptor", help="receptor coordinates (.npy)")
    ap.add_argument("ligand", help="ligand coordinates (.npy; (nconf, nlig, 3) for a conformer library)")
    ap.add_argument("--atrec", required=True, help="receptor atom types (.npy)")
    ap.add_argument("--atlig", required=True, help="ligand atom types (.npy)")
    ap.add_argument("--grid", required=True, help="ATTRACT docking grid")
    ap.add_argument("--par", required=True, help="pair parameters rc, ac, ivor (.npz, as attract-par.npz)")
    ap.add_argument("--conformers", help="ligand conformer indices (.npy, 1-based)")
    ap.add_argument("--out_poses", required=True)
    ap.add_argument("--out_ene", required=True)
    args = ap.parse_args(argv)
    par = np.load(args.par)
    lig = np.load(args.ligand)
    conf = np.load(args.conformers) if args.conformers is not None else None
    if lig.ndim == 2 and conf is not None:
        raise SystemExit("--conformers needs a conformer library (nconf, nlig, 3)")
    with open(args.grid, "rb") as f:
        grid = read_grid2(f.read(), read_gradients=True)
    poses, E = dock_test5(np.load(args.poses), np.load(args.receptor), np.load(args.atrec), lig,
                          np.load(args.atlig), grid, (par["rc"], par["ac"], par["ivor"]), conf)
    np.save(args.out_poses, poses)
    np.save(args.out_ene, E)


if __name__ == "__main__":
    main()
