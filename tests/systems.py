"""Small synthetic docking systems for the tests, built entirely by the oracle (CPU)."""
from collections import namedtuple

import numpy as np

from oracle import oracle as O


def tiny_table(rng, ntypes=4):
    rmin = rng.uniform(3.0, 5.0, size=(ntypes, ntypes))
    rmin = (rmin + rmin.T) / 2
    eabs = rng.uniform(0.1, 1.5, size=(ntypes, ntypes))
    eabs = (eabs + eabs.T) / 2
    ivor = np.where(rng.random((ntypes, ntypes)) < 0.25, -1.0, 1.0)
    ivor = np.triu(ivor) + np.triu(ivor, 1).T
    rbc = rmin * np.sqrt(3) / 2
    abc = 256 / 27 * eabs
    return O.ff_from_par(rbc, abc, ivor)


def build_reference_grid(coor_rec, rec_types, ff, ntypes, origin, spacing, dim, dim2, gridextension, plateaudis,
                         neighbourdis, cutoff=50.0, with_potentials=True):
    """A reference-format (read_grid, CSR) Grid built by the oracle's grid semantics."""
    origin = np.asarray(origin, np.float64)
    ii = np.stack(np.meshgrid(*[np.arange(d) for d in dim], indexing="ij"), -1).reshape(-1, 3)
    pts = origin + spacing * ii
    d = {}
    d["gridspacing"] = float(spacing)
    d["gridextension"] = int(gridextension)
    d["plateaudis"] = float(plateaudis)
    d["neighbourdis"] = float(neighbourdis)
    alphabet = np.zeros(99, bool)
    alphabet[:ntypes] = True
    d["alphabet"] = alphabet
    d["origin"] = origin
    d["dim"] = np.array(dim)
    d["dim2"] = np.array(dim2)
    d["natoms"] = len(coor_rec)
    lists = [O.grid_point_neighbours(p, coor_rec, neighbourdis) for p in pts]
    counts = np.array([len(l) for l in lists])
    d["neighbours"] = np.concatenate(lists).astype(np.uint32) if counts.sum() else np.zeros(0, np.uint32)
    if with_potentials:
        inner = np.stack([O.grid_point_potential(pts, coor_rec, rec_types, t, ff, plateaudis, cutoff)
                          for t in range(ntypes)]).reshape(ntypes, *dim, 4)
        d["inner_potential_grid"] = inner.astype(np.float32)
    starts = np.concatenate(([0], np.cumsum(counts)[:-1]))
    ng = np.stack([counts, np.where(counts > 0, starts + 1, 0)], -1).astype(np.int32).reshape(*dim, 2)
    d["neighbour_grid"] = ng
    if with_potentials:
        jj = np.stack(np.meshgrid(*[np.arange(v) for v in dim2], indexing="ij"), -1).reshape(-1, 3)
        pts2 = origin + spacing * (2 * jj - gridextension)
        outer = np.stack([O.grid_point_potential(pts2, coor_rec, rec_types, t, ff, plateaudis, cutoff)
                          for t in range(ntypes)]).reshape(ntypes, *dim2, 4)
        d["outer_potential_grid"] = outer.astype(np.float32)
    return namedtuple("Grid", d.keys())(*d.values())


class TinySystem:
    """40-bead receptor in a 6 A sphere, 4 types, inner 14^3 @1 A, outer 10^3 @2 A."""

    def __init__(self, seed=0, nlig=5, npose=64, with_potentials=True, plateau=3.0):
        rng = np.random.default_rng(seed)
        from attract_jax_amd.workloads import random_chain, random_receptor
        self.coor_rec = random_receptor(40, 6.0, 2.0, rng)
        self.rec_types = rng.integers(0, 4, size=40)
        self.ff = tiny_table(rng)
        self.ntypes = 4
        self.spacing = 1.0
        nbdis = plateau + self.spacing * np.sqrt(3) / 2 + 1e-3
        self.grid = build_reference_grid(self.coor_rec, self.rec_types, self.ff, 4, origin=(-7.0, -7.0, -7.0),
                                         spacing=1.0, dim=(14, 14, 14), dim2=(10, 10, 10), gridextension=6,
                                         plateaudis=plateau, neighbourdis=nbdis, with_potentials=with_potentials)
        self.coor_lig = random_chain(nlig, rng, bond=3.0, min_sep=2.5)
        self.lig_types = rng.integers(0, 4, size=nlig)
        from attract_jax_amd.pose import random_poses, mat_from_qt
        q, t = random_poses(npose, rng, 3.0, 9.0)
        self.quat, self.trans = q, t
        self.mat = mat_from_qt(q, t)

    def oracle_energrad(self, mat=None, **kw):
        return O.energrad(self.mat if mat is None else mat, self.coor_lig, self.lig_types, self.lig_types,
                          self.coor_rec, self.rec_types, self.ff, self.grid, **kw)


def pose_rmsd(Pa, Pb, coor_lig):
    """Ligand RMSD between two sets of (quaternion, translation) poses (N, 7)."""
    xa = O.transform(O.mat_from_qt(np.asarray(Pa, np.float64)[:, :4], np.asarray(Pa, np.float64)[:, 4:]), coor_lig)
    xb = O.transform(O.mat_from_qt(np.asarray(Pb, np.float64)[:, :4], np.asarray(Pb, np.float64)[:, 4:]), coor_lig)
    return O.rmsd(xa, xb)


def _nudged(P, cols, ulps, direction):
    Q = P.copy()
    for _ in range(ulps):
        Q[:, cols] = np.nextafter(Q[:, cols], np.float32(direction))
    return Q


def twin_conditioning(twin, pose7, nsteps, lr_rot, lr_trans, coor_lig, ulps=(1, 4, 16), tol=0.005, **kw):
    """Minimize with the FP32 twin from the given starts, once more with its bias corrections
    rounded otherwise at every step, and from nudged copies: the translations, then the
    quaternion, each up and down by every count of units in the last place in `ulps`.  A pose whose nudged results differ from its own by more than `tol` A RMSD
    amplifies rounding differences (the optimizer's normalised steps flip with the sign of
    gradient components near zero, a bead crosses the inner grid's edge a step earlier or later):
    another FP32 implementation, whose sums round differently, cannot be expected to follow it.
    Returns (poses, energies, well_conditioned, largest nudged RMSD)."""
    P0 = np.array(pose7, np.float32).reshape(-1, 7)
    Pa, Ea = twin.minimize(P0, nsteps, lr_rot, lr_trans, **kw)
    # and the same starts with the bias corrections rounded otherwise at every step
    Pb, _ = twin.minimize(P0, nsteps, lr_rot, lr_trans, bias_pow=True, **kw)
    r = pose_rmsd(Pa, Pb, coor_lig)
    for u in ulps:
        for cols, direction in ((slice(4, 7), np.inf), (slice(4, 7), -np.inf), (slice(0, 4), np.inf),
                                (slice(0, 4), -np.inf)):
            Pb, _ = twin.minimize(_nudged(P0, cols, u, direction), nsteps, lr_rot, lr_trans, **kw)
            r = np.maximum(r, pose_rmsd(Pa, Pb, coor_lig))
    return Pa, Ea, r < tol, r


def smooth_poses(mat, coor_lig, coor_rec, grid, margin):
    """Poses whose beads all lie more than `margin` A from every voxel face of the grid they are
    looked up in, from the inner grid's boundary (and the outer grid's, for beads outside the
    inner one), and (for every receptor bead) from the plateau sphere."""
    xp = O.transform(mat, coor_lig)
    v, inside = O.voxelize(xp.reshape(-1, 3), grid.origin, grid.gridspacing, grid.dim)
    vo = (v + grid.gridextension) / 2
    frac = np.where(inside[:, None], v, vo)
    cell = np.where(inside[:, None], grid.gridspacing, 2 * grid.gridspacing)
    face = (np.abs(frac - np.round(frac)) * cell).min(axis=1)
    edge = np.minimum(np.abs(v), np.abs(v - (np.asarray(grid.dim) - 1))).min(axis=1) * grid.gridspacing
    if getattr(grid, "dim2", None) is not None:
        edge2 = np.minimum(np.abs(vo), np.abs(vo - (np.asarray(grid.dim2) - 1))).min(axis=1) * 2 * grid.gridspacing
        edge = np.minimum(edge, np.where(inside, np.inf, edge2))
    d = np.linalg.norm(xp[:, :, None, :] - np.asarray(coor_rec)[None, None], axis=-1)
    shell = np.abs(d - grid.plateaudis).min(axis=2).reshape(-1)
    ok = (face > margin) & (edge > margin) & (shell > margin)
    return ok.reshape(len(mat), -1).all(axis=1)
