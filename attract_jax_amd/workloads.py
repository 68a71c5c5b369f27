"""Synthetic docking systems of the five benchmark configurations.

All randomness is numpy.random.default_rng(1000 * config + stream).  Receptor
beads: random sequential addition inside a sphere with a minimum separation.
Pair table: symmetric 32x32 with rmin ~ U(3, 5) A, |emin| ~ U(0.1, 1.5), a pair
repulsive (ivor = -1) with probability 0.1; mapped to the ATTRACT parameters
rbc = rmin sqrt(3)/2, abc = 256/27 |emin|, rc = abc rbc^8, ac = abc rbc^6
(attract-original/convert-attract-par.py:16-22).  Bead types are 1..32; the pair
table and the potential grids are indexed by type - 1.
"""
from dataclasses import dataclass, field

import numpy as np

from . import pose as posemod

NTYPES = 32


def rng_for(config, stream):
    return np.random.default_rng(1000 * config + stream)


def random_receptor(n, radius, min_sep, rng, max_tries=200000):
    """n points uniformly inside a sphere, pairwise >= min_sep apart (sequential addition)."""
    pts = np.empty((n, 3))
    k = 0
    tries = 0
    ms2 = min_sep * min_sep
    while k < n:
        batch = rng.uniform(-radius, radius, size=(256, 3))
        batch = batch[(batch * batch).sum(1) <= radius * radius]
        for c in batch:
            tries += 1
            if k == 0 or ((pts[:k] - c) ** 2).sum(1).min() >= ms2:
                pts[k] = c
                k += 1
                if k == n:
                    break
        if tries > max_tries:
            raise RuntimeError("random sequential addition did not converge; lower min_sep")
    return pts


def jittered_receptor(n, radius, min_sep, rng):
    """n points inside a sphere, pairwise >= min_sep apart: random sites of a cubic lattice (as
    dense as the sphere needs for n sites, with room to spare) each moved by a uniform jitter
    that keeps neighbours min_sep apart.  Vectorised: for the many receptors of a serving
    stream, where sequential addition would take seconds each."""
    a = (4.0 / 3.0 * np.pi * radius ** 3 / (1.3 * n)) ** (1.0 / 3.0)
    if a < min_sep:
        raise ValueError("too many beads for the sphere at this separation")
    k = int(np.ceil(radius / a)) + 1
    g = np.stack(np.meshgrid(*[np.arange(-k, k + 1)] * 3, indexing="ij"), -1).reshape(-1, 3) * a
    j = (a - min_sep) / 2.0
    g = g[np.linalg.norm(g, axis=1) <= radius - j * np.sqrt(3.0)]
    if len(g) < n:
        raise ValueError("lattice too coarse for n beads")
    pts = g[rng.choice(len(g), n, replace=False)]
    return pts + rng.uniform(-j, j, size=pts.shape)


def random_pair_table(rng, ntypes=NTYPES, p_repulsive=0.1):
    """(rc, ac, ivor, rmin, emin_abs) symmetric ntypes x ntypes float64 tables."""
    iu = np.triu_indices(ntypes)
    rmin = np.zeros((ntypes, ntypes))
    eabs = np.zeros((ntypes, ntypes))
    ivor = np.ones((ntypes, ntypes))
    rmin[iu] = rng.uniform(3.0, 5.0, size=len(iu[0]))
    eabs[iu] = rng.uniform(0.1, 1.5, size=len(iu[0]))
    ivor[iu] = np.where(rng.random(len(iu[0])) < p_repulsive, -1.0, 1.0)
    for a in (rmin, eabs, ivor):
        a.T[iu] = a[iu]
    rc, ac = rc_ac_from_rmin_emin(rmin, eabs)
    return rc, ac, ivor, rmin, eabs


def rc_ac_from_rmin_emin(rmin, eabs):
    rbc = rmin * np.sqrt(3.0) / 2.0
    abc = 256.0 / 27.0 * np.abs(eabs)
    return abc * rbc ** 8, abc * rbc ** 6


def random_chain(n, rng, bond=3.8, min_sep=3.0):
    """A compact self-avoiding bead chain of n beads, centred on its centroid."""
    pts = [np.zeros(3)]
    while len(pts) < n:
        for _ in range(1000):
            d = rng.normal(size=3)
            d *= bond / np.linalg.norm(d)
            c = pts[-1] + d
            others = np.array(pts[:-1])
            if len(others) == 0 or ((others - c) ** 2).sum(1).min() >= min_sep ** 2:
                # keep the chain compact: prefer steps towards the centroid
                if len(pts) > 3 and np.linalg.norm(c - np.mean(pts, 0)) > 1.3 * bond * len(pts) ** (1 / 3):
                    continue
                pts.append(c)
                break
        else:
            raise RuntimeError("chain growth failed")
    pts = np.array(pts)
    return pts - pts.mean(0)


LEUCINE_LIKE = np.array([[-1.45, 0.55, 0.30],   # N
                         [0.00, 0.70, 0.10],    # CA
                         [0.55, 2.10, 0.35],    # C
                         [0.80, -0.45, -0.55]])  # side-chain bead (CSE-like)


@dataclass
class LigandSpec:
    coor: np.ndarray     # (nbeads, 3) local coordinates
    types: np.ndarray    # (nbeads,) 1..32


@dataclass
class Workload:
    name: str
    config: int
    coor_rec: np.ndarray
    rec_types: np.ndarray
    rc: np.ndarray
    ac: np.ndarray
    ivor: np.ndarray
    rmin: np.ndarray
    emin_abs: np.ndarray
    grid_kw: dict
    ligands: list
    pose_counts: list
    quat: np.ndarray
    trans: np.ndarray
    nsteps: int = 0
    lr_rot: float = 1e-3
    lr_trans: float = 1e-3
    extra: dict = field(default_factory=dict)

    @property
    def npose(self):
        return int(sum(self.pose_counts))


def c1(npose=1000):
    """Plain method, small: 300 beads in R=15 A (min sep 3 A); inner grid 32^3 @0.9 A
    with an outer 32^3 @1.8 A (gridextension 16) so that ligand centres at 18-25 A
    lie inside a grid; neighbour lists within 5 A (plateau 4.2 A = 5 A minus half a
    voxel diagonal); 4-bead leucine-like ligand; 100 ADAM steps, lr 0.1 / 0.01."""
    rng = rng_for(1, 0)
    coor_rec = random_receptor(300, 15.0, 3.0, rng)
    rec_types = rng_for(1, 1).integers(1, NTYPES + 1, size=300)
    rc, ac, ivor, rmin, eabs = random_pair_table(rng_for(1, 2))
    lig = LigandSpec(LEUCINE_LIKE.copy(), rng_for(1, 3).integers(1, NTYPES + 1, size=4))
    q, t = posemod.random_poses(npose, rng_for(1, 4), 18.0, 25.0)
    grid_kw = dict(origin=(-14.4, -14.4, -14.4), spacing=0.9, dim=(32, 32, 32), dim2=(32, 32, 32),
                   gridextension=16, plateaudis=4.2, neighbourdis=5.0)
    return Workload("C1 plain small: 300-bead receptor, 32^3 grid, 4-bead ligand", 1, coor_rec, rec_types, rc, ac,
                    ivor, rmin, eabs, grid_kw, [lig], [npose], q, t, nsteps=100, lr_rot=0.01, lr_trans=0.1)


def _c2_receptor():
    rng = rng_for(2, 0)
    coor_rec = random_receptor(2000, 25.0, 2.25, rng)
    rec_types = rng_for(2, 1).integers(1, NTYPES + 1, size=2000)
    rc, ac, ivor, rmin, eabs = random_pair_table(rng_for(2, 2))
    grid_kw = dict(origin=(-43.2, -43.2, -43.2), spacing=0.9, dim=(96, 96, 96), dim2=(64, 64, 64),
                   gridextension=16, plateaudis=10.0, neighbourdis=10.8)
    return coor_rec, rec_types, rc, ac, ivor, rmin, eabs, grid_kw


def c2(npose=131072):
    """Single-shot energrad at scale: 2000 beads in R=25 A, inner 96^3 @0.9 A + outer
    64^3 @1.8 A, plateau 10 A / lists 10.8 A; 21-bead trinucleotide-like ligand;
    centres at 25-35 A."""
    coor_rec, rec_types, rc, ac, ivor, rmin, eabs, grid_kw = _c2_receptor()
    lig = LigandSpec(random_chain(21, rng_for(2, 3)), rng_for(2, 4).integers(1, NTYPES + 1, size=21))
    q, t = posemod.random_poses(npose, rng_for(2, 5), 25.0, 35.0)
    return Workload("C2 energrad: 2000-bead receptor, 96^3+64^3 grids, 21-bead ligand", 2, coor_rec, rec_types, rc,
                    ac, ivor, rmin, eabs, grid_kw, [lig], [npose], q, t)


def c3(nligands=64, poses_per_ligand=4096, nsteps=200):
    """Batching: the C2 receptor grid, 64 ligands of U{5..60} beads, 4096 poses each,
    200 fused ADAM steps (lr 1e-3, the reference's single rate)."""
    coor_rec, rec_types, rc, ac, ivor, rmin, eabs, grid_kw = _c2_receptor()
    r = rng_for(3, 0)
    sizes = r.integers(5, 61, size=nligands)
    ligs = [LigandSpec(random_chain(int(n), rng_for(3, 100 + i)), rng_for(3, 200 + i).integers(1, NTYPES + 1, size=n))
            for i, n in enumerate(sizes)]
    q, t = posemod.random_poses(nligands * poses_per_ligand, rng_for(3, 1), 25.0, 35.0)
    return Workload("C3 batched minimization: 64 ligands x 4096 poses, 5-60 beads, 200 ADAM steps", 3, coor_rec,
                    rec_types, rc, ac, ivor, rmin, eabs, grid_kw, ligs, [poses_per_ligand] * nligands, q, t,
                    nsteps=nsteps, lr_rot=1e-3, lr_trans=1e-3)


def c4(npose=65536):
    """Differentiation: C2 energy, 65536 poses, d/d(pose DOFs) and d/d(rmin, emin)."""
    w = c2(npose)
    w.name = "C4 adjoint: C2 system, 65536 poses"
    w.config = 4
    return w


def serving_jobs(njobs, seed=5000, rate=20.0):
    """Config-5 job stream: Poisson arrivals (mean `rate` jobs/s), receptors of
    500-4000 beads, ligands of 5-60 beads, log-uniform 1e3..2e5 poses, 100-300 steps."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        nrec = int(rng.integers(500, 4001))
        jobs.append(dict(arrival=t, nrec=nrec, nlig=int(rng.integers(5, 61)),
                         nposes=int(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))),
                         nsteps=int(rng.integers(100, 301)), seed=seed + 1 + i))
    return jobs


def build_system(w, device=None, layout=None, keep_plain=False, flags=None, launch=None, nshell=None):
    """Device objects of a workload: (ScoringSystem, [Ligand]); the grid is computed on the GPU
    (neighbour lists in nshell distance shells, ScoringSystem's default when None; 0 = whole lists)."""
    from . import _lib
    from .device import DeviceGrid, ForceField, Ligand, Receptor
    from .score import ScoringSystem
    if layout is None:
        layout = _lib.ATJ_LAYOUT_CORNER
    ff = ForceField.create(w.rc, w.ac, w.ivor, device=device)
    rec = Receptor.create(w.coor_rec, w.rec_types - 1, device)
    grid = DeviceGrid.build(rec, ff, np.arange(NTYPES), layout=layout, device=device, keep_plain=keep_plain,
                            **w.grid_kw)
    ligs = [Ligand.create(l.coor, l.types - 1, l.types - 1, device) for l in w.ligands]
    kw = {} if nshell is None else dict(nshell=nshell)
    return ScoringSystem(grid, ff, rec, flags=flags, launch=launch, **kw), ligs


@dataclass
class ReceptorSpec:
    """A serving job's receptor: beads, types 1..32, its own pair table and grid geometry."""
    key: object
    coor: np.ndarray
    types: np.ndarray
    rc: np.ndarray
    ac: np.ndarray
    ivor: np.ndarray
    grid_kw: dict

    @property
    def radius(self):
        return float(np.linalg.norm(self.coor, axis=1).max())

    def plain_bytes(self):
        d, d2 = self.grid_kw["dim"], self.grid_kw["dim2"]
        return NTYPES * 16 * (int(np.prod(d)) + int(np.prod(d2)))

    def corner_bytes(self):
        d, d2 = self.grid_kw["dim"], self.grid_kw["dim2"]
        return NTYPES * 128 * (int(np.prod(np.subtract(d, 1))) + int(np.prod(np.subtract(d2, 1))))

    def list_bytes(self):
        """Neighbour-list memory, estimated from the bead density (lists of about 4/3 pi
        neighbourdis^3 x density entries near the receptor, padded to 32), plus the per-point
        records."""
        v = int(np.prod(self.grid_kw["dim"]))
        r = self.radius + self.grid_kw["neighbourdis"]
        near = min(1.0, 4.0 / 3.0 * np.pi * r ** 3 / (v * self.grid_kw["spacing"] ** 3))
        dens = len(self.coor) / (4.0 / 3.0 * np.pi * max(self.radius, 1.0) ** 3)
        per = 4.0 / 3.0 * np.pi * self.grid_kw["neighbourdis"] ** 3 * dens
        return int(v * (8 + 16) + v * near * 2 * 32 * np.ceil(per / 32 + 1))


def c5_receptor(key, nrec, rng):
    """Receptor of a serving job: nrec beads at the C2 bead density (radius 25 A x (nrec/2000)^(1/3)),
    own 32-type pair table, inner grid at 0.9 A reaching 12-26 A beyond the receptor (more for
    larger receptors), outer grid at 1.8 A another 14.4 A out: plain fp32 grids of 0.18 GB (500
    beads) to 1.4 GB (4000 beads)."""
    radius = 25.0 * (nrec / 2000.0) ** (1.0 / 3.0)
    coor = jittered_receptor(nrec, radius, 2.25, rng)
    types = rng.integers(1, NTYPES + 1, size=nrec)
    rc, ac, ivor, _, _ = random_pair_table(rng)
    margin = 12.0 + 14.0 * (nrec - 500) / 3500.0
    d = int(np.ceil(2 * (radius + margin) / 0.9)) + 1
    d2 = (d + 1) // 2 + 16
    o = -0.9 * (d - 1) / 2
    grid_kw = dict(origin=(o, o, o), spacing=0.9, dim=(d, d, d), dim2=(d2, d2, d2), gridextension=16,
                   plateaudis=10.0, neighbourdis=10.8)
    return ReceptorSpec(key, coor, types, rc, ac, ivor, grid_kw)


def c5_stream(njobs=200, rate=20.0, seed=5999, pose_scale=1.0, nrec_range=(500, 4001)):
    """The serving workload: Poisson arrivals (mean `rate` jobs/s); each job brings its own
    receptor (500-4000 beads, its own pair table and grid), a ligand of 5-60 beads, log-uniform
    1e3..2e5 poses (times pose_scale) around the receptor surface and 100-300 ADAM steps
    (lr 1e-3).  Returns the JobSpecs (receptor_key = the job's receptor) and {key: ReceptorSpec}."""
    from .serve import JobSpec
    rng = np.random.default_rng(seed)
    t = 0.0
    specs, receptors = [], {}
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        r = c5_receptor(("c5", i), int(rng.integers(*nrec_range)), np.random.default_rng(seed + 7000 + i))
        receptors[r.key] = r
        nb = int(rng.integers(5, 61))
        coor = random_chain(nb, np.random.default_rng(seed + 1000 + i))
        types = rng.integers(1, NTYPES + 1, size=nb)
        npose = max(1, int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))) * pose_scale))
        rad = r.radius
        q, tr = posemod.random_poses(npose, np.random.default_rng(seed + 5000 + i), rad, rad + 10.0)
        specs.append(JobSpec(r.key, coor, types, q, tr, int(rng.integers(100, 301)), arrival=t))
    return specs, receptors


def c5_receptors(n=16, seed=5000):
    """Receptor pool of the serving workload: 500-4000 beads at the C2 bead density
    (radius 25 A * (n/2000)^(1/3)), own 32-type pair table, inner grid at 0.9 A covering
    the receptor plus 22 A, outer 64^3 at 1.8 A; plain fp32 grids of 0.2-1.1 GB."""
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + 100 * i)
        nrec = int(rng.integers(500, 4001))
        radius = 25.0 * (nrec / 2000.0) ** (1.0 / 3.0)
        coor = random_receptor(nrec, radius, 2.25, rng)
        types = rng.integers(1, NTYPES + 1, size=nrec)
        rc, ac, ivor, rmin, eabs = random_pair_table(rng)
        d = int(np.ceil(2 * (radius + 22.0) / 0.9)) + 1
        o = -0.9 * (d - 1) / 2
        grid_kw = dict(origin=(o, o, o), spacing=0.9, dim=(d, d, d), dim2=(64, 64, 64), gridextension=16,
                       plateaudis=10.0, neighbourdis=10.8)
        out.append(dict(key=i, nrec=nrec, radius=radius, coor=coor, types=types, rc=rc, ac=ac, ivor=ivor,
                        grid_kw=grid_kw))
    return out


def c5_jobs(receptors, njobs=200, rate=20.0, seed=5999, pose_scale=1.0):
    """Poisson arrivals (mean `rate` jobs/s); each job: a receptor of the pool, a ligand of
    5-60 beads, log-uniform 1e3..2e5 poses (times pose_scale) around the receptor surface,
    100-300 ADAM steps (lr 1e-3)."""
    from .serve import JobSpec
    rng = np.random.default_rng(seed)
    t = 0.0
    specs = []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        r = receptors[int(rng.integers(0, len(receptors)))]
        nb = int(rng.integers(5, 61))
        coor = random_chain(nb, np.random.default_rng(seed + 1000 + i))
        types = rng.integers(1, NTYPES + 1, size=nb)
        npose = max(1, int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))) * pose_scale))
        q, tr = posemod.random_poses(npose, np.random.default_rng(seed + 5000 + i), r["radius"],
                                     r["radius"] + 10.0)
        specs.append(JobSpec(r["key"], coor, types, q, tr, int(rng.integers(100, 301)), arrival=t))
    return specs


def c5_backend_receptors(receptors, device, layout=None, keep_plain=False):
    """Backend receptor table: key -> {'grid_bytes', 'build'} for serve.GpuBackend."""
    from . import _lib
    from .device import DeviceGrid, ForceField, Receptor
    from .score import ScoringSystem
    layout = _lib.ATJ_LAYOUT_CORNER if layout is None else layout
    table = {}
    for r in receptors:
        d = r["grid_kw"]["dim"][0]
        plain = NTYPES * 16 * (d ** 3 + 64 ** 3)
        corner = NTYPES * 128 * ((d - 1) ** 3 + 63 ** 3)
        lists = d ** 3 * 8 + d ** 3 * 2 * 48

        def build(r=r):
            ff = ForceField.create(r["rc"], r["ac"], r["ivor"], device=device)
            rec = Receptor.create(r["coor"], r["types"] - 1, device)
            grid = DeviceGrid.build(rec, ff, np.arange(NTYPES), layout=layout, device=device, keep_plain=keep_plain,
                                    **r["grid_kw"])
            return ScoringSystem(grid, ff, rec)
        table[r["key"]] = dict(grid_bytes=(corner if layout == _lib.ATJ_LAYOUT_CORNER else plain) + lists,
                               build=build)
    return table
