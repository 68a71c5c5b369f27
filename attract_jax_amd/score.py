"""Reference-named scoring entry points, computed by the HIP kernels.

Names, argument order and meaning follow the reference scripts:
  potential_energrad / potential_energy   test-grid-energrad.py:214-244
  neighbour_energy                         test-grid-energrad.py:262-296
  main / value_and_grad_main               test-grid-energrad.py:298-317
  main_quat / value_and_grad_quat          test4.py:389-401,415
`grid` may be a reference Grid namedtuple (read_grid / read_grid2 / pad_grid output,
neighbour indices remapped) or a DeviceGrid; `ff` a namedtuple with fields rc, ac,
ivor, emin, rmin2 (compacted, test-grid-energrad.py:67-79) or a ForceField.  Results
are float32 torch tensors on the GPU.
"""
import ctypes
import hashlib

import numpy as np
import torch

from . import _lib
from .device import DeviceGrid, ForceField, Ligand, Receptor, _ptr, _stream

_grid_cache = {}
DEFAULT_NSHELL = 6  # distance shells of the neighbour lists a ScoringSystem reads (0: whole lists)


def device_grid(grid, layout=_lib.ATJ_LAYOUT_CORNER):
    """DeviceGrid of a reference Grid, uploaded once per grid object."""
    if isinstance(grid, DeviceGrid):
        return grid
    key = (id(grid), layout)
    hit = _grid_cache.get(key)
    if hit is not None and hit[0] is grid:
        return hit[1]
    if len(_grid_cache) > 8:
        _grid_cache.clear()
    dg = DeviceGrid.from_grid(grid, layout=layout)
    _grid_cache[key] = (grid, dg)
    return dg


def device_ff(ff):
    if isinstance(ff, ForceField):
        return ff
    return ForceField.create(ff.rc, ff.ac, ff.ivor, ff.emin, ff.rmin2)


def _f32(x, device):
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=device, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(x), dtype=np.float32), device=device)


def _default_flags(dg, terms):
    flags = 0
    if "grid" in terms and dg.has_potentials:
        flags |= _lib.ATJ_TERM_GRID
    if "nb" in terms:
        flags |= _lib.ATJ_TERM_NB
        if dg.has_potentials:
            flags |= _lib.ATJ_NB_PLATEAU
    return flags


def point_terms(coords, grid_types, ff_types, grid, ff=None, receptor=None, flags=None, terms=("grid", "nb")):
    """Per-point grid4 (M,4), voxel (M,), nb energy (M,), nb gradient (M,3) at explicit coordinates."""
    lib = _lib.load()
    dg = device_grid(grid)
    d = dg.device
    c = _f32(coords, d).reshape(-1, 3)
    M = c.shape[0]
    gt = np.asarray(grid_types, np.int64).reshape(-1)
    lt = np.asarray(ff_types, np.int64).reshape(-1) if ff_types is not None else np.zeros_like(gt)
    if len(gt) and (gt.min() < 0 or lt.min() < 0 or gt.max() > 0xFFFF or lt.max() > 0x7FFF):
        raise ValueError("point types out of range")
    if ff is None:
        ff = ForceField.create(np.ones((1, 1)), np.ones((1, 1)), np.ones((1, 1)))
    ffd = device_ff(ff)
    if len(gt) and dg.has_potentials and gt.max() >= dg.ntypes:
        raise ValueError(f"point grid type {gt.max()} beyond the {dg.ntypes} potential grids")
    if len(lt) and lt.max() >= ffd.nlt:
        raise ValueError(f"point type {lt.max()} beyond the {ffd.nlt} columns of the pair table")
    types = torch.as_tensor((gt | (lt << 16)).astype(np.int32), device=d)
    if flags is None:
        flags = _default_flags(dg, terms)
    if receptor is None:
        receptor = Receptor(torch.zeros((1, 4), device=d), 0)
    grid4 = torch.zeros((M, 4), dtype=torch.float32, device=d)
    voxel = torch.empty(M, dtype=torch.int32, device=d)
    nbe = torch.zeros(M, dtype=torch.float32, device=d)
    nbg = torch.zeros((M, 3), dtype=torch.float32, device=d)
    whole = dg.nshell > 0 and not dg.shells_valid_for(receptor)  # shells of other coordinates: whole lists
    _lib.check(lib.atj_point_terms(ctypes.byref(dg.struct(whole_lists=whole)), ctypes.byref(ffd.struct()),
                                   ctypes.byref(receptor.struct()), _ptr(c), _ptr(types), M, _ptr(grid4),
                                   _ptr(voxel), _ptr(nbe), _ptr(nbg), flags, _stream(d)))
    return grid4, voxel, nbe, nbg


def potential_energrad(all_coors_lig, lig_atomtype_pos, grid):
    """(N, nlig, 4) interpolated [E, gx, gy, gz] of each bead (stored ATTRACT sign)."""
    c = _f32(all_coors_lig, device_grid(grid).device)
    N, nlig = c.shape[0], c.shape[1]
    gt = np.broadcast_to(np.asarray(lig_atomtype_pos).reshape(1, nlig), (N, nlig))
    grid4, _, _, _ = point_terms(c, gt, None, grid, flags=_lib.ATJ_TERM_GRID)
    return grid4.reshape(N, nlig, 4)


def potential_energy(all_coors_lig, lig_atomtype_pos, grid):
    """(N,) sum over beads of the interpolated grid energy."""
    return potential_energrad(all_coors_lig, lig_atomtype_pos, grid)[:, :, 0].sum(dim=1)


def neighbour_energy(coor_rec, rec_atomtypes, all_coors_lig, lig_atomtypes, ff, grid, plateau=None):
    """(N,) neighbour-list pair energy.  plateau defaults to 'grid has potentials'
    (test-grid-energrad.py:129-131 vs test-nb-energrad.py:142-145)."""
    dg = device_grid(grid)
    c = _f32(all_coors_lig, dg.device)
    N, nlig = c.shape[0], c.shape[1]
    if plateau is None:
        plateau = dg.has_potentials
    flags = _lib.ATJ_TERM_NB | (_lib.ATJ_NB_PLATEAU if plateau else 0)
    rec = Receptor.create(coor_rec, rec_atomtypes, dg.device)
    lt = np.broadcast_to(np.asarray(lig_atomtypes).reshape(1, nlig), (N, nlig))
    _, _, nbe, _ = point_terms(c, np.zeros_like(lt), lt, dg, ff, rec, flags=flags)
    return nbe.reshape(N, nlig).sum(dim=1)


class ScoringSystem:
    """Receptor grid, pair table and receptor beads resident on the device, plus a
    work-counter workspace: the state the energrad / minimize launches share."""

    def __init__(self, grid, ff, receptor, flags=None, launch=None, nshell=DEFAULT_NSHELL):
        """nshell > 0 puts the grid's lists in distance shells for this receptor (same
        contributing pairs, fewer list entries read per bead) unless they already are; a
        grid in shell order for other receptor coordinates is read whole."""
        self.grid = device_grid(grid)
        if nshell and self.grid.nshell == 0:
            self.grid.shell_lists(receptor, nshell)
        self.ff = device_ff(ff)
        self.receptor = receptor
        self.device = self.grid.device
        self.flags = flags if flags is not None else _default_flags(self.grid, ("grid", "nb"))
        self.launch = launch or _lib.AtjLaunch(8, 0, 0)
        self.counter = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.stats = torch.zeros(_lib.ATJ_NSTATS + _lib.ATJ_NPHASES, dtype=torch.int64, device=self.device)
        self.refresh()

    def refresh(self):
        """Re-read the device pointers of grid, pair table and receptor (after any of them
        was replaced); the tensors they point to are kept alive with the structs."""
        g = self.grid
        self.receptor.check_types(self.ff)
        whole = g.nshell > 0 and not g.shells_valid_for(self.receptor)
        self._gs = g.struct(whole_lists=whole)
        self._fs = self.ff.struct()
        self._rs = self.receptor.struct()
        self._alive = (g.inner, g.outer, g.nbgrid, g.nblist, None if whole else g.nbshell, self.receptor.beads,
                       self.ff)
        self._last_jobs = None  # (ligand, its beads' address and count, its job table) of the last energrad

    def jobs_tensor(self, ligands, nsteps=0, lr_rot=1e-3, lr_trans=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
                    pose_counts=None, optimizer=_lib.ATJ_OPT_ADAM, eps_root=0.0, gravity=0.0):
        """Device array of AtjJob records for the given ligands (per-ligand arrays or
        scalars for nsteps, lr_rot, lr_trans, gravity)."""
        if not isinstance(ligands, (list, tuple)):
            ligands = [ligands]
        arr = (_lib.AtjJob * len(ligands))()
        begin = 0
        pick = lambda v, i: float(v[i]) if np.ndim(v) else float(v)
        for i, lig in enumerate(ligands):
            lig.check_types(self.grid, self.ff)
            n = int(pose_counts[i]) if pose_counts is not None else 0
            ns = int(nsteps[i]) if np.ndim(nsteps) else int(nsteps)
            opt = _lib.AtjOptimizer(int(optimizer), pick(lr_rot, i), pick(lr_trans, i), float(beta1), float(beta2),
                                    float(eps), float(eps_root), pick(gravity, i))
            arr[i] = _lib.AtjJob(lig.struct(), begin, n, ns, 0, opt)
            begin += n
        raw = np.frombuffer(bytes(arr), np.uint8)
        return torch.from_numpy(raw.copy()).to(self.device)

    def energrad(self, poses, ligand, *, pose_job=None, jobs=None, njobs=1, order=None, grad=True, flags=None,
                 stats=False, qt=False, grad_qt=False, stream=None):
        """energies (N,), dE/dM (N,4,4) of poses (N,4,4) [or (N,7) with qt=True; with grad_qt
        as well the gradient is dE/d(q, t) (N,7), computed by the kernel]."""
        lib = _lib.load()
        d = self.device
        P = _f32(poses, d)
        N = P.shape[0]
        P = P.reshape(N, 7 if qt else 16)
        if grad_qt and not qt:
            raise ValueError("grad_qt needs quaternion poses (qt=True)")
        if jobs is None:  # the job table of the last ligand is kept: repeated calls upload nothing
            key = (ligand.beads.data_ptr(), ligand.nbeads)
            if self._last_jobs is None or self._last_jobs[0] is not ligand or self._last_jobs[1] != key:
                self._last_jobs = (ligand, key, self.jobs_tensor([ligand]))
            jobs = self._last_jobs[2]
            njobs = 1
        energy = torch.empty(N, dtype=torch.float32, device=d)
        dEdM = torch.empty((N, 4, 4), dtype=torch.float32, device=d) if grad else None
        fl = ((self.flags if flags is None else flags) | (_lib.ATJ_POSE_QT if qt else 0)
              | (_lib.ATJ_GRAD_QT if grad_qt else 0))
        st = stream if stream is not None else _stream(d)
        _lib.check(lib.atj_energrad(ctypes.byref(self._gs), ctypes.byref(self._fs), ctypes.byref(self._rs),
                                    _ptr(jobs), njobs, _ptr(pose_job), _ptr(order), _ptr(P), N, _ptr(energy),
                                    _ptr(dEdM), fl, ctypes.byref(self.launch), _ptr(self.counter),
                                    _ptr(self.stats) if stats else ctypes.c_void_p(0), st))
        if grad and grad_qt:
            return energy, dEdM.reshape(N, 16)[:, :7]
        return energy, dEdM

    def bead_terms(self, poses, ligand, flags=None, qt=False):
        lib = _lib.load()
        d = self.device
        P = _f32(poses, d)
        N = P.shape[0]
        P = P.reshape(N, 7 if qt else 16)
        ligand.check_types(self.grid, self.ff)
        nb = ligand.nbeads
        grid4 = torch.zeros((N, nb, 4), dtype=torch.float32, device=d)
        voxel = torch.empty((N, nb), dtype=torch.int32, device=d)
        nbe = torch.zeros((N, nb), dtype=torch.float32, device=d)
        nbg = torch.zeros((N, nb, 3), dtype=torch.float32, device=d)
        fl = (self.flags if flags is None else flags) | (_lib.ATJ_POSE_QT if qt else 0)
        _lib.check(lib.atj_bead_terms(ctypes.byref(self._gs), ctypes.byref(self._fs), ctypes.byref(self._rs),
                                      ctypes.byref(ligand.struct()), _ptr(P), N, _ptr(grid4), _ptr(voxel),
                                      _ptr(nbe), _ptr(nbg), fl, _stream(d)))
        return grid4, voxel, nbe, nbg


_system_cache = {}


def _content_key(a):
    """Key of a small host array (or tensor, or namedtuple of them) by its values."""
    if isinstance(a, tuple):
        return tuple(_content_key(x) for x in a)
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    a = np.ascontiguousarray(np.asarray(a))
    return (a.dtype.str, a.shape, hashlib.blake2b(a.tobytes(), digest_size=16).digest())


def _system(coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid):
    """ScoringSystem and Ligand of the reference-named calls.  They are made once per grid
    object and per receptor / ligand / pair-table values: repeated calls (the reference's
    jitted main called again on new poses) reuse the device copies."""
    dg = device_grid(grid)
    ffk = _content_key(tuple(getattr(ff, f) for f in ("rc", "ac", "ivor", "emin", "rmin2")))
    key = (id(dg), _content_key(coor_rec), _content_key(rec_atomtypes), _content_key(coor_lig),
           _content_key(lig_atomtypes), _content_key(lig_atomtype_pos), ffk)
    hit = _system_cache.get(key)
    if hit is not None and hit[0] is dg:
        return hit[1], hit[2]
    rec = Receptor.create(coor_rec, rec_atomtypes, dg.device)
    lig = Ligand.create(coor_lig, lig_atomtype_pos, lig_atomtypes, dg.device)
    sysm = ScoringSystem(dg, device_ff(ff), rec)
    if len(_system_cache) > 8:
        _system_cache.clear()
    _system_cache[key] = (dg, sysm, lig)
    return sysm, lig


def main(mat, coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid):
    """(sum of energies, energies (N,)) of the poses mat (N,4,4) (test-grid-energrad.py:298-306)."""
    sysm, lig = _system(coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid)
    e, _ = sysm.energrad(mat, lig, grad=False)
    return e.sum(), e


def value_and_grad_main(mat, coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid):
    """((total, energies), dE/dmat (N,4,4)) -- jax.value_and_grad(main, has_aux=True)
    with the ATTRACT gradient convention (test-grid-energrad.py:316-327)."""
    sysm, lig = _system(coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid)
    e, g = sysm.energrad(mat, lig, grad=True)
    return (e.sum(), e), g


def main_quat(quat, trans, coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid):
    """test4.py:396-401 with a single (non-conformer) ligand: (total, energies)."""
    sysm, lig = _system(coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid)
    d = sysm.device
    qt = torch.cat([_f32(quat, d), _f32(trans, d)], dim=1)
    e, _ = sysm.energrad(qt, lig, grad=False, qt=True)
    return e.sum(), e


def value_and_grad_quat(quat, trans, coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid):
    """((total, energies), (dE/dquat (N,4), dE/dtrans (N,3))) -- test4.py:415."""
    sysm, lig = _system(coor_rec, rec_atomtypes, coor_lig, lig_atomtypes, lig_atomtype_pos, ff, grid)
    d = sysm.device
    qt = torch.cat([_f32(quat, d), _f32(trans, d)], dim=1)
    e, g = sysm.energrad(qt, lig, grad=True, qt=True, grad_qt=True)
    return (e.sum(), e), (g[:, :4].contiguous(), g[:, 4:7].contiguous())
