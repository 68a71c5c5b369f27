"""Pose and structure I/O around the docking path.

* ATTRACT `.dat` pose files (e.g. mono-leu/minim-sorted.dat): per structure one line of
  degrees of freedom per molecule, `phi ssi rot x y z` (an optional leading conformer
  number), after `#pivot`/`#centered` header lines.  `dat_to_mat` gives the reference's
  4x4 row-vector pose matrices of the ligand in the receptor frame
  (x' = [x, 1] @ M, M[:3,:3] = R^T, M[3,:3] = translation): the form mono-leu/conv.sh
  converts minim-sorted.dat into minim-sorted.npy.  `write_dat` writes matrices back.
* ATTRACT's Euler rotation R(phi, ssi, rot) and its inverse, and the ATTRACT gradients
  -dE/d(phi, ssi, rot, x, y, z) of a pose from dE/dM (the Gradients lines of
  mono-leu/minim-sorted-grid.score; test-grid-energrad.py:319-327 prints -dE/dM because
  "sign is opposite of ATTRACT").  `write_score` writes ATTRACT's Energy/Gradients format.
* PDB bead files: ATOM/HETATM records with the ATTRACT bead type in columns 58-59
  (1B7F-bound/score-bound.sh:4).
* Ligand RMSD of poses (test-rmsd.py:13-17) on the device, and energy sort / top-k
  (ATTRACT's sort.py / top, mono-leu/dock.sh:13-15).
"""
import re
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .device import _ptr, _stream


# --------------------------------------------------------------------------- Euler angles

def euler_to_rotation(phi, ssi, rot):
    """ATTRACT's rotation matrix (column-vector convention, x' = R x) of Euler angles
    (phi, ssi, rot), any broadcastable shapes -> (..., 3, 3)."""
    phi, ssi, rot = np.broadcast_arrays(*(np.asarray(a, np.float64) for a in (phi, ssi, rot)))
    cs, ss = np.cos(ssi), np.sin(ssi)
    cp, sp = np.cos(phi), np.sin(phi)
    cr, sr = np.cos(rot), np.sin(rot)
    R = np.empty(phi.shape + (3, 3))
    R[..., 0, 0] = cr * cs * cp + sr * sp
    R[..., 0, 1] = sr * cs * cp - cr * sp
    R[..., 0, 2] = ss * cp
    R[..., 1, 0] = cr * cs * sp - sr * cp
    R[..., 1, 1] = sr * cs * sp + cr * cp
    R[..., 1, 2] = ss * sp
    R[..., 2, 0] = -cr * ss
    R[..., 2, 1] = -sr * ss
    R[..., 2, 2] = cs
    return R


def euler_derivatives(phi, ssi, rot):
    """dR/dphi, dR/dssi, dR/drot of euler_to_rotation -> (3, ..., 3, 3)."""
    phi, ssi, rot = np.broadcast_arrays(*(np.asarray(a, np.float64) for a in (phi, ssi, rot)))
    cs, ss = np.cos(ssi), np.sin(ssi)
    cp, sp = np.cos(phi), np.sin(phi)
    cr, sr = np.cos(rot), np.sin(rot)
    z = np.zeros_like(phi)
    dphi = np.stack([np.stack([-cr * cs * sp + sr * cp, -sr * cs * sp - cr * cp, -ss * sp], -1),
                     np.stack([cr * cs * cp + sr * sp, sr * cs * cp - cr * sp, ss * cp], -1),
                     np.stack([z, z, z], -1)], -2)
    dssi = np.stack([np.stack([-cr * ss * cp, -sr * ss * cp, cs * cp], -1),
                     np.stack([-cr * ss * sp, -sr * ss * sp, cs * sp], -1),
                     np.stack([-cr * cs, -sr * cs, -ss], -1)], -2)
    drot = np.stack([np.stack([-sr * cs * cp + cr * sp, cr * cs * cp + sr * sp, z], -1),
                     np.stack([-sr * cs * sp - cr * cp, cr * cs * sp - sr * cp, z], -1),
                     np.stack([sr * ss, -cr * ss, z], -1)], -2)
    return np.stack([dphi, dssi, drot])


def rotation_to_euler(R):
    """Euler angles (phi, ssi, rot) of rotation matrices (..., 3, 3) (column-vector
    convention), ssi in [0, pi].  At ssi = 0 only phi - rot is defined and at ssi = pi only
    phi + rot; there rot = 0 is taken."""
    R = np.asarray(R, np.float64)
    cs = np.clip(R[..., 2, 2], -1.0, 1.0)
    ssi = np.arccos(cs)
    gen = np.sin(ssi) > 1e-9
    phi = np.where(gen, np.arctan2(R[..., 1, 2], R[..., 0, 2]),
                   np.where(cs > 0, np.arctan2(R[..., 1, 0], R[..., 0, 0]),
                            np.arctan2(-R[..., 1, 0], -R[..., 0, 0])))
    rot = np.where(gen, np.arctan2(-R[..., 2, 1], -R[..., 2, 0]), 0.0)
    return phi, ssi, rot


def attract_gradients(euler, dEdM):
    """ATTRACT's gradients -dE/d(phi, ssi, rot, x, y, z) (N, 6) of poses with Euler angles
    euler (N, 3) and reference-convention dE/dM (N, 4, 4) (M[:3,:3] = R^T): the chain
    rule through euler_to_rotation.  The angles must be the ones the pose was given in:
    the two Euler triples of one rotation give different angle derivatives."""
    e = np.asarray(euler, np.float64).reshape(-1, 3)
    G = np.asarray(dEdM, np.float64).reshape(-1, 4, 4)
    dR = euler_derivatives(e[:, 0], e[:, 1], e[:, 2])  # (3, N, 3, 3)
    dang = np.einsum("nkl,anlk->na", G[:, :3, :3], dR)  # d(R^T)_kl = dR_lk
    return -np.concatenate([dang, G[:, 3, :3]], axis=1)


# --------------------------------------------------------------------------- .dat files

@dataclass
class DatFile:
    """Parsed ATTRACT .dat: header lines, pivots, and per structure the comment lines,
    the energy (from '## Energy:' if present) and the (nmol, 6) degrees of freedom."""
    header: list = field(default_factory=list)
    pivots: dict = field(default_factory=dict)
    centered_receptor: bool = False
    centered_ligands: bool = False
    comments: list = field(default_factory=list)
    energies: np.ndarray = None
    dofs: np.ndarray = None          # (N, nmol, 6): phi ssi rot x y z
    conformers: np.ndarray = None    # (N, nmol) 1-based conformer numbers, 0 where absent


def read_dat(text):
    """Parse ATTRACT .dat text (mono-leu/minim-sorted.dat is the format)."""
    d = DatFile()
    structs, comments, energies, confs = [], [], [], []
    cur = None
    for line in text.splitlines():
        s = line.strip()
        if not s:
            continue
        if s.startswith("#"):
            if re.fullmatch(r"#\d+", s):
                cur = []
                structs.append(cur)
                comments.append([])
                energies.append(np.nan)
                confs.append([])
            elif cur is None:
                d.header.append(line)
                m = re.match(r"#pivot\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)", s)
                if m:
                    d.pivots[int(m.group(1))] = np.array([float(m.group(k)) for k in (2, 3, 4)])
                m = re.match(r"#centered (receptor|ligands):\s*(\w+)", s)
                if m:
                    setattr(d, "centered_receptor" if m.group(1) == "receptor" else "centered_ligands",
                            m.group(2).lower() == "true")
            else:
                comments[-1].append(line)
                m = re.match(r"##\s*Energy:\s*(\S+)", s)
                if m:
                    energies[-1] = float(m.group(1))
            continue
        if cur is None:
            raise ValueError(f"degrees of freedom before the first structure: {line!r}")
        v = s.split()
        if len(v) == 7:
            confs[-1].append(int(v[0]))
            v = v[1:]
        elif len(v) == 6:
            confs[-1].append(0)
        else:
            raise ValueError(f"expected 6 or 7 values per molecule line: {line!r}")
        cur.append([float(x) for x in v])
    nmol = {len(sv) for sv in structs}
    if len(nmol) > 1:
        raise ValueError("structures with different numbers of molecules")
    d.comments = comments
    d.energies = np.array(energies)
    d.dofs = np.array(structs, np.float64).reshape(len(structs), -1, 6)
    d.conformers = np.array(confs, np.int64).reshape(len(structs), -1)
    return d


def _pivot(d, k):
    """Rotation centre of molecule k: its header pivot, the origin for a centered molecule."""
    centered = d.centered_ligands if k > 1 else d.centered_receptor
    return np.zeros(3) if centered else d.pivots.get(k, np.zeros(3))


def dat_to_mat(d, ligand=2, receptor=1):
    """Pose matrices (N, 4, 4) of molecule `ligand` in the frame of molecule `receptor`,
    in the reference's row-vector form.  Each molecule k is placed as
    x' = R_k (x - p_k) + p_k + t_k (pivot p_k, zero for centered molecules); pinned for zero
    pivots against mono-leu/minim-sorted.npy."""
    dof = d.dofs
    Rl = euler_to_rotation(dof[:, ligand - 1, 0], dof[:, ligand - 1, 1], dof[:, ligand - 1, 2])
    Rr = euler_to_rotation(dof[:, receptor - 1, 0], dof[:, receptor - 1, 1], dof[:, receptor - 1, 2])
    pl, pr = _pivot(d, ligand), _pivot(d, receptor)
    tl = dof[:, ligand - 1, 3:] + pl - Rl @ pl
    tr = dof[:, receptor - 1, 3:] + pr - Rr @ pr
    Rrel = np.swapaxes(Rr, 1, 2) @ Rl                     # x_rel = Rr^T (Rl x + tl - tr)
    trel = np.einsum("nlk,nl->nk", Rr, tl - tr)           # Rr^T (tl - tr)
    M = np.ones((len(dof), 4, 4))
    M[:, :3, :3] = np.swapaxes(Rrel, 1, 2)
    M[:, 3, :3] = trel
    return M


def write_dat(mats, energies=None, euler=None):
    """ATTRACT .dat text of ligand poses (N, 4, 4) with the receptor fixed at the origin
    (zero pivots, as mono-leu/minim-sorted.dat).  euler (N, 3) selects the Euler triple to
    write (default: rotation_to_euler's)."""
    M = np.asarray(mats, np.float64).reshape(-1, 4, 4)
    if euler is None:
        phi, ssi, rot = rotation_to_euler(np.swapaxes(M[:, :3, :3], 1, 2))
    else:
        e = np.asarray(euler, np.float64).reshape(-1, 3)
        phi, ssi, rot = e[:, 0], e[:, 1], e[:, 2]
    out = ["#pivot 1 0.000 0.000 0.000", "#pivot 2 0.000 0.000 0.000", "#centered receptor: false",
           "#centered ligands: false"]
    for i in range(len(M)):
        out.append(f"#{i + 1}")
        if energies is not None:
            out.append(f"## Energy: {float(energies[i]):.3f}")
        out.append("   0.000000 0.000000 0.000000 0.0000 0.0000 0.0000")
        t = M[i, 3, :3]
        out.append(f"   {phi[i]:.6f} {ssi[i]:.6f} {rot[i]:.6f} {t[0]:.4f} {t[1]:.4f} {t[2]:.4f}")
    return "\n".join(out) + "\n"


def write_score(energies, gradients=None):
    """ATTRACT --score output: per structure 'Energy:', the energy-term line and, with
    gradients (N, 6) = attract_gradients(...), a 'Gradients:' line."""
    out = []
    E = np.asarray(energies, np.float64).reshape(-1)
    for i in range(len(E)):
        out.append(f" Energy: {float(E[i])!r}")
        out.append(f"   {E[i]:11.3f}" + "       0.000" * 5)
        if gradients is not None:
            out.append(" Gradients: " + " ".join(repr(float(g)) for g in gradients[i]))
    return "\n".join(out) + "\n"


def read_score(text):
    """(energies (N,), gradients (N, 6) or None) of ATTRACT --score output."""
    E, G = [], []
    for line in text.splitlines():
        s = line.split()
        if s and s[0] == "Energy:":
            E.append(float(s[1]))
        elif s and s[0] == "Gradients:":
            G.append([float(x) for x in s[1:]])
    return np.array(E), (np.array(G) if G else None)


# --------------------------------------------------------------------------- PDB beads

def read_pdb(text):
    """ATOM/HETATM records of a bead PDB: dict(name, resname, resid, xyz (n, 3),
    atomtype (n,) int from columns 58-59, 0 where blank)."""
    name, resname, resid, xyz, at = [], [], [], [], []
    for line in text.splitlines():
        if not (line.startswith("ATOM") or line.startswith("HETATM")):
            continue
        name.append(line[12:16].strip())
        resname.append(line[17:20].strip())
        resid.append(int(line[22:26]))
        xyz.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
        t = line[57:59].strip()
        at.append(int(t) if t else 0)
    return dict(name=np.array(name), resname=np.array(resname), resid=np.array(resid, np.int64),
                xyz=np.array(xyz, np.float64).reshape(-1, 3), atomtype=np.array(at, np.int64))


# --------------------------------------------------------------------------- ranking / RMSD

def sort_poses(energies, k=None):
    """Indices of the poses by ascending energy (stable, as ATTRACT's sort.py), the first k
    only when k is given (ATTRACT's top).  numpy arrays or torch tensors."""
    if isinstance(energies, torch.Tensor):
        idx = torch.sort(energies, stable=True).indices
    else:
        idx = np.argsort(np.asarray(energies), kind="stable")
    return idx if k is None else idx[:k]


def ligand_rmsd(poses, coor_lig, coor_ref, device="cuda:0", stream=None):
    """Ligand RMSD sqrt(sum_j |x'_j - ref_j|^2 / n) of every pose (test-rmsd.py:13-17):
    poses (N, 4, 4) row-vector matrices applied to coor_lig (n, 3), against coor_ref (n, 3).
    Computed on the device (atj_ligand_rmsd); returns a float32 tensor (N,)."""
    lib = _lib.load()
    P = torch.as_tensor(np.asarray(poses, np.float32).reshape(-1, 16), device=device).contiguous()
    L = torch.as_tensor(np.asarray(coor_lig, np.float32).reshape(-1, 3), device=device).contiguous()
    R = torch.as_tensor(np.asarray(coor_ref, np.float32).reshape(-1, 3), device=device).contiguous()
    if L.shape != R.shape:
        raise ValueError("ligand and reference need the same beads")
    out = torch.empty(P.shape[0], dtype=torch.float32, device=device)
    st = stream if stream is not None else _stream(torch.device(device))
    _lib.check(lib.atj_ligand_rmsd(_ptr(P), P.shape[0], _ptr(L), _ptr(R), L.shape[0], _ptr(out), st))
    return out
