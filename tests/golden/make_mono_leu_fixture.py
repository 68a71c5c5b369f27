"""Writes mono_leu.npz: the reference's own neighbour-list and RMSD outputs on its mono-leu
system, with the inputs they were computed from (run in a container that has the reference
checkout; the fixture is committed, this script is not needed at test time).

  rec_xyz, rec_type     mono-leu/mono-leuc-pdb.npy: 1275 beads, all ATTRACT type 15; the
                        scripts use this molecule as receptor and as ligand
                        (test-nb-energrad.py:43-44)
  backbone              names CA, C, O, N (test-rmsd.py:8)
  best_xyz              mono-leu/minim-best-ligand.pdb coordinates (the reference ligand of
                        test-rmsd.py; its parsed .npy lost 31 of the 1275 records, the PDB
                        holds them all, in the same order as mono-leuc.pdb)
  par_rbc/abc/ipon      attract-original/attract.par entry (15, 15), the table the scripts
                        load after convert-attract-par.py:4-22
  poses                 the first NPOSE matrices of mono-leu/minim-sorted.npy (float64)
  lrmsd                 test-rmsd.lrmsd for those poses (test-rmsd.py:13-19, 3 decimals)
  nb_E/nb_T/nb_Q        test-nb-energrad.out: energies, -dE/dM translation rows and torque
                        blocks of the first 10 poses (test-nb-energrad.py:253-261)
  nbp_E/nbp_T/nbp_Q     the same from test-nb-energrad-plateau.out
  dat_dofs              Euler angles and translations of the first NDAT structures of
                        mono-leu/minim-sorted.dat (receptor row, ligand row)
  att_E, att_G          ATTRACT's Energy and Gradients lines for them in
                        mono-leu/minim-sorted-nbgrid.score (mono-leu/dock.sh:22-27)
  list_total, list_max  neighbour-list entries of the grid the scripts read, and its longest
                        list: test-grid-energrad-retile.out:2-3 (printed by
                        test-grid-energrad-retile.py:95-101)
"""
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
NPOSE = 2000
NDAT = 1000
NUM = r"-?\d+\.?\d*(?:e[+-]?\d+)?"


def pdb_xyz(path):
    rows = [l for l in open(path).read().split("\n") if l.startswith("ATOM") or l.startswith("HETATM")]
    xyz = np.array([[float(l[30:38]), float(l[38:46]), float(l[46:54])] for l in rows])
    names = np.array([l[12:16].strip() for l in rows])
    return xyz, names


def jax_out(path):
    txt = open(path).read()
    en = txt.split("Energies:")[1].split("Gradients")[0]
    tr = txt.split("Gradients (translational):")[1].split("Gradients (torque)")[0]
    tq = txt.split("Gradients (torque):")[1].split("Run argsort")[0]
    f = lambda s, n: np.array([float(x) for x in re.findall(NUM, s)])[:n]
    return f(en, 10), f(tr, 30).reshape(10, 3), f(tq, 90).reshape(10, 3, 3)


def dat_structures(text, n):
    out, cur = [], None
    for line in text.split("\n"):
        s = line.strip()
        if s.startswith("#"):
            if re.fullmatch(r"#\d+", s):
                if len(out) == n:
                    break
                cur = []
                out.append(cur)
            continue
        if s and cur is not None:
            cur.append([float(x) for x in s.split()])
    return np.array(out)


def main():
    mol = np.load(os.path.join(REF, "mono-leu/mono-leuc-pdb.npy"), allow_pickle=False)
    rec_xyz = np.stack([mol["x"], mol["y"], mol["z"]], axis=1)
    best_xyz, best_names = pdb_xyz(os.path.join(REF, "mono-leu/minim-best-ligand.pdb"))
    names = mol["name"].astype("U4")
    assert (best_names == names).all()
    backbone = np.isin(names, ("CA", "C", "O", "N"))

    par_path = os.path.join(REF, "attract-original/attract.par")
    nt = int(open(par_path).readline().split()[1])
    par = np.loadtxt(par_path, skiprows=1).reshape(3, nt, nt)
    t = int(mol["occupancy"][0]) - 1
    assert (mol["occupancy"] == t + 1).all()

    poses = np.load(os.path.join(REF, "mono-leu/minim-sorted.npy"), allow_pickle=False)[:NPOSE]
    lrmsd = np.loadtxt(os.path.join(REF, "test-rmsd.lrmsd"))[:NPOSE]
    nb = jax_out(os.path.join(REF, "test-nb-energrad.out"))
    nbp = jax_out(os.path.join(REF, "test-nb-energrad-plateau.out"))

    dofs = dat_structures(open(os.path.join(REF, "mono-leu/minim-sorted.dat")).read(), NDAT)
    sc = open(os.path.join(REF, "mono-leu/minim-sorted-nbgrid.score")).read().split("\n")
    att_E = np.array([float(l.split()[1]) for l in sc if l.strip().startswith("Energy:")])[:NDAT]
    att_G = np.array([[float(x) for x in l.split()[1:]] for l in sc if l.strip().startswith("Gradients:")])[:NDAT]

    retile = open(os.path.join(REF, "test-grid-energrad-retile.out")).read()
    list_max = int(re.search(r"Maximum padding from \d+ to (\d+)", retile).group(1))
    list_total = int(re.search(r"Padding increase: (\d+)", retile).group(1))

    np.savez_compressed(os.path.join(OUT, "mono_leu.npz"), rec_xyz=rec_xyz.astype(np.float32),
                        rec_type=mol["occupancy"].astype(np.int64), backbone=backbone, best_xyz=best_xyz,
                        par_rbc=par[0, t, t], par_abc=par[1, t, t], par_ipon=par[2, t, t], poses=poses,
                        lrmsd=lrmsd, nb_E=nb[0], nb_T=nb[1], nb_Q=nb[2], nbp_E=nbp[0], nbp_T=nbp[1],
                        nbp_Q=nbp[2], dat_dofs=dofs, att_E=att_E, att_G=att_G, list_total=list_total,
                        list_max=list_max)


if __name__ == "__main__":
    main()
