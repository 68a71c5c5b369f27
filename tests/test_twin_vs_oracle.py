"""The FP32 CPU twin (oracle/twin.c) against the FP64 NumPy oracle: energies and gradients
within 1e-4 relative, voxel indices identical except at FP32/FP64 rounding boundaries."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.twin import T_EXACT, T_GRID, T_NB, T_PLATEAU, Twin
from tests.systems import TinySystem


@pytest.fixture(scope="module")
def tiny():
    s = TinySystem(seed=21, nlig=8, npose=200)
    tw = Twin(s.grid, s.ff, s.coor_rec, s.rec_types, s.coor_lig, s.lig_types, s.lig_types)
    return s, tw


def _scale(G):
    return np.maximum(np.abs(G).max(axis=(1, 2), keepdims=True), 1.0)


@pytest.mark.parametrize("exact", [False, True])
def test_twin_energrad(tiny, exact):
    s, tw = tiny
    E, G = s.oracle_energrad(exact=exact)
    flags = T_GRID | T_NB | T_PLATEAU | (T_EXACT if exact else 0)
    Et, Gt = tw.energrad(s.mat.astype(np.float32), flags=flags)
    den = np.maximum(np.abs(E), 1.0)
    assert (np.abs(Et - E) / den).max() < 1e-4
    assert (np.abs(Gt - G) / _scale(G)).max() < 1e-4


def test_twin_nb_only_no_plateau(tiny):
    s, tw = tiny
    E, G = s.oracle_energrad(terms=("nb",), plateau=False)
    Et, Gt = tw.energrad(s.mat.astype(np.float32), flags=T_NB)
    assert (np.abs(Et - E) / np.maximum(np.abs(E), 1.0)).max() < 1e-4
    assert (np.abs(Gt - G) / _scale(G)).max() < 1e-4


def test_twin_quaternion_poses(tiny):
    s, tw = tiny
    qt = np.concatenate([s.quat, s.trans], 1).astype(np.float32)
    E, _ = s.oracle_energrad(O.mat_from_qt(qt[:, :4].astype(np.float64), qt[:, 4:].astype(np.float64)))
    Et, _ = tw.energrad(qt, qt=True)
    assert (np.abs(Et - E) / np.maximum(np.abs(E), 1.0)).max() < 1e-4


def test_twin_voxels(tiny):
    s, tw = tiny
    out = tw.bead_terms(s.mat.astype(np.float32))
    vox64 = O.nb_voxel(O.transform(s.mat, s.coor_lig), s.grid)
    assert (out["voxel"] != vox64).mean() < 1e-3
    assert (out["voxel"] >= 0).mean() > 0.3


def test_twin_minimize_vs_oracle(tiny):
    s, tw = tiny
    q0, t0 = s.quat[:50], s.trans[:50]
    lr_r, lr_t, nst = 0.01, 0.1, 30

    def fn(q, t):
        return s.oracle_energrad(O.mat_from_qt(q, t))

    q, t, E = O.adam_minimize(q0, t0, nst, lr_r, lr_t, fn)
    P, Et = tw.minimize(np.concatenate([q0, t0], 1), nst, lr_r, lr_t)
    x_o = O.transform(O.mat_from_qt(q, t), s.coor_lig)
    x_t = O.transform(O.mat_from_qt(P[:, :4].astype(np.float64), P[:, 4:].astype(np.float64)), s.coor_lig)
    r = O.rmsd(x_o, x_t)
    assert np.median(r) < 0.01
    assert (r < 0.05).mean() >= 0.9


def test_twin_adabelief_gravity_vs_oracle(tiny):
    """test5.py's optimizer: adabelief with the gravity restraint, twin against oracle."""
    s, tw = tiny
    q0, t0 = s.quat[:40], s.trans[:40]
    lr, nst, w = 0.05, 20, 0.1

    def fn(q, t):
        return s.oracle_energrad(O.mat_from_qt(q, t))

    q, t, E = O.optax_minimize(q0, t0, nst, lr, lr, fn, "adabelief", eps=1e-16, eps_root=1e-16, gravity=w)
    P, Et = tw.minimize(np.concatenate([q0, t0], 1), nst, lr, lr, eps=1e-16, kind=1, eps_root=1e-16, gravity=w)
    x_o = O.transform(O.mat_from_qt(q, t), s.coor_lig)
    x_t = O.transform(O.mat_from_qt(P[:, :4].astype(np.float64), P[:, 4:].astype(np.float64)), s.coor_lig)
    r = O.rmsd(x_o, x_t)
    assert np.median(r) < 0.01 and (r < 0.05).mean() >= 0.9
    close = r < 1e-4
    assert close.mean() >= 0.5
    assert (np.abs(Et[close] - E[close]) / np.maximum(np.abs(E[close]), 1.0)).max() < 1e-3
