"""Pose / structure I/O against data the reference holds (fixtures written by
tests/golden/make_structio_fixtures.py): ATTRACT .dat parsing and the .dat -> 4x4 matrix
conversion (mono-leu/minim-sorted.dat vs minim-sorted.npy), ATTRACT gradients from dE/dM
(test-grid-energrad.out vs mono-leu/minim-sorted-grid.score), the PDB bead reader
(mono-leuc.pdb vs mono-leuc-pdb.npy); closed forms for the rest."""
import os

import numpy as np
import pytest

from attract_jax_amd import structio as S

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _text(name):
    with open(os.path.join(GOLD, name)) as f:
        return f.read()


def test_read_dat_excerpt():
    d = S.read_dat(_text("dat_excerpt.dat"))
    assert d.dofs.shape == (12, 2, 6)
    assert set(d.pivots) == {1, 2} and not d.centered_receptor and not d.centered_ligands
    assert d.energies[0] == -1632.033 and np.all(np.isfinite(d.energies))
    assert np.allclose(d.dofs[0, 1], [-4.486760, -1.217123, 5.320105, 7.1279, -20.0625, -8.6534])
    assert np.all(d.conformers == 0)


def test_dat_to_mat_matches_reference_conversion():
    g = np.load(os.path.join(GOLD, "dat_pairs.npz"))
    d = S.DatFile(dofs=np.stack([g["rec"], g["lig"]], axis=1))
    M = S.dat_to_mat(d)
    err = np.abs(M - g["mat"]).max(axis=(1, 2))
    assert np.median(err) < 1e-15 and err.max() < 1e-6  # a few entries carry float32-level rounding


def test_dat_write_read_round_trip_and_pivots():
    rng = np.random.default_rng(5)
    e = rng.uniform(-np.pi, np.pi, (20, 3))
    R = S.euler_to_rotation(*e.T)
    M = np.ones((20, 4, 4))
    M[:, :3, :3] = np.swapaxes(R, 1, 2)
    M[:, 3, :3] = rng.uniform(-30, 30, (20, 3))
    E = rng.normal(size=20)
    d = S.read_dat(S.write_dat(M, E))
    assert np.abs(S.dat_to_mat(d)[:, :3, :3] - M[:, :3, :3]).max() < 2e-6
    assert np.abs(S.dat_to_mat(d)[:, 3, :3] - M[:, 3, :3]).max() < 1e-4
    assert np.allclose(d.energies, np.round(E, 3))
    # a pivot p: x' = R (x - p) + p + t, so the matrix translation is t + p - R p
    d.pivots[2] = np.array([1.0, -2.0, 0.5])
    Mp = S.dat_to_mat(d)
    Rl = S.euler_to_rotation(*d.dofs[:, 1, :3].T)
    assert np.allclose(Mp[:, 3, :3], d.dofs[:, 1, 3:] + d.pivots[2] - Rl @ d.pivots[2])


def test_euler_round_trip_including_degenerate():
    rng = np.random.default_rng(6)
    e = rng.uniform(-7, 7, (500, 3))
    e[:3, 1] = [0.0, np.pi, -np.pi]
    R = S.euler_to_rotation(*e.T)
    assert np.allclose(R @ np.swapaxes(R, 1, 2), np.eye(3), atol=1e-12)
    assert np.abs(S.euler_to_rotation(*S.rotation_to_euler(R)) - R).max() < 1e-12


def test_attract_gradients_vs_attract_score():
    """JAX FP32 -dE/dM of the printed poses through the Euler chain rule reproduce ATTRACT's FP64
    Gradients lines to within the JAX-vs-ATTRACT differences (the translational columns
    show the same spread)."""
    g = np.load(os.path.join(GOLD, "score_pairs.npz"))
    n = len(g["euler"])
    assert n >= 6 and np.allclose(g["energy_jax"], g["energy_attract"], rtol=1e-4)
    dEdM = np.zeros((n, 4, 4))
    dEdM[:, :3, :3] = -g["neg_dEdR"]
    dEdM[:, 3, :3] = -g["neg_dEdt"]
    G = S.attract_gradients(g["euler"], dEdM)
    ref = g["gradients_attract"]
    assert np.all(np.abs(G - ref) <= 0.03 * np.abs(ref) + 0.3)
    assert np.all(np.sign(G[np.abs(ref) > 1]) == np.sign(ref[np.abs(ref) > 1]))


def test_attract_gradients_are_minus_euler_derivatives():
    """For E(M) = sum G.M, -dE/d(euler) by central differences equals attract_gradients."""
    rng = np.random.default_rng(7)
    Gm = rng.normal(size=(4, 4))
    e = rng.uniform(-3, 3, 3)

    def E(a):
        M = np.zeros((4, 4))
        M[:3, :3] = S.euler_to_rotation(*a).T
        return (Gm * M).sum()

    h = 1e-6
    fd = np.array([(E(e + h * np.eye(3)[k]) - E(e - h * np.eye(3)[k])) / (2 * h) for k in range(3)])
    ga = S.attract_gradients(e[None], Gm[None])[0]
    assert np.allclose(ga[:3], -fd, rtol=1e-7, atol=1e-8) and np.allclose(ga[3:], -Gm[3, :3])


def test_score_round_trip():
    E = np.array([-1632.0227230234073, 7.68988])
    G = np.arange(12.0).reshape(2, 6) - 5.5
    E2, G2 = S.read_score(S.write_score(E, G))
    assert np.array_equal(E2, E) and np.array_equal(G2, G)


def test_read_pdb_matches_reference_parse():
    p = S.read_pdb(_text("pdb_excerpt.pdb"))
    ex = np.load(os.path.join(GOLD, "pdb_expected.npz"))
    assert np.allclose(p["xyz"], ex["xyz"], atol=5e-4)
    assert list(p["name"]) == list(ex["name"])
    assert np.all(p["atomtype"] == 15) and p["resname"][0] == "SER"


def test_sort_and_top():
    E = np.array([3.0, -1.0, 2.0, -1.0, 0.5])
    assert S.sort_poses(E).tolist() == [1, 3, 4, 2, 0]
    assert S.sort_poses(E, 2).tolist() == [1, 3]
