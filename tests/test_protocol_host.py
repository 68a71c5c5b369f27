"""Host logic of the docking protocol (test5.py:88-146 preprocessing): type-99 removal,
alphabet compaction, pair-table slicing, grid-type positions, neighbour remapping."""
from collections import namedtuple

import numpy as np
import pytest

from attract_jax_amd.protocol import SCHEMA_TEST5, prepare, remap_grid_neighbours


def _par(rng):
    rc = rng.uniform(1, 5, (99, 99))
    ac = rng.uniform(1, 5, (99, 99))
    ivor = np.where(rng.random((99, 99)) < 0.2, -1.0, 1.0)
    return rc, ac, ivor


def test_prepare_compacts_types_and_drops_dummies():
    rng = np.random.default_rng(3)
    rec_types = np.array([5, 99, 12, 5, 99, 30, 12])
    lig_types = np.array([7, 99, 3, 7, 40])
    coor_rec = rng.normal(size=(7, 3))
    coor_lig = rng.normal(size=(2, 5, 3))  # a two-conformer library
    alphabet = np.zeros(99, bool)
    alphabet[[2, 6, 20, 39]] = True  # grid types 3, 7, 21, 40
    par = _par(rng)
    P = prepare(coor_rec, rec_types, coor_lig, lig_types, par, alphabet)
    assert np.array_equal(P.coor_rec, coor_rec[[0, 2, 3, 5, 6]])
    assert np.array_equal(P.rec_mapping[[0, 2, 3, 5, 6]], np.arange(5))
    assert np.array_equal(P.rec_atomtypes, [0, 1, 0, 2, 1])  # alphabet 5, 12, 30
    assert P.coor_lig.shape == (2, 4, 3) and np.array_equal(P.coor_lig[1], coor_lig[1][[0, 2, 3, 4]])
    assert np.array_equal(P.lig_atomtypes, [1, 0, 1, 2])  # alphabet 3, 7, 40
    assert np.array_equal(P.lig_atomtype_pos, [1, 0, 1, 3])  # positions of 7, 3, 7, 40 in (3, 7, 21, 40)
    rc, ac, ivor, emin, rmin2 = P.ff
    assert rc.shape == (3, 3) and rc[2, 1] == par[0][29, 6] and ivor[0, 2] == par[2][4, 39]
    assert np.allclose(emin, -27 * ac ** 4 / (256 * rc ** 3)) and np.allclose(rmin2, 4 * rc / (3 * ac))


def test_prepare_rejects_ligand_type_without_grid():
    rng = np.random.default_rng(4)
    alphabet = np.zeros(99, bool)
    alphabet[2] = True
    with pytest.raises(ValueError):
        prepare(rng.normal(size=(2, 3)), [1, 2], rng.normal(size=(2, 3)), [3, 4], _par(rng), alphabet)


def test_remap_neighbours_dense_and_csr():
    mapping = np.array([0, 0, 1, 2, 2, 3])  # beads 1 and 4 were type 99
    Dense = namedtuple("Dense", ["neighbour_grid", "nr_neighbours"])
    g = Dense(np.array([[[[0, 2, 65535], [5, 3, 65535]]]], np.uint16), np.array([[[2, 2]]]))
    r = remap_grid_neighbours(g, mapping)
    assert r.neighbour_grid.tolist() == [[[[0, 1, 65535], [3, 2, 65535]]]]
    Csr = namedtuple("Csr", ["neighbours", "neighbour_grid"])
    c = remap_grid_neighbours(Csr(np.array([5, 0, 3], np.uint32), None), mapping)
    assert c.neighbours.tolist() == [3, 0, 2] and c.neighbours.dtype == np.uint32


def test_schema_is_test5s_stage_list():
    assert [s[0] for s in SCHEMA_TEST5] == [5, 10, 10, 50, 100, 100, 100]
    assert sum(s[0] for s in SCHEMA_TEST5) == 375
    assert [s[2] for s in SCHEMA_TEST5] == [True] * 4 + [False] * 3
