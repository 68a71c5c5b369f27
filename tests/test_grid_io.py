"""Grid reader / writer / pad_grid against the reference reader's golden vectors
(tests/golden/reader_golden.npz, produced by make_reader_golden.py) and round trips."""
import os
import struct

import numpy as np
import pytest

from attract_jax_amd.grid import GridFormatError, pad_grid, read_grid, read_grid2, write_grid
from tests.conftest import GOLDEN


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "reader_golden.npz"))


@pytest.fixture(scope="module")
def data():
    return open(os.path.join(GOLDEN, "reader_small.grid"), "rb").read()


def _compare(golden, prefix, mine):
    fields = list(golden[prefix + "__fields"])
    assert list(mine._fields) == fields
    for f in fields:
        v = getattr(mine, f)
        k = prefix + "__" + f
        if k not in golden:
            assert v is None, f
            continue
        ref = golden[k]
        a = np.asarray(v)
        assert a.shape == ref.shape, f
        assert a.dtype == ref.dtype, (f, a.dtype, ref.dtype)
        assert np.array_equal(a, ref), f


@pytest.mark.parametrize("grad", [True, False])
def test_read_grid_golden(golden, data, grad):
    _compare(golden, "rg_grad" if grad else "rg_nograd", read_grid(data, read_gradients=grad))


def test_read_grid2_golden(golden, data):
    _compare(golden, "rg2_grad", read_grid2(data, read_gradients=True))


def test_pad_grid_golden(golden, data):
    _compare(golden, "pad_list", pad_grid(read_grid(data, True), [2, 1, 3]))
    _compare(golden, "pad_int", pad_grid(read_grid(data, True), 2))


def test_write_read_round_trip():
    rng = np.random.default_rng(3)
    dim, dim2 = (7, 6, 5), (5, 5, 4)
    alphabet = np.zeros(99, bool)
    alphabet[[2, 3, 10, 20]] = True
    inner = rng.normal(size=(4, *dim, 4)).astype(np.float32)
    outer = rng.normal(size=(4, *dim2, 4)).astype(np.float32)
    counts = rng.integers(0, 6, size=dim)
    nb = rng.integers(0, 1000, size=counts.sum())
    data = write_grid(gridspacing=0.9, gridextension=16, plateaudis=10.0, neighbourdis=10.8, alphabet=alphabet,
                      origin=(1.5, -2.25, 3.0), dim=dim, dim2=dim2, natoms=1000, inner_potential_grid=inner,
                      outer_potential_grid=outer, nb_counts=counts, neighbours=nb)
    g = read_grid(data, read_gradients=True)
    assert np.array_equal(g.inner_potential_grid, inner)
    assert np.array_equal(g.outer_potential_grid, outer)
    assert np.array_equal(g.neighbour_grid[..., 0], counts)
    assert np.array_equal(g.neighbours, nb)
    assert g.gridspacing == 0.9 and g.gridextension == 16 and np.allclose(g.origin, (1.5, -2.25, 3.0))
    ng = g.neighbour_grid.reshape(-1, 2)
    c = ng[:, 0]
    for v in np.nonzero(c)[0][:20]:
        s = ng[v, 1] - 1
        assert np.array_equal(g.neighbours[s:s + c[v]], nb[np.cumsum(c)[v] - c[v]:np.cumsum(c)[v]])
    g2 = read_grid2(data, read_gradients=False)
    assert np.array_equal(g2.inner_potential_grid, inner[..., 0])
    assert g2.max_nr_neighbours == counts.max()


def test_nbgrid_without_potentials():
    counts = np.ones((3, 3, 3), int)
    data = write_grid(gridspacing=1.0, gridextension=0, plateaudis=5.0, neighbourdis=7.0,
                      alphabet=np.zeros(99, bool), origin=(0, 0, 0), dim=(3, 3, 3), dim2=(0, 0, 0), natoms=27,
                      nb_counts=counts, neighbours=np.arange(27))
    g = read_grid(data)
    assert "inner_potential_grid" not in g._fields and "outer_potential_grid" not in g._fields
    assert np.array_equal(g.neighbours, np.arange(27))


def test_reader_errors(data):
    bad = bytearray(data)
    bad[0] = 1  # torque grid
    with pytest.raises(AssertionError):
        read_grid(bytes(bad))
    bad = bytearray(data)
    struct.pack_into("<h", bad, 1, 32)  # arch 32
    with pytest.raises(GridFormatError):
        read_grid(bytes(bad))
    with pytest.raises(ValueError):
        read_grid(data[:100])
    # too many atoms for the dense uint16 table (read_grid2.py:86)
    counts = np.ones((2, 2, 2), int)
    big = write_grid(gridspacing=1.0, gridextension=0, plateaudis=5.0, neighbourdis=7.0,
                     alphabet=np.zeros(99, bool), origin=(0, 0, 0), dim=(2, 2, 2), dim2=(0, 0, 0), natoms=70000,
                     nb_counts=counts, neighbours=np.full(8, 70000))
    with pytest.raises(AssertionError):
        read_grid2(big)
