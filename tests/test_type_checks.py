"""Host-side type validation: receptor, ligand and point types beyond the pair table or the
potential grids are refused with ValueError before any device call.  CPU tensors stand in
for device memory; nothing here launches a kernel.  (The kernels also read such types as
the last row / column / grid, so a C caller that skips these checks cannot make them
address memory outside the tables.)"""
import types

import numpy as np
import pytest

from attract_jax_amd.device import ForceField, Ligand, Receptor


def _ff(nrt, nlt):
    one = np.ones((nrt, nlt))
    return ForceField.create(one, one, one, device="cpu")


def test_receptor_types():
    rec = Receptor.create(np.zeros((4, 3)), [0, 1, 2, 3], device="cpu")
    assert rec.max_type == 3
    rec.check_types(_ff(4, 2))
    with pytest.raises(ValueError, match="rows"):
        rec.check_types(_ff(3, 2))
    with pytest.raises(ValueError, match="out of range"):
        Receptor.create(np.zeros((1, 3)), [-1], device="cpu")


def test_ligand_types():
    lig = Ligand.create(np.zeros((3, 3)), [0, 5, 2], [1, 0, 3], device="cpu")
    assert (lig.max_grid_type, lig.max_ff_type) == (5, 3)
    grid = types.SimpleNamespace(has_potentials=True, ntypes=6)
    lig.check_types(grid, _ff(2, 4))
    with pytest.raises(ValueError, match="potential grids"):
        lig.check_types(types.SimpleNamespace(has_potentials=True, ntypes=5), _ff(2, 4))
    lig.check_types(types.SimpleNamespace(has_potentials=False, ntypes=0), _ff(2, 4))  # grid types unused
    with pytest.raises(ValueError, match="columns"):
        lig.check_types(grid, _ff(2, 3))
    with pytest.raises(ValueError, match="out of range"):
        Ligand.create(np.zeros((1, 3)), [0], [1 << 15], device="cpu")


def test_ligands_without_host_types_are_not_checked():
    lig = Ligand(None, 0)  # built from device data: the kernels' limits apply
    lig.check_types(types.SimpleNamespace(has_potentials=True, ntypes=1), _ff(1, 1))
