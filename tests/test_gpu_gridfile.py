"""The drop-in file path on the GPU: an ATTRACT grid file (written by grid.write_grid from a
grid built on the device) read back by read_grid2 (the dense-table reader test5.py uses,
test5.py:32; read_grid2.py:113-124) and by read_grid + pad_grid (read_grid.py:141-204), then
scored and minimized by the kernels through the reference-named calls.  Scores equal those
of the grid the file was written from, bit for bit, and the FP64 oracle's on the read grid."""
import numpy as np
import pytest
import torch

import attract_jax_amd as A
from attract_jax_amd import grid as G
from attract_jax_amd.minimize import BatchPlan, PoseState, minimize
from attract_jax_amd.pose import mat_from_qt
from attract_jax_amd.device import DeviceGrid
from attract_jax_amd.score import ScoringSystem
from attract_jax_amd.workloads import build_system, c1
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _file_of(ref, potentials=True):
    counts = ref.neighbour_grid[..., 0]
    return G.write_grid(gridspacing=ref.gridspacing, gridextension=ref.gridextension, plateaudis=ref.plateaudis,
                        neighbourdis=ref.neighbourdis, alphabet=ref.alphabet, origin=ref.origin, dim=ref.dim,
                        dim2=ref.dim2, natoms=300,
                        inner_potential_grid=ref.inner_potential_grid if potentials else None,
                        outer_potential_grid=ref.outer_potential_grid if potentials else None,
                        nb_counts=counts, neighbours=ref.neighbours)


@pytest.fixture(scope="module")
def c1sys():
    w = c1(npose=512)
    system, ligs = build_system(w, DEV, keep_plain=True, nshell=0)
    return w, system, ligs, system.grid.to_reference_grid()


@pytest.mark.parametrize("reader", ["read_grid2", "read_grid+pad_grid"])
def test_grid_file_through_the_reference_readers(c1sys, reader):
    w, system, ligs, ref = c1sys
    data = bytes(_file_of(ref))
    if reader == "read_grid2":
        g = A.read_grid2(data, read_gradients=True)
        assert g.neighbour_grid.dtype == np.uint16 and g.neighbour_grid.shape[:3] == tuple(ref.dim)
    else:
        g = A.pad_grid(A.read_grid(data, read_gradients=True), padding=[0] * 5 + [3] * 10 + [7])
    lig = w.ligands[0]
    mat = mat_from_qt(w.quat, w.trans).astype(np.float32)
    ff = O.ff_complete(w.rc, w.ac, w.ivor)
    # the grid the file came from and the grid read from the file, both read whole (no shell
    # order): the same bits
    E0, G0 = system.energrad(mat, ligs[0])
    sys_file = ScoringSystem(DeviceGrid.from_grid(g), system.ff, system.receptor, nshell=0)
    E1, G1 = sys_file.energrad(mat, ligs[0])
    assert torch.equal(E0, E1) and torch.equal(G0, G1)
    (tot, E), Gm = A.value_and_grad_main(mat, w.coor_rec, w.rec_types - 1, lig.coor, lig.types - 1, lig.types - 1,
                                         ff, g)
    # the reference-named call (shell-ordered lists): same pairs, other summation order
    assert (np.abs(E.cpu().numpy() - E0.cpu().numpy()) / np.maximum(np.abs(E0.cpu().numpy()), 1.0)).max() < 1e-5
    # and the FP64 oracle on the grid read from the file
    Eo, Go = O.energrad(mat.astype(np.float64), lig.coor, lig.types - 1, lig.types - 1, w.coor_rec,
                        w.rec_types - 1, ff, g)
    assert (np.abs(E.cpu().numpy() - Eo) / np.maximum(np.abs(Eo), 1.0)).max() < 1e-4
    sc = np.maximum(np.abs(Go).max(axis=(1, 2), keepdims=True), 1.0)
    assert (np.abs(Gm.cpu().numpy() - Go) / sc).max() < 1e-4


def test_nbgrid_file_and_minimization_from_file(c1sys):
    """A neighbour-only .nbgrid (no energrads) through read_grid2 gives the neighbour-list
    energies without plateau; a full grid file read by read_grid2 minimizes exactly as the
    grid it was written from."""
    w, system, ligs, ref = c1sys
    g_nb = A.read_grid2(bytes(_file_of(ref, potentials=False)))
    assert not hasattr(g_nb, "inner_potential_grid")
    lig = w.ligands[0]
    mat = mat_from_qt(w.quat, w.trans).astype(np.float32)
    ff = O.ff_complete(w.rc, w.ac, w.ivor)
    _, E = A.main(mat, w.coor_rec, w.rec_types - 1, lig.coor, lig.types - 1, lig.types - 1, ff, g_nb)
    xp = O.transform(mat.astype(np.float64), lig.coor)
    En, _, _ = O.nb_energrad(xp, lig.types - 1, w.coor_rec, w.rec_types - 1, ff, g_nb, plateau=False)
    assert (np.abs(E.cpu().numpy() - En.sum(1)) / np.maximum(np.abs(En.sum(1)), 1.0)).max() < 1e-4

    g = A.read_grid2(bytes(_file_of(ref)), read_gradients=True)
    sys_file = ScoringSystem(DeviceGrid.from_grid(g), system.ff, system.receptor, nshell=0)
    out = []
    for s in (system, sys_file):
        plan = BatchPlan(s, ligs, [w.npose], 30, w.lr_rot, w.lr_trans)
        st = PoseState.create(w.quat, w.trans, DEV)
        minimize(s, plan, st, 30)
        out.append(st)
    torch.cuda.synchronize()
    assert torch.equal(out[0].pose, out[1].pose) and torch.equal(out[0].energy, out[1].energy)
