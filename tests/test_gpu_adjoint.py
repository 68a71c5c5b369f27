"""Differentiation (torch.autograd.Function + adjoint kernel) against FP64 central
finite differences of the oracle (h = 1e-4): pose DOFs and pair-table parameters."""
import numpy as np
import pytest
import torch

from attract_jax_amd import _lib
from attract_jax_amd.autodiff import dof_gradient, pair_table_from_rmin_emin, pose_energy
from attract_jax_amd.pose import mat_from_qt, perturb_pose
from attract_jax_amd.workloads import build_system, c1
from oracle import oracle as O
from tests.systems import smooth_poses

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
EXACT = _lib.ATJ_TERM_GRID | _lib.ATJ_TERM_NB | _lib.ATJ_NB_PLATEAU | _lib.ATJ_GRAD_EXACT


@pytest.fixture(scope="module")
def sysc1():
    w = c1(npose=256)
    system, ligs = build_system(w, DEV, keep_plain=True)
    grid = system.grid.to_reference_grid()
    return w, system, ligs, grid


def _oracle_total(w, grid, ff, mat, weights):
    l = w.ligands[0]
    E, _ = O.energrad(mat, l.coor, l.types - 1, l.types - 1, w.coor_rec, w.rec_types - 1, ff, grid, exact=True)
    return float((weights * E).sum()), E


def test_param_gradient_vs_fd(sysc1):
    w, system, ligs, grid = sysc1
    mat = mat_from_qt(w.quat, w.trans)
    # pull the poses towards the receptor so that many type pairs meet within the plateau
    mat[:, 3, :3] *= 0.55
    rng = np.random.default_rng(0)
    weights = rng.uniform(0.5, 1.5, len(mat))
    rmin = torch.tensor(w.rmin, dtype=torch.float64, device=DEV, requires_grad=True)
    eabs = torch.tensor(w.emin_abs, dtype=torch.float64, device=DEV, requires_grad=True)
    ivor = torch.tensor(w.ivor, dtype=torch.float64, device=DEV)
    rc, ac, iv, emin, rmin2 = pair_table_from_rmin_emin(rmin, eabs, ivor)
    E = pose_energy(torch.tensor(mat, dtype=torch.float32, device=DEV), rc, ac, iv, emin, rmin2, system, ligs[0],
                    flags=EXACT)
    L = (torch.tensor(weights, dtype=torch.float32, device=DEV) * E).sum()
    L.backward()
    g_rmin = rmin.grad.cpu().numpy()
    g_eabs = eabs.grad.cpu().numpy()
    idx = np.argsort(-np.abs(g_rmin).ravel())[:8]
    assert np.abs(g_rmin).max() > 0

    def total(rm, ea):
        ffp = O.ff_from_par(rm * np.sqrt(3) / 2, 256 / 27 * ea, w.ivor)
        return _oracle_total(w, grid, ffp, mat, weights)[0]

    for flat in idx:
        a, b = np.unravel_index(flat, g_rmin.shape)
        for arr, g in ((w.rmin, g_rmin), (w.emin_abs, g_eabs)):
            h = 1e-4 * max(1.0, abs(arr[a, b]))
            p_, m_ = arr.copy(), arr.copy()
            p_[a, b] += h
            m_[a, b] -= h
            if arr is w.rmin:
                fd = (total(p_, w.emin_abs) - total(m_, w.emin_abs)) / (2 * h)
            else:
                fd = (total(w.rmin, p_) - total(w.rmin, m_)) / (2 * h)
            assert abs(g[a, b] - fd) <= 2e-3 * max(abs(fd), 1e-2 * np.abs(g).max()), (a, b, g[a, b], fd)


def test_pose_dof_gradient_vs_fd(sysc1):
    w, system, ligs, grid = sysc1
    mat = mat_from_qt(w.quat[:64], w.trans[:64])
    mat[:, 3, :3] *= 0.6
    ff = system.ff
    M = torch.tensor(mat, dtype=torch.float32, device=DEV, requires_grad=True)
    E = pose_energy(M, ff.rc, ff.ac, ff.ivor, ff.emin, ff.rmin2, system, ligs[0], flags=EXACT)
    E.sum().backward()
    dof = dof_gradient(M.detach(), M.grad).cpu().numpy()
    fff = O.FF(*[t.cpu().numpy().astype(np.float64) for t in (ff.rc, ff.ac, ff.ivor, ff.emin, ff.rmin2)])
    l = w.ligands[0]
    h = 1e-4
    fd = np.zeros((len(mat), 6))
    for a in range(6):
        e = np.zeros(6)
        e[a] = h
        Ep = O.energrad(perturb_pose(mat, e[:3], e[3:]), l.coor, l.types - 1, l.types - 1, w.coor_rec,
                        w.rec_types - 1, fff, grid, exact=True)[0]
        Em = O.energrad(perturb_pose(mat, -e[:3], -e[3:]), l.coor, l.types - 1, l.types - 1, w.coor_rec,
                        w.rec_types - 1, fff, grid, exact=True)[0]
        fd[:, a] = (Ep - Em) / (2 * h)
    scale = np.maximum(np.abs(fd).max(axis=1, keepdims=True), 1.0)
    err = np.abs(dof - fd) / scale
    # the energy is smooth in the pose only between voxel faces, the inner/outer boundary and
    # the plateau sphere: poses whose beads keep 2e-3 A from all of them (the perturbations
    # move a bead by at most 1e-4 x its distance from the ligand centre) must all agree
    smooth = smooth_poses(mat, l.coor, w.coor_rec, grid, margin=2e-3)
    assert smooth.sum() >= 32, smooth.sum()
    assert err[smooth].max() < 2e-3, err[smooth].max()
    assert (err < 2e-3).mean() >= 0.9


def test_adjoint_deterministic_and_attract_mode(sysc1):
    """Repeated adjoint launches give the same bits; the ATTRACT gradient convention leaves the
    parameter gradient of the energy unchanged (the same bits too)."""
    w, system, ligs, grid = sysc1
    mat = mat_from_qt(w.quat, w.trans)
    mat[:, 3, :3] *= 0.55
    M = torch.tensor(mat, dtype=torch.float32, device=DEV)
    ff = system.ff
    outs = []
    for flags in (EXACT, EXACT, EXACT & ~_lib.ATJ_GRAD_EXACT):
        rc = ff.rc.clone().requires_grad_(True)
        E = pose_energy(M, rc, ff.ac, ff.ivor, ff.emin, ff.rmin2, system, ligs[0], flags=flags)
        E.sum().backward()
        outs.append(rc.grad.clone())
    assert outs[0].abs().max() > 0
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_c4_param_gradient_vs_fd():
    """The C4 system itself (2000-bead receptor, 96^3 + 64^3 grids, 21-bead ligand): the pair-table
    gradient of a weighted sum over 64 poses through the autograd Function, w.r.t. (rmin, |emin|)
    of the eight type pairs with the largest rmin gradient, against FP64 central differences
    (h = 1e-4) of the oracle's neighbour-list energy (the potential grids are data)."""
    from attract_jax_amd.workloads import c4
    w = c4(npose=64)
    system, ligs = build_system(w, DEV)
    g = system.grid
    mat = mat_from_qt(w.quat, w.trans)
    mat[:, 3, :3] *= 0.8  # into contact with the receptor surface
    rng = np.random.default_rng(3)
    weights = rng.uniform(0.5, 1.5, len(mat))
    rmin = torch.tensor(w.rmin, dtype=torch.float64, device=DEV, requires_grad=True)
    eabs = torch.tensor(w.emin_abs, dtype=torch.float64, device=DEV, requires_grad=True)
    ivor = torch.tensor(w.ivor, dtype=torch.float64, device=DEV)
    rc, ac, iv, emin, rmin2 = pair_table_from_rmin_emin(rmin, eabs, ivor)
    E = pose_energy(torch.tensor(mat, dtype=torch.float32, device=DEV), rc, ac, iv, emin, rmin2, system, ligs[0])
    (torch.tensor(weights, dtype=torch.float32, device=DEV) * E).sum().backward()
    g_rmin, g_eabs = rmin.grad.cpu().numpy(), eabs.grad.cpu().numpy()
    l = w.ligands[0]

    def total(rm, ea):
        ffp = O.ff_from_par(rm * np.sqrt(3) / 2, 256 / 27 * ea, w.ivor)
        En, _, _ = O.nb_energrad_complete(mat, l.coor, l.types - 1, w.coor_rec, w.rec_types - 1, ffp, g.origin,
                                          g.spacing, g.dim, g.plateaudis, plateau=True)
        return float((weights * En).sum())

    idx = np.argsort(-np.abs(g_rmin).ravel())[:8]
    assert np.abs(g_rmin).max() > 0
    errs = []
    for flat in idx:
        a, b = np.unravel_index(flat, g_rmin.shape)
        for arr, gr in ((w.rmin, g_rmin), (w.emin_abs, g_eabs)):
            h = 1e-4 * max(1.0, abs(arr[a, b]))
            p_, m_ = arr.copy(), arr.copy()
            p_[a, b] += h
            m_[a, b] -= h
            if arr is w.rmin:
                fd = (total(p_, w.emin_abs) - total(m_, w.emin_abs)) / (2 * h)
            else:
                fd = (total(w.rmin, p_) - total(w.rmin, m_)) / (2 * h)
            errs.append(abs(gr[a, b] - fd) / max(abs(fd), 1e-2 * np.abs(gr).max()))
            assert abs(gr[a, b] - fd) <= 5e-4 * max(abs(fd), 1e-2 * np.abs(gr).max()), (a, b, gr[a, b], fd)
    print(f"C4 pair-table gradient vs FP64 FD (8 type pairs x (rmin, |emin|)): rel max {max(errs):.1e} median "
          f"{np.median(errs):.1e}")


def test_adjoint_fixed_point_sums_do_not_depend_on_order(sysc1):
    """The block's fixed-point sums are exact for terms below 2^19: with one block, the poses
    spread over 1, 5 or 16 wavefronts (different orders of the adds) give the same bits.  The
    per-wavefront float tables (the form for pair tables too large for the fixed-point one)
    agree to float rounding."""
    from attract_jax_amd.autodiff import param_adjoint
    w, system, ligs, grid = sysc1
    mat = mat_from_qt(w.quat, w.trans)
    mat[:, 3, :3] *= 0.55
    M = torch.tensor(mat, dtype=torch.float32, device=DEV).reshape(-1, 16)
    rng = np.random.default_rng(5)
    wts = torch.tensor(rng.uniform(-1.0, 1.5, len(mat)), dtype=torch.float32, device=DEV)
    ff = system.ff
    one = [param_adjoint(system, ligs[0], M, ff, wts, EXACT, blocks=1, waves_per_block=n) for n in (1, 5, 16)]
    assert one[0].abs().max() > 0
    assert torch.equal(one[0], one[1]) and torch.equal(one[0], one[2])
    fixed = param_adjoint(system, ligs[0], M, ff, wts, EXACT, blocks=64, waves_per_block=16)
    waves = param_adjoint(system, ligs[0], M, ff, wts, EXACT, blocks=64, waves_per_block=16,
                          options=_lib.ATJ_LAUNCH_ADJOINT_WAVE_TABLES)
    scale = one[0].abs().max()
    assert ((fixed - one[0]).abs().max() / scale).item() < 1e-6
    assert ((waves - one[0]).abs().max() / scale).item() < 1e-5


def test_adjoint_terms_beyond_the_fixed_point_range(sysc1):
    """Terms of 2^19 and above go to the float sums beside the fixed-point words: with the pose
    weights scaled by 2^24 (every in-plateau term of magnitude >= 2^-5 then leaves the
    fixed-point range) the sums scale by 2^24 to float rounding, and the float sums agree with
    the per-wavefront float tables."""
    from attract_jax_amd.autodiff import param_adjoint
    w, system, ligs, grid = sysc1
    mat = mat_from_qt(w.quat, w.trans)
    mat[:, 3, :3] *= 0.55
    M = torch.tensor(mat, dtype=torch.float32, device=DEV).reshape(-1, 16)
    rng = np.random.default_rng(7)
    wts = torch.tensor(rng.uniform(0.5, 1.5, len(mat)), dtype=torch.float32, device=DEV)
    ff = system.ff
    base = param_adjoint(system, ligs[0], M, ff, wts, EXACT, blocks=64, waves_per_block=16)
    big = param_adjoint(system, ligs[0], M, ff, wts * 2.0 ** 24, EXACT, blocks=64, waves_per_block=16)
    waves = param_adjoint(system, ligs[0], M, ff, wts * 2.0 ** 24, EXACT, blocks=64, waves_per_block=16,
                          options=_lib.ATJ_LAUNCH_ADJOINT_WAVE_TABLES)
    scale = big.abs().max()
    assert scale > 2.0 ** 19  # the float path was taken
    assert torch.isfinite(big).all()
    assert ((big - base * 2.0 ** 24).abs().max() / scale).item() < 1e-5
    assert ((big - waves).abs().max() / scale).item() < 1e-5
