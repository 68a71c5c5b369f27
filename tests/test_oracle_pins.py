"""Pins of the oracle: golden vectors, closed forms, invariants and finite differences
that a plausible mistake in the restatement (dropped term, wrong sign or index) fails."""
import json
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from tests.conftest import GOLDEN
from tests.systems import TinySystem, tiny_table


def test_nonbon_energy_init_golden():
    """energy-init/score.out: ATTRACT FP64 single-pair energies (energy-init/energy-init.py:35-40)."""
    d = json.load(open(os.path.join(GOLDEN, "energy_init.json")))
    ff = O.ff_from_par(d["rbc"], d["abc"], d["ipon"], d["potshape"])
    r = np.array(d["distance"])
    e = np.array(d["attract_energy"])
    E = O.nonbon(r * r, *ff)
    rel = np.abs(E - e) / np.abs(e)
    assert rel.max() < 1e-5
    assert np.median(rel) < 1e-6
    # both branches are exercised by the golden distances
    assert (r * r < ff.rmin2).any() and (r * r > ff.rmin2).any()


@pytest.mark.parametrize("ivor", [1.0, -1.0])
def test_nonbon_continuity_and_sign(ivor):
    rng = np.random.default_rng(1)
    for _ in range(20):
        rbc, abc = rng.uniform(2.5, 5.0), rng.uniform(1.0, 15.0)
        ff = O.ff_from_par(rbc, abc, ivor)
        r2 = ff.rmin2
        lo = O.nonbon(r2 * (1 - 1e-12), *ff)
        hi = O.nonbon(r2 * (1 + 1e-12), *ff)
        assert abs(lo - hi) < 1e-9 * max(1, abs(lo))
        assert np.isclose(hi, ivor * ff.emin, rtol=1e-9)
        d = np.linspace(0.5, 12, 400) ** 2
        E = O.nonbon(d, *ff)
        if ivor < 0:  # purely repulsive: positive and decreasing
            assert (E > 0).all() and (np.diff(E) < 0).all()
        else:  # minimum emin at rmin
            assert np.isclose(E.min(), ff.emin, rtol=1e-3)


def test_nonbon_dsq_grad_fd():
    rng = np.random.default_rng(2)
    ff = tiny_table(rng)
    for a in range(4):
        for b in range(4):
            p = [x[a, b] for x in ff]
            dsq = rng.uniform(4, 100, size=50)
            dsq = dsq[np.abs(dsq - p[4]) > 0.01]
            h = 1e-6 * dsq
            fd = (O.nonbon(dsq + h, *p) - O.nonbon(dsq - h, *p)) / (2 * h)
            assert np.allclose(O.nonbon_dsq_grad(dsq, *p), fd, rtol=1e-6, atol=1e-12)


def test_force_shift_is_gradient_of_shifted_energy():
    """ATTRACT-mode pair gradient 2 dnb/ddsq d - 2 pG d P/|d| (test-grid-energrad.py:140-146)
    equals the gradient of nonbon(|d|^2) - 2 pG P |d|."""
    sys_ = TinySystem(seed=3, nlig=1, npose=1)
    ff, P = sys_.ff, sys_.grid.plateaudis
    pE, pG = O.plateau_tables(ff, P)
    rng = np.random.default_rng(3)
    # one receptor bead at the origin, a probe point at distance < P
    grid = sys_.grid._replace(neighbours=np.array([0], np.uint32))
    ng = np.zeros_like(grid.neighbour_grid)
    ng[..., 0] = 1
    ng[..., 1] = 1
    grid = grid._replace(neighbour_grid=ng)
    for _ in range(10):
        x = rng.normal(size=3)
        x *= rng.uniform(1.5, P - 0.2) / np.linalg.norm(x)
        E, g, n = O.nb_energrad(x[None, None], [1], np.zeros((1, 3)), [2], ff, grid, plateau=True)
        assert n == 1
        f = lambda y: O.nonbon((y * y).sum(), *[a[2, 1] for a in ff]) - 2 * pG[2, 1] * P * np.linalg.norm(y)
        h = 1e-6
        fd = np.array([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.allclose(g[0, 0], fd, rtol=1e-5, atol=1e-8)
        assert np.isclose(E[0, 0], O.nonbon((x * x).sum(), *[a[2, 1] for a in ff]) - pE[2, 1])


def test_trilinear_affine_and_integer_quirk():
    rng = np.random.default_rng(4)
    dim = (6, 5, 7)
    a, b = rng.normal(size=4), rng.normal(size=(3, 4))
    ii = np.stack(np.meshgrid(*[np.arange(d) for d in dim], indexing="ij"), -1)
    g = (a + ii @ b)[None]
    v = rng.uniform(0.01, 3.99, size=(200, 3))
    v = v[np.all(np.abs(v - np.round(v)) > 1e-3, axis=1)]
    res, dres = O.trilinear(g, np.zeros(len(v), int), v, exact=True)
    assert np.allclose(res, a + v @ b, atol=1e-12)
    assert np.allclose(dres, np.broadcast_to(b[:, 0], dres.shape), atol=1e-12)
    vi = v.copy()
    vi[:, 1] = 2.0  # one integer coordinate: both weights vanish (test-grid-energrad.py:183-194)
    res, _ = O.trilinear(g, np.zeros(len(v), int), vi)
    assert np.all(res == 0)


def test_grid_gradient_sign_and_exact_mode():
    s = TinySystem(seed=5, nlig=6, npose=16)
    xp = O.transform(s.mat, s.coor_lig)
    raw, E, g = O.grid_energrad(xp, s.lig_types, s.grid)
    assert np.array_equal(g, -raw[..., 1:4])
    _, E2, ge = O.grid_energrad(xp, s.lig_types, s.grid, exact=True)
    h = 1e-5
    for k in range(3):
        dx = np.zeros(3)
        dx[k] = h
        Ep = O.grid_energrad(xp + dx, s.lig_types, s.grid)[1]
        Em = O.grid_energrad(xp - dx, s.lig_types, s.grid)[1]
        fd = (Ep - Em) / (2 * h)
        ok = np.isclose(ge[..., k], fd, rtol=1e-5, atol=1e-6)
        assert ok.mean() > 0.97  # voxel faces inside the +-h stencil are rare


def test_outside_grids_is_zero():
    s = TinySystem(seed=6, nlig=3, npose=4)
    far = s.mat.copy()
    far[:, 3, :3] = 500.0
    E, G = s.oracle_energrad(far)
    assert np.all(E == 0) and np.all(G == 0)


def test_as_matrix_and_from_matrix_vs_scipy():
    rng = np.random.default_rng(7)
    q = rng.normal(size=(100, 4))
    qn = q / np.linalg.norm(q, axis=1, keepdims=True)
    assert np.allclose(O.as_matrix(qn), Rotation.from_quat(qn).as_matrix(), atol=1e-12)
    from attract_jax_amd.pose import from_matrix
    R = Rotation.from_quat(qn).as_matrix()
    q2 = from_matrix(R)
    q3 = Rotation.from_matrix(R).as_quat()
    assert np.allclose(np.abs((q2 * q3).sum(1)), 1, atol=1e-12)
    assert np.allclose(O.as_matrix(q2), R, atol=1e-12)


def test_quat_grad_fd():
    rng = np.random.default_rng(8)
    q = rng.normal(size=(20, 4))
    A = rng.normal(size=(20, 3, 3))
    f = lambda qq: (O.as_matrix(qq) * A).sum(axis=(1, 2))
    g = O.quat_grad(q, A)
    h = 1e-6
    fd = np.stack([(f(q + h * e) - f(q - h * e)) / (2 * h) for e in np.eye(4)], 1)
    assert np.allclose(g, fd, rtol=1e-7, atol=1e-8)


def test_nb_term_vs_brute_force():
    """Lists cover every bead within plateaudis (neighbourdis >= plateaudis + half diagonal),
    so the list walk equals all pairs with dsq < P^2 (test-1B7F-explicit.py pattern)."""
    s = TinySystem(seed=9, nlig=7, npose=40)
    xp = O.transform(s.mat, s.coor_lig)
    inner = np.all((O.voxelize(xp, s.grid.origin, s.grid.gridspacing, s.grid.dim)[1]), axis=1)
    for plateau in (False, True):
        E, _, n = O.nb_energrad(xp, s.lig_types, s.coor_rec, s.rec_types, s.ff, s.grid, plateau=plateau)
        pE = O.plateau_tables(s.ff, s.grid.plateaudis)[0] if plateau else None
        bf = O.brute_force_energy(xp[inner], s.lig_types, s.coor_rec, s.rec_types, s.ff,
                                  cutoff2=s.grid.plateaudis ** 2, plateau=pE)
        assert n > 50
        assert np.allclose(E.sum(1)[inner], bf, rtol=1e-10, atol=1e-9)


def test_energrad_dEdM_exact_mode_fd():
    s = TinySystem(seed=10, nlig=5, npose=12)
    E, G = s.oracle_energrad(exact=True)
    h = 1e-6
    fd = np.zeros_like(G)
    for k in range(4):
        for l in range(3):
            mp, mm = s.mat.copy(), s.mat.copy()
            mp[:, k, l] += h
            mm[:, k, l] -= h
            fd[:, k, l] = (s.oracle_energrad(mp, exact=True)[0] - s.oracle_energrad(mm, exact=True)[0]) / (2 * h)
    scale = np.maximum(np.abs(fd).max(axis=(1, 2), keepdims=True), 1.0)
    ok = np.all(np.abs(G - fd) / scale < 1e-4, axis=(1, 2))
    assert ok.mean() >= 0.9


def test_dof_gradient_fd():
    from attract_jax_amd.pose import perturb_pose
    s = TinySystem(seed=11, nlig=5, npose=12)
    E, G = s.oracle_energrad(exact=True)
    dof = O.dof_gradient(s.mat, G)
    h = 1e-6
    fd = np.zeros((len(E), 6))
    for a in range(6):
        e = np.zeros(6)
        e[a] = h
        Ep = s.oracle_energrad(perturb_pose(s.mat, e[:3], e[3:]), exact=True)[0]
        Em = s.oracle_energrad(perturb_pose(s.mat, -e[:3], -e[3:]), exact=True)[0]
        fd[:, a] = (Ep - Em) / (2 * h)
    scale = np.maximum(np.abs(fd).max(axis=1, keepdims=True), 1.0)
    assert (np.all(np.abs(dof - fd) / scale < 1e-4, axis=1)).mean() >= 0.9


def test_adam_textbook_steps():
    """Kingma & Ba with optax defaults: step 1 moves each parameter by -lr*g/(|g|+eps);
    step 2 follows the bias-corrected moments."""
    g1 = np.array([[0.5, -2.0, 0.0, 1.0, 3.0, -0.1, 0.2]])
    g2 = np.array([[0.1, 1.0, 0.5, -1.0, 3.0, 0.3, 0.0]])
    seq = iter([g1, g2])
    q0 = np.array([[0.0, 0.0, 0.0, 1.0]])
    t0 = np.zeros((1, 3))
    lr = 0.01

    def fn(q, t):
        g = next(seq)
        # energy whose dE/dM gives the requested (dE/dq, dE/dt): pick A so quat_grad(q, A) = g_q
        # at q = identity-ish: use a least-squares solve of the linear map A -> quat_grad(q, A)
        J = np.stack([O.quat_grad(q, e.reshape(1, 3, 3))[0] for e in np.eye(9)], 1)
        A = np.linalg.lstsq(J, g[0, :4], rcond=None)[0].reshape(1, 3, 3)
        G = np.zeros((1, 4, 4))
        G[:, :3, :3] = A
        G[:, 3, :3] = g[:, 4:]
        return np.zeros(1), G

    # the quaternion components are renormalised; check translations exactly and quaternion direction
    q, t, _ = O.adam_minimize(q0, t0, 1, lr, lr, fn)
    exp_t = -lr * g1[0, 4:] / (np.abs(g1[0, 4:]) + 1e-8)
    assert np.allclose(t[0], exp_t, rtol=1e-12, atol=1e-15)
    seq = iter([g1, g2])
    q, t, _ = O.adam_minimize(q0, t0, 2, lr, lr, fn)
    m = 0.1 * g1 * 0.9 + 0.1 * g2
    v = 0.001 * g1 ** 2 * 0.999 + 0.001 * g2 ** 2
    upd = (m / (1 - 0.81)) / (np.sqrt(v / (1 - 0.999 ** 2)) + 1e-8)
    assert np.allclose(t[0], exp_t - lr * upd[0, 4:], rtol=1e-10, atol=1e-14)


def _translation_gradient_fn(g_t, energy=0.0):
    """Energy function with a constant translation gradient and no rotational gradient."""
    def fn(q, t):
        G = np.zeros((len(q), 4, 4))
        G[:, 3, :3] = g_t
        return np.full(len(q), energy), G
    return fn


def test_adabelief_constant_gradient_closed_form():
    """optax.adabelief under a constant gradient g: m_hat = g exactly, the prediction error
    of step k is g b1^(k-1) (taken against the previous first moment), so
    s_k = sum_j b2^(k-j) ((1-b2) g^2 b1^(2(j-1)) + eps_root) and each step moves t by
    -lr g / (sqrt(s_k / (1-b2^k)) + eps).  The quaternion has zero gradient and stays put."""
    g = np.array([0.7, -2.0, 0.05])
    b1, b2, eps, eps_root, lr, K = 0.9, 0.999, 1e-16, 1e-16, 0.5, 12
    q0 = np.array([[0.0, 0.0, 0.0, 1.0]])
    t0 = np.array([[1.0, 2.0, 3.0]])
    q, t, _ = O.optax_minimize(q0, t0, K, lr, lr, _translation_gradient_fn(g), "adabelief", b1, b2, eps, eps_root)
    exp = t0[0].copy()
    for k in range(1, K + 1):
        j = np.arange(1, k + 1)
        s = (b2 ** (k - j) * ((1 - b2) * g[:, None] ** 2 * b1 ** (2 * (j - 1)) + eps_root)).sum(axis=1)
        exp = exp - lr * g / (np.sqrt(s / (1 - b2 ** k)) + eps)
    assert np.allclose(t[0], exp, rtol=1e-12, atol=1e-12)
    assert np.allclose(q, q0)


def test_adam_constant_gradient_moves_lr_per_step():
    """optax.adam under a constant gradient: m_hat = g and v_hat = g^2 exactly, so every step
    moves each translation component by -lr g / (|g| + eps)."""
    g = np.array([0.7, -2.0, 0.05])
    q0 = np.array([[0.0, 0.0, 0.0, 1.0]])
    t0 = np.array([[1.0, 2.0, 3.0]])
    q, t, _ = O.optax_minimize(q0, t0, 9, 0.1, 0.1, _translation_gradient_fn(g), "adam")
    assert np.allclose(t[0], t0[0] - 9 * 0.1 * g / (np.abs(g) + 1e-8), rtol=1e-12)


@pytest.mark.parametrize("kind", ["adam", "adabelief"])
def test_gravity_first_step(kind):
    """Gravity restraint w |t|^2 (test5.py:418-420): the first evaluation reports w |t0|^2 and
    the first step sees the gradient 2 w t0."""
    w, lr = 0.1, 0.5
    q0 = np.array([[0.0, 0.0, 0.0, 1.0]])
    t0 = np.array([[3.0, -4.0, 12.0]])
    q, t, E = O.optax_minimize(q0, t0, 1, lr, lr, _translation_gradient_fn(np.zeros(3), energy=2.5), kind,
                               eps_root=1e-16 if kind == "adabelief" else 0.0, gravity=w)
    assert np.isclose(E[0], 2.5 + w * 169.0)
    g = 2 * w * t0[0]
    vh = g * g + (1e-16 / (1 - 0.999) if kind == "adabelief" else 0.0)
    assert np.allclose(t[0], t0[0] - lr * g / (np.sqrt(vh) + 1e-8), rtol=1e-12)


def test_rmsd_closed_form():
    rng = np.random.default_rng(12)
    x = rng.normal(size=(3, 9, 3))
    shift = np.array([1.0, -2.0, 2.0])
    assert np.allclose(O.rmsd(x, x + shift), 3.0)


def test_grid_point_potential_single_bead():
    ff = tiny_table(np.random.default_rng(13))
    P = 3.0
    pts = np.array([[4.0, 0, 0], [2.0, 0, 0], [0, 0, 60.0]])
    pot = O.grid_point_potential(pts, np.zeros((1, 3)), [1], 2, ff, P)
    p = [a[1, 2] for a in ff]
    assert np.isclose(pot[0, 0], O.nonbon(16.0, *p))
    assert np.isclose(pot[0, 1], -2 * O.nonbon_dsq_grad(16.0, *p) * 4.0)  # force = -dE/dx
    assert np.isclose(pot[1, 0], O.nonbon(P * P, *p))  # plateau cap
    assert np.isclose(pot[1, 1], -2 * O.nonbon_dsq_grad(P * P, *p) * P)  # projected plateau gradient
    assert np.all(pot[2] == 0)  # beyond the 50 A cutoff
