"""FP64 NumPy restatement of the reference's scoring + minimization semantics.

TEST INFRASTRUCTURE ONLY.  This module is the correctness oracle: only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it, and only
to check or time against it.  The product path (attract_jax_amd) never calls it.

Each function cites the reference file:line it restates (sjdv1982/attract-jax).
Pins (what fixes each function independently of this restatement) are listed in
PINS below; functions whose pin is weaker than a golden vector say so.
"""
from collections import namedtuple

import numpy as np

PINS = {
    "nonbon": "energy-init golden (296 ATTRACT FP64 single-pair energies) + rmin2 continuity",
    "nonbon_dsq_grad": "central finite differences of nonbon",
    "plateau_tables": "closed form at plateaudis**2 (nonbon / nonbon_dsq_grad)",
    "transform": "identity / pure translation invariants; as_matrix vs scipy.spatial.transform",
    "trilinear": "affine fields reproduced exactly; integer coordinate -> 0; outside -> 0",
    "grid_energrad": "same as trilinear + gradient sign convention (-interpolated stored gradient)",
    "nb_energrad": "brute-force all-pairs within plateaudis on tiny systems; force-shift closed form; "
                   "the reference's own outputs test-nb-energrad.out / test-nb-energrad-plateau.out on "
                   "mono-leu (nb_energrad_complete)",
    "nb_energrad_complete": "test-nb-energrad.out and test-nb-energrad-plateau.out (10 poses: energies, "
                            "-dE/dM translation rows and torque blocks) on the mono-leu system",
    "energrad": "dE/dM = sum_j xh_j (x) g_j closed form; exact mode vs FP64 central differences",
    "optax_minimize": "adam: textbook Kingma & Ba steps and the constant-gradient property "
                      "(each step moves -lr g/(|g|+eps)); adabelief: constant-gradient closed form of "
                      "the belief recurrence; gravity: first-step closed form.  Against optax itself "
                      "parity unpinned (optax absent, version unpinned)",
    "rmsd": "closed form; test-rmsd.lrmsd (mono-leu poses, 3 decimals)",
    "read_grid": "golden vectors from the reference reader on a synthetic grid file",
    "build_grid": "brute force over receptor beads at sampled grid points",
}

FF = namedtuple("FF", ["rc", "ac", "ivor", "emin", "rmin2"])


# --------------------------------------------------------------------------- pair potential

def ff_from_par(rbc, abc, ipon, potshape=8):
    """rc = abc rbc^8, ac = abc rbc^6, ivor = ipon (attract-original/convert-attract-par.py:16-22);
    emin, rmin2 derived as in test-grid-energrad.py:71-72."""
    rbc = np.asarray(rbc, np.float64)
    abc = np.asarray(abc, np.float64)
    rc = abc * rbc ** potshape
    ac = abc * rbc ** 6
    return ff_complete(rc, ac, ipon)


def ff_complete(rc, ac, ivor):
    rc = np.asarray(rc, np.float64)
    ac = np.asarray(ac, np.float64)
    emin = -27.0 * ac ** 4 / (256.0 * rc ** 3)
    rmin2 = 4.0 * rc / (3.0 * ac)
    return FF(rc, ac, np.asarray(ivor, np.float64), emin, rmin2)


def nonbon(dsq, rc, ac, ivor, emin, rmin2):
    """ATTRACT 8-6 potential of squared distance (test-grid-energrad.py:95-111)."""
    rr2 = 1.0 / dsq
    rr23 = rr2 * rr2 * rr2
    vlj = (rc * rr2 - ac) * rr23
    return np.where(dsq < rmin2, vlj + (ivor - 1) * emin, ivor * vlj)


def nonbon_dsq_grad(dsq, rc, ac, ivor, emin, rmin2):
    """d nonbon / d dsq (jax.grad of nonbon, test-grid-energrad.py:113)."""
    u = 1.0 / dsq
    dv = -(4 * rc * u - 3 * ac) * u ** 4
    return np.where(dsq < rmin2, dv, ivor * dv)


def plateau_tables(ff, plateaudis):
    """plateau energies / dsq-gradients of every type pair (test-grid-energrad.py:124-127)."""
    p2 = plateaudis ** 2
    return nonbon(p2, *ff), nonbon_dsq_grad(p2, *ff)


# --------------------------------------------------------------------------- poses

def transform(mat, coor_lig):
    """x' = [x, 1] . mat for every pose (test-grid-energrad.py:299-301)."""
    coor_lig = np.asarray(coor_lig, np.float64)
    xh = np.concatenate([coor_lig, np.ones((len(coor_lig), 1))], axis=1)
    return np.einsum("jk,ikl->ijl", xh, np.asarray(mat, np.float64))[:, :, :3]


def as_matrix(q):
    """jax.scipy / scipy Rotation.as_matrix polynomial for scalar-last quaternions (test4.py:399)."""
    q = np.asarray(q, np.float64)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = x * x - y * y - z * z + w * w
    R[..., 0, 1] = 2 * (x * y - z * w)
    R[..., 0, 2] = 2 * (x * z + y * w)
    R[..., 1, 0] = 2 * (x * y + z * w)
    R[..., 1, 1] = -x * x + y * y - z * z + w * w
    R[..., 1, 2] = 2 * (y * z - x * w)
    R[..., 2, 0] = 2 * (x * z - y * w)
    R[..., 2, 1] = 2 * (y * z + x * w)
    R[..., 2, 2] = -x * x - y * y + z * z + w * w
    return R


def mat_from_qt(q, t):
    """test4.py:398-400: ones((N,4,4)), rotation block, translation row."""
    q = np.asarray(q, np.float64)
    M = np.ones(q.shape[:-1] + (4, 4))
    M[..., :3, :3] = as_matrix(q)
    M[..., 3, :3] = t
    return M


def quat_grad(q, A):
    """dE/dq through as_matrix given A = dE/dR (chain rule of the polynomial)."""
    q = np.asarray(q, np.float64)
    J = np.empty(q.shape[:-1] + (4, 3, 3))
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    # dR/dx
    J[..., 0, :, :] = 2 * np.stack([np.stack([x, y, z], -1), np.stack([y, -x, -w], -1), np.stack([z, w, -x], -1)], -2)
    J[..., 1, :, :] = 2 * np.stack([np.stack([-y, x, w], -1), np.stack([x, y, z], -1), np.stack([-w, z, -y], -1)], -2)
    J[..., 2, :, :] = 2 * np.stack([np.stack([-z, -w, x], -1), np.stack([w, -z, y], -1), np.stack([x, y, z], -1)], -2)
    J[..., 3, :, :] = 2 * np.stack([np.stack([w, -z, y], -1), np.stack([z, w, -x], -1), np.stack([-y, x, w], -1)], -2)
    return np.einsum("...ikl,...kl->...i", J, A)


# --------------------------------------------------------------------------- grid term

def voxelize(xp, origin, spacing, dim):
    """v = (x' - origin)/spacing; inside = 0 <= v < dim-1 (test-grid-energrad.py:216-218)."""
    v = (np.asarray(xp, np.float64) - np.asarray(origin, np.float64)) / spacing
    dim = np.asarray(dim)
    inside = np.all((v >= 0) & (v < dim - 1), axis=-1)
    return v, inside


def trilinear(pgrid, t, v, exact=False):
    """Interpolate pgrid[t] (ntypes, dx, dy, dz, 4) at voxel coords v (M,3)
    (eval_potential_grid, test-grid-energrad.py:183-194).  Returns (M,4) and,
    with exact=True, the derivative of the E channel w.r.t. v (M,3)."""
    v = np.asarray(v, np.float64)
    v0 = np.floor(v).astype(np.int64)
    v1 = np.ceil(v).astype(np.int64)
    w = [v1 - v, v - v0]
    vox = [v0, v1]
    res = np.zeros((len(v), 4))
    dres = np.zeros((len(v), 3))
    for cx in (0, 1):
        for cy in (0, 1):
            for cz in (0, 1):
                wgt = w[cx][:, 0] * w[cy][:, 1] * w[cz][:, 2]
                val = pgrid[t, vox[cx][:, 0], vox[cy][:, 1], vox[cz][:, 2]]
                res += wgt[:, None] * val
                if exact:
                    s = [-1.0, 1.0]
                    dres[:, 0] += s[cx] * w[cy][:, 1] * w[cz][:, 2] * val[:, 0]
                    dres[:, 1] += s[cy] * w[cx][:, 0] * w[cz][:, 2] * val[:, 0]
                    dres[:, 2] += s[cz] * w[cx][:, 0] * w[cy][:, 1] * val[:, 0]
    return res, dres


def grid_energrad(xp, gtypes, grid, exact=False):
    """Per-bead grid term (potential_energrad + potential_energies JVP,
    test-grid-energrad.py:214-239).  xp (N, nlig, 3), gtypes (nlig,).
    Returns raw (N,nlig,4) interpolation, E (N,nlig), dE/dx' (N,nlig,3)."""
    N, nl = xp.shape[:2]
    flat = xp.reshape(-1, 3)
    t = np.broadcast_to(np.asarray(gtypes), (N, nl)).reshape(-1)
    v, inside = voxelize(flat, grid.origin, grid.gridspacing, grid.dim)
    raw = np.zeros((len(flat), 4))
    dEdv = np.zeros((len(flat), 3))
    scale = np.zeros(len(flat))
    if inside.any():
        r, d = trilinear(grid.inner_potential_grid, t[inside], v[inside], exact)
        raw[inside], dEdv[inside], scale[inside] = r, d, 1.0 / grid.gridspacing
    outer = getattr(grid, "outer_potential_grid", None)
    if outer is not None:
        vo = (v + grid.gridextension) / 2
        in_outer = ~inside & np.all((vo >= 0) & (vo < np.asarray(grid.dim2) - 1), axis=-1)
        if in_outer.any():
            r, d = trilinear(outer, t[in_outer], vo[in_outer], exact)
            raw[in_outer], dEdv[in_outer], scale[in_outer] = r, d, 0.5 / grid.gridspacing
    E = raw[:, 0]
    g = dEdv * scale[:, None] if exact else -raw[:, 1:4]
    return raw.reshape(N, nl, 4), E.reshape(N, nl), g.reshape(N, nl, 3)


# --------------------------------------------------------------------------- neighbour term

def nb_voxel(xp, grid):
    """Flat nearest-voxel index of the neighbour lookup, -1 outside the inner grid
    (generate_nb_table, test-grid-energrad.py:246-253)."""
    v, inside = voxelize(xp, grid.origin, grid.gridspacing, grid.dim)
    pos = np.floor(v + 0.5).astype(np.int64)
    dim = np.asarray(grid.dim)
    flat = (pos[..., 0] * dim[1] + pos[..., 1]) * dim[2] + pos[..., 2]
    return np.where(inside, flat, -1)


def csr_lists(grid):
    """(counts, 0-based starts, flat receptor indices) of a reference Grid (CSR or dense)."""
    if hasattr(grid, "nr_neighbours"):
        counts = np.asarray(grid.nr_neighbours).reshape(-1).astype(np.int64)
        table = np.asarray(grid.neighbour_grid)
        k = table.shape[-1]
        return counts, np.arange(len(counts)) * k, table.reshape(-1).astype(np.int64)
    ng = np.asarray(grid.neighbour_grid).reshape(-1, 2)
    counts = ng[:, 0].astype(np.int64)
    return counts, ng[:, 1].astype(np.int64) - 1, np.asarray(grid.neighbours).astype(np.int64)


def nb_energrad(xp, lig_fftypes, coor_rec, rec_fftypes, ff, grid, plateau=True, exact=False):
    """Per-bead neighbour-list energy and gradient (nb_energy_single + nonbon_dif JVP,
    test-grid-energrad.py:129-162; plateau=False: test-nb-energrad.py:142-145).
    Padding sentinels (65535 / 0xFFFFFFFF, read_grid2 / pad_grid) are skipped.
    Returns E (N,nlig), dE/dx' (N,nlig,3), evaluated-pair count."""
    N, nl = xp.shape[:2]
    flat = np.asarray(xp, np.float64).reshape(-1, 3)
    lt = np.broadcast_to(np.asarray(lig_fftypes), (N, nl)).reshape(-1)
    vox = nb_voxel(flat, grid)
    counts, starts, lists = csr_lists(grid)
    item = np.nonzero(vox >= 0)[0]
    c = counts[vox[item]]
    c = np.maximum(c, 0)
    pair_item = np.repeat(item, c)
    within = np.arange(c.sum()) - np.repeat(np.cumsum(c) - c, c)
    r = lists[np.repeat(starts[vox[item]], c) + within]
    valid = (r < 65535) & (r < len(coor_rec))
    pair_item, r = pair_item[valid], r[valid]
    d = flat[pair_item] - np.asarray(coor_rec, np.float64)[r]
    dsq = (d * d).sum(axis=1)
    P2 = grid.plateaudis ** 2
    keep = dsq < P2
    pair_item, r, d, dsq = pair_item[keep], r[keep], d[keep], dsq[keep]
    at1 = np.asarray(rec_fftypes)[r]
    at2 = lt[pair_item]
    prm = [np.asarray(a, np.float64)[at1, at2] for a in ff]
    e = nonbon(dsq, *prm)
    gs = 2 * nonbon_dsq_grad(dsq, *prm)
    if plateau:
        pE, pG = plateau_tables(ff, grid.plateaudis)
        e = e - pE[at1, at2]
        if not exact:
            gs = gs - 2 * pG[at1, at2] * np.sqrt(P2 / dsq)
    E = np.zeros(len(flat))
    G = np.zeros((len(flat), 3))
    np.add.at(E, pair_item, e)
    np.add.at(G, pair_item, gs[:, None] * d)
    return E.reshape(N, nl), G.reshape(N, nl, 3), len(e)


def brute_force_energy(xp, lig_fftypes, coor_rec, rec_fftypes, ff, cutoff2=None, plateau=None):
    """All-pairs energy per pose (test-1B7F-explicit.py:66-75); with cutoff2 only pairs
    with dsq < cutoff2 count, and plateau=(pE table) subtracts it per counted pair."""
    xp = np.asarray(xp, np.float64)
    d = xp[:, :, None, :] - np.asarray(coor_rec, np.float64)[None, None]
    dsq = (d * d).sum(-1)
    at1 = np.asarray(rec_fftypes)[None, None, :]
    at2 = np.asarray(lig_fftypes)[None, :, None]
    prm = [np.asarray(a, np.float64)[at1, at2] for a in ff]
    e = nonbon(dsq, *prm)
    if plateau is not None:
        e = e - plateau[at1, at2]
    if cutoff2 is not None:
        e = np.where(dsq < cutoff2, e, 0.0)
    return e.sum(axis=(1, 2))


# --------------------------------------------------------------------------- pose energy / gradient

def energrad(mat, coor_lig, gtypes, lig_fftypes, coor_rec, rec_fftypes, ff, grid, terms=("grid", "nb"),
             plateau=None, exact=False):
    """Energies (N,) and dE/dmat (N,4,4) (main + value_and_grad, test-grid-energrad.py:298-317).
    dE/dM[k,l] = sum_j xh_jk (dE/dx'_j)_l; column 3 is zero."""
    xp = transform(mat, coor_lig)
    N, nl = xp.shape[:2]
    E = np.zeros(N)
    g = np.zeros((N, nl, 3))
    has_pot = getattr(grid, "inner_potential_grid", None) is not None
    if "grid" in terms and has_pot:
        _, Eg, gg = grid_energrad(xp, gtypes, grid, exact)
        E += Eg.sum(axis=1)
        g += gg
    if "nb" in terms:
        pl = has_pot if plateau is None else plateau
        En, gn, _ = nb_energrad(xp, lig_fftypes, coor_rec, rec_fftypes, ff, grid, pl, exact)
        E += En.sum(axis=1)
        g += gn
    xh = np.concatenate([np.asarray(coor_lig, np.float64), np.ones((nl, 1))], axis=1)
    dEdM = np.zeros((N, 4, 4))
    dEdM[:, :, :3] = np.einsum("jk,ijl->ikl", xh, g)
    return E, dEdM


def nb_energrad_complete(mat, coor_lig, lig_fftypes, coor_rec, rec_fftypes, ff, origin, spacing, dim, plateaudis,
                         plateau=False, exact=False):
    """Neighbour term of energrad for a grid whose voxel lists are complete: every list holds
    each receptor bead within plateaudis + half the voxel diagonal of its grid point (ATTRACT
    lists built with neighbourdis at least that).  The list walk of nb_energy_single
    (test-nb-energrad.py:148-158; plateau variant test-grid-energrad.py:129-159) then evaluates
    exactly the pairs (bead inside the inner grid, receptor bead with dsq < plateaudis^2), so
    they are found here with a k-d tree instead of lists.  Returns energies (N,), dE/dM
    (N,4,4) and the sum of |pair energy| per pose (the scale of an energy that is a sum of
    terms of both signs)."""
    from scipy.spatial import cKDTree
    xp = transform(mat, coor_lig)
    N, nl = xp.shape[:2]
    coor_rec = np.asarray(coor_rec, np.float64)
    tree = cKDTree(coor_rec)
    lt = np.asarray(lig_fftypes)
    rt = np.asarray(rec_fftypes)
    P2 = plateaudis ** 2
    pE, pG = plateau_tables(ff, plateaudis)
    E = np.zeros(N)
    S = np.zeros(N)
    g = np.zeros((N, nl, 3))
    for i in range(N):
        _, inside = voxelize(xp[i], origin, spacing, dim)
        idx = np.nonzero(inside)[0]
        lists = tree.query_ball_point(xp[i][idx], plateaudis)
        c = np.array([len(l) for l in lists], np.int64)
        if c.sum() == 0:
            continue
        j = np.repeat(idx, c)
        r = np.concatenate([np.asarray(l, np.int64) for l in lists])
        d = xp[i][j] - coor_rec[r]
        dsq = (d * d).sum(1)
        keep = dsq < P2
        j, r, d, dsq = j[keep], r[keep], d[keep], dsq[keep]
        at1, at2 = rt[r], lt[j]
        prm = [np.asarray(a, np.float64)[at1, at2] for a in ff]
        e = nonbon(dsq, *prm)
        gs = 2 * nonbon_dsq_grad(dsq, *prm)
        if plateau:
            e = e - pE[at1, at2]
            if not exact:
                gs = gs - 2 * pG[at1, at2] * np.sqrt(P2 / dsq)
        E[i] = e.sum()
        S[i] = np.abs(e).sum()
        np.add.at(g[i], j, gs[:, None] * d)
    xh = np.concatenate([np.asarray(coor_lig, np.float64), np.ones((nl, 1))], axis=1)
    dEdM = np.zeros((N, 4, 4))
    dEdM[:, :, :3] = np.einsum("jk,ijl->ikl", xh, g)
    return E, dEdM, S


def dof_gradient(mat, dEdM):
    """(dE/dt, dE/domega): omega = world-frame rotation vector about the ligand centre;
    dE/domega_a = eps_abc (R^T A)_bc with A = dE/dM[0:3,0:3]."""
    R = np.asarray(mat, np.float64)[:, :3, :3]
    A = np.asarray(dEdM, np.float64)[:, :3, :3]
    B = np.einsum("nkb,nkc->nbc", R, A)
    tau = np.stack([B[:, 1, 2] - B[:, 2, 1], B[:, 2, 0] - B[:, 0, 2], B[:, 0, 1] - B[:, 1, 0]], axis=1)
    return np.concatenate([np.asarray(dEdM)[:, 3, :3], tau], axis=1)


# --------------------------------------------------------------------------- minimization

def optax_minimize(q, t, nsteps, lr_rot, lr_trans, energy_fn, kind="adam", beta1=0.9, beta2=0.999, eps=1e-8,
                   eps_root=0.0, gravity=0.0):
    """Optimizer loops of test4.py:445-455 (optax.adam) and test5.py:450-458 (optax.adabelief)
    on (quaternion, translation), the published optax algorithms (version unpinned):
      adam:      m = (1-b1) g + b1 m;  v = (1-b2) g^2 + b2 v
      adabelief: m = (1-b1) g + b1 m;  s = (1-b2) (g - m_prev)^2 + b2 s + eps_root
                 (optax's scale_by_belief takes the prediction error against the previous
                 first moment and adds eps_root to the stored second moment)
      p += -lr (m / (1 - b1^k)) / (sqrt(v / (1 - b2^k)) + eps); q renormalised after each step.
    gravity > 0: energy += gravity |t|^2, dE/dt += 2 gravity t (test5.py:418-420).
    energy_fn(q, t) -> (E (N,), dE/dM (N,4,4)).  Returns final q, t and the energies of
    the last evaluation (restraint included)."""
    q = np.array(q, np.float64)
    t = np.array(t, np.float64)
    N = len(q)
    m = np.zeros((N, 7))
    v = np.zeros((N, 7))
    E = np.zeros(N)
    lr = np.array([lr_rot] * 4 + [lr_trans] * 3)
    for k in range(1, nsteps + 1):
        E, G = energy_fn(q, t)
        g = np.concatenate([quat_grad(q, G[:, :3, :3]), G[:, 3, :3]], axis=1)
        if gravity:
            E = E + gravity * (t * t).sum(axis=1)
            g[:, 4:] += 2 * gravity * t
        pe = g - m if kind == "adabelief" else g
        m = (1 - beta1) * g + beta1 * m
        v = (1 - beta2) * pe * pe + beta2 * v
        if kind == "adabelief":
            v = v + eps_root
        mh = m / (1 - beta1 ** k)
        vh = v / (1 - beta2 ** k)
        p = np.concatenate([q, t], axis=1) - lr * mh / (np.sqrt(vh) + eps)
        q = p[:, :4] / np.linalg.norm(p[:, :4], axis=1, keepdims=True)
        t = p[:, 4:]
    return q, t, E


def adam_minimize(q, t, nsteps, lr_rot, lr_trans, energy_fn, beta1=0.9, beta2=0.999, eps=1e-8):
    """optax.adam loop of test4.py:445-455 (optax_minimize with kind="adam")."""
    return optax_minimize(q, t, nsteps, lr_rot, lr_trans, energy_fn, "adam", beta1, beta2, eps)


def rmsd(xa, xb):
    """Ligand RMSD sqrt(sum |xa - xb|^2 / n) per pose (test-rmsd.py:13-17)."""
    d = np.asarray(xa, np.float64) - np.asarray(xb, np.float64)
    return np.sqrt((d * d).sum(axis=(-1, -2)) / d.shape[-2])


# --------------------------------------------------------------------------- grid construction

def grid_point_potential(points, coor_rec, rec_fftypes, lig_fftype, ff, plateaudis, cutoff=50.0):
    """Potential and force (ATTRACT sign) of one ligand type at points (M,3):
    plateau-capped nonbon summed over beads within cutoff (test-calc-grid-energy.py:132-148);
    capped beads contribute the plateau gradient projected on the plateau sphere."""
    d = np.asarray(points, np.float64)[:, None, :] - np.asarray(coor_rec, np.float64)[None]
    dsq = (d * d).sum(-1)
    prm = [np.asarray(a, np.float64)[np.asarray(rec_fftypes), lig_fftype][None] for a in ff]
    P2 = plateaudis ** 2
    dd = np.maximum(dsq, P2)
    e = nonbon(dd, *prm)
    gd = nonbon_dsq_grad(dd, *prm)
    with np.errstate(divide="ignore", invalid="ignore"):
        gs = np.where(dsq < P2, np.where(dsq > 0, 2 * gd * plateaudis / np.sqrt(dsq), 0.0), 2 * gd)
    inc = dsq < cutoff ** 2
    E = np.where(inc, e, 0).sum(1)
    F = -(np.where(inc, gs, 0)[:, :, None] * d).sum(1)
    return np.concatenate([E[:, None], F], axis=1)


def grid_point_neighbours(point, coor_rec, neighbourdis):
    """Receptor beads strictly within neighbourdis of a grid point, ascending index."""
    d = np.asarray(coor_rec, np.float64) - np.asarray(point, np.float64)
    return np.nonzero((d * d).sum(-1) < neighbourdis ** 2)[0]
