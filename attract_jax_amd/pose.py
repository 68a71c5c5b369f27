"""Pose conventions of the reference, on the host (numpy).

A pose is a 4x4 matrix M applied to row vectors, x' = [x, y, z, 1] . M
(test-grid-energrad.py:299-301); test4.py:396-401 builds M from a quaternion with
M[0:3, 0:3] = Rotation.as_matrix() and M[3, 0:3] = translation.  Quaternions are
scalar-last (x, y, z, w) as in scipy / jax.scipy.spatial.transform.Rotation.
"""
import numpy as np


def as_matrix(quat):
    """Rotation.as_matrix of (possibly unnormalised) quaternions (..., 4) -> (..., 3, 3)."""
    q = np.asarray(quat)
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    R = np.empty(q.shape[:-1] + (3, 3), q.dtype)
    R[..., 0, 0] = x2 - y2 - z2 + w2
    R[..., 0, 1] = 2 * (xy - zw)
    R[..., 0, 2] = 2 * (xz + yw)
    R[..., 1, 0] = 2 * (xy + zw)
    R[..., 1, 1] = -x2 + y2 - z2 + w2
    R[..., 1, 2] = 2 * (yz - xw)
    R[..., 2, 0] = 2 * (xz - yw)
    R[..., 2, 1] = 2 * (yz + xw)
    R[..., 2, 2] = -x2 - y2 + z2 + w2
    return R


def from_matrix(R):
    """Rotation.from_matrix (largest-of-diagonal/trace branch, then normalised) -> (..., 4)."""
    R = np.asarray(R)
    m = R.reshape(-1, 3, 3)
    tr = m[:, 0, 0] + m[:, 1, 1] + m[:, 2, 2]
    decision = np.stack([m[:, 0, 0], m[:, 1, 1], m[:, 2, 2], tr], axis=1)
    choice = np.argmax(decision, axis=1)
    q = np.empty((len(m), 4), R.dtype)
    for n in range(len(m)):
        c = choice[n]
        if c != 3:
            i, j, k = c, (c + 1) % 3, (c + 2) % 3
            q[n, i] = 1 - decision[n, 3] + 2 * m[n, i, i]
            q[n, j] = m[n, j, i] + m[n, i, j]
            q[n, k] = m[n, k, i] + m[n, i, k]
            q[n, 3] = m[n, k, j] - m[n, j, k]
        else:
            q[n, 0] = m[n, 2, 1] - m[n, 1, 2]
            q[n, 1] = m[n, 0, 2] - m[n, 2, 0]
            q[n, 2] = m[n, 1, 0] - m[n, 0, 1]
            q[n, 3] = 1 + decision[n, 3]
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.reshape(R.shape[:-2] + (4,))


def mat_from_qt(quat, trans, fill=0.0):
    """4x4 pose matrices from quaternion + translation (test4.py:398-400 uses fill=1)."""
    quat = np.asarray(quat)
    trans = np.asarray(trans)
    M = np.full(quat.shape[:-1] + (4, 4), fill, quat.dtype)
    M[..., :3, :3] = as_matrix(quat)
    M[..., 3, :3] = trans
    M[..., 3, 3] = 1
    return M


def qt_from_mat(mat):
    mat = np.asarray(mat)
    return from_matrix(mat[..., :3, :3]), np.array(mat[..., 3, :3])


def random_quaternions(n, rng):
    """Uniform random rotations (Shoemake 1992), scalar-last."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                     b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)


def random_poses(n, rng, rmin, rmax):
    """n poses with uniform rotation and centre at distance U(rmin, rmax) in a uniform direction.
    Returns (quat (n,4), trans (n,3)) float64."""
    q = random_quaternions(n, rng)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = d * rng.uniform(rmin, rmax, size=(n, 1))
    return q, t


def dof_gradient(mat, dEdM):
    """6-DOF gradient (dE/dt, dE/domega) from dE/dM.

    omega is a rotation vector applied in the world frame about the ligand centre t:
    x'' = t + exp([omega]x)(x' - t), i.e. R -> R . exp([omega]x)^T; then
    dE/domega_a = eps_abc (R^T A)_bc with A = dE/dM[0:3, 0:3]."""
    mat = np.asarray(mat, np.float64)
    G = np.asarray(dEdM, np.float64)
    R = mat[..., :3, :3]
    A = G[..., :3, :3]
    B = np.einsum("...kb,...kc->...bc", R, A)
    tau = np.stack([B[..., 1, 2] - B[..., 2, 1], B[..., 2, 0] - B[..., 0, 2], B[..., 0, 1] - B[..., 1, 0]], axis=-1)
    return np.concatenate([G[..., 3, :3], tau], axis=-1)


def rotvec_matrix(omega):
    """exp([omega]x) for (..., 3) rotation vectors (Rodrigues)."""
    omega = np.asarray(omega, np.float64)
    th = np.linalg.norm(omega, axis=-1, keepdims=True)[..., None]
    K = np.zeros(omega.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2], K[..., 1, 2] = -omega[..., 2], omega[..., 1], -omega[..., 0]
    K[..., 1, 0], K[..., 2, 0], K[..., 2, 1] = omega[..., 2], -omega[..., 1], omega[..., 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(th > 0, np.sin(th) / th, 1.0)
        c = np.where(th > 0, (1 - np.cos(th)) / th ** 2, 0.5)
    return np.eye(3) + s * K + c * (K @ K)


def perturb_pose(mat, dt, omega):
    """Apply the 6-DOF perturbation of dof_gradient to pose matrices."""
    mat = np.array(mat, np.float64)
    Q = rotvec_matrix(omega)
    mat[..., :3, :3] = mat[..., :3, :3] @ np.swapaxes(Q, -1, -2)
    mat[..., 3, :3] += dt
    return mat
