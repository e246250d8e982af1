"""Pins the block-layer oracle (oracle/block_oracle.py) against geometry and
the reference's own rot_to_quat_batch outputs."""
import os

import numpy as np

from oracle import block_oracle as BO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
VERTS = np.eye(3)  # three surfels per face, one on each vertex


def quat_to_matrix(q):
    """Standard (w, x, y, z) -> rotation matrix, written independently."""
    w, x, y, z = q
    return np.array([
        [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z],
    ])


def test_ellipsoid_closed_form():
    rng = np.random.default_rng(0)
    params, occ = BO.random_blocks(rng, 3)
    params[:, 10:12] = 0.0  # sigmoid(0) * 1.8 + 0.1 = 1: an ellipsoid
    eta, om = BO.grid(12, 16)
    means, _, _, _ = BO.generate(params, occ, eta, om, bary=VERTS)
    per = 12 * 16 * 3
    for b in range(3):
        S = np.exp(params[b, 0:3]) + 0.2
        q = params[b, 3:7] / np.linalg.norm(params[b, 3:7])
        R = quat_to_matrix(q)
        # vertices are ((v * 0.25 * S) @ R_ref) + t with R_ref = R(q) in row form
        local = (means[b * per:(b + 1) * per] - params[b, 7:10]) @ R.T
        unit = local / (0.25 * S)
        np.testing.assert_allclose(np.linalg.norm(unit, axis=1), 1.0, atol=1e-12)


def test_frames_orthonormal_and_aligned():
    rng = np.random.default_rng(1)
    params, occ = BO.random_blocks(rng, 2)
    eta, om = BO.grid(10, 12)
    verts, _, _, _ = BO.generate(params, occ, eta, om, bary=VERTS)
    _, scales, rots, opac = BO.generate(params, occ, eta, om)
    V = verts.reshape(-1, 3, 3)  # per face: V0, V1, V2
    for f in range(V.shape[0]):
        M = quat_to_matrix(rots[f])
        np.testing.assert_allclose(M.T @ M, np.eye(3), atol=1e-12)
        n = np.cross(V[f, 1] - V[f, 0], V[f, 2] - V[f, 0])
        c = V[f].mean(0)
        d1 = V[f, 1] - c
        # the reference divides by (|v| + 1e-8), which skews the frame of tiny
        # triangles; allow that bias
        tol = 1e-9 + 1e-7 / np.linalg.norm(n) + 1e-7 / np.linalg.norm(d1)
        np.testing.assert_allclose(M[:, 2], n / np.linalg.norm(n), atol=tol)
        np.testing.assert_allclose(M[:, 0], d1 / np.linalg.norm(d1), atol=tol)
        np.testing.assert_allclose(scales[f, 0], np.linalg.norm(d1) / 2, rtol=1e-7, atol=2e-8)
        np.testing.assert_allclose(scales[f, 1], np.dot(V[f, 2] - c, M[:, 1]) / 2, rtol=1e-5, atol=1e-8 + tol * np.linalg.norm(V[f, 2] - c))
        assert rots[f, 0] >= 0
    np.testing.assert_allclose(opac[:, 0], np.repeat(1 / (1 + np.exp(-occ)), 10 * 12))


def test_rot_to_quat_matches_reference_golden():
    g = np.load(os.path.join(GOLDEN, "rot_to_quat.npz"))
    q = BO.rot_to_quat_batch(g["R"].astype(np.float64))
    np.testing.assert_allclose(q, g["quat"], atol=2e-6)


def test_rot_to_quat_round_trip():
    rng = np.random.default_rng(2)
    q = rng.normal(size=(500, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    R = np.stack([quat_to_matrix(x) for x in q])
    np.testing.assert_allclose(BO.rot_to_quat_batch(R), q, atol=1e-12)


def test_superquadric_vertices_at_unequal_exponents():
    """eps1 != eps2 (utils/superquadric.py:10-14): every vertex, mapped back
    to the block frame and divided by the semi-axes 0.25 S, satisfies the
    superquadric implicit equation (|x|^(2/e2) + |z|^(2/e2))^(e2/e1) +
    |y|^(2/e1) = 1 (superquadric.py:25-46), and inverting the
    parametrisation gives back the grid angles: sign(y) = sign(sin eta),
    |y|^(1/e1) = |sin eta|, sign(x) = sign(sin omega), sign(z) =
    sign(cos omega), (|x| / |z|)^(1/e2) = |tan omega|.  Exchanging e1 and
    e2, or sin and cos of either angle, fails one of these."""
    rng = np.random.default_rng(5)
    params, occ = BO.random_blocks(rng, 4, eps_range=(0.2, 1.8))
    e = np.array([[0.3, 1.6], [1.7, 0.4], [0.6, 1.1], [1.25, 0.25]])
    params[:, 10:12] = np.log((e - 0.1) / 1.8 / (1 - (e - 0.1) / 1.8))
    params[:, 0:3] = np.log(rng.uniform(0.1, 0.5, size=(4, 3)))  # unequal semi-axes
    ne, no = 15, 20  # 16 latitudes, none at sin(eta) = 0
    eta, om = BO.grid(ne, no)
    means, _, _, _ = BO.generate(params, occ, eta, om, bary=VERTS)
    per = ne * no * 3
    ii, jj = np.meshgrid(np.arange(ne), np.arange(no), indexing="ij")
    # the surfel on vertex k of face (i, j) sits on grid node (i + [0,1,0][k], (j + [0,0,1][k]) % no)
    node_i = np.stack([ii, ii + 1, ii], -1).reshape(-1)
    node_j = np.stack([jj, jj, (jj + 1) % no], -1).reshape(-1)
    for b in range(4):
        e1, e2 = e[b]
        S = np.exp(params[b, 0:3]) + 0.2
        R = quat_to_matrix(params[b, 3:7] / np.linalg.norm(params[b, 3:7]))
        u = ((means[b * per:(b + 1) * per] - params[b, 7:10]) @ R.T) / (0.25 * S)
        x, y, z = np.abs(u[:, 0]), np.abs(u[:, 1]), np.abs(u[:, 2])
        F = (x ** (2 / e2) + z ** (2 / e2)) ** (e2 / e1) + y ** (2 / e1)
        np.testing.assert_allclose(F, 1.0, atol=1e-10)
        et, omg = eta[node_i], om[node_j]
        np.testing.assert_array_equal(np.sign(u[:, 1]), np.sign(np.sin(et)))
        np.testing.assert_allclose(y ** (1 / e1), np.abs(np.sin(et)), atol=1e-10)
        np.testing.assert_array_equal(np.sign(u[:, 0]), np.sign(np.sin(omg)))
        np.testing.assert_array_equal(np.sign(u[:, 2]), np.sign(np.cos(omg)))
        np.testing.assert_allclose((x / z) ** (1 / e2), np.abs(np.tan(omg)), rtol=1e-9)
