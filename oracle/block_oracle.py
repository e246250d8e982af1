"""numpy restatement of the superquadric block layer (CPU, FP64 by default).

TEST INFRASTRUCTURE ONLY (imported by tests/).  Restates
BlockGaussianModel.prepare_scaling_rot and the activations the rasteriser
sees, games/block_mesh_splatting/scene/block_gaussian_model.py:103-121,
178-256, with its helpers parametric_sq (utils/superquadric.py:10-14),
signed_pow (utils/pytorch.py:28-29), quaternion_to_rotation_matrix
(superquadric.py:93-101) and rot_to_quat_batch (utils/general_utils.py:10-87).
The mesh is a regular (eta, omega) grid: face (i, j) is the triangle
(V[i][j], V[i+1][j], V[i][(j+1) % n_omega]).

Pinned by tests/test_oracle_block.py: ellipsoid closed form (eps = 1),
orthonormal frames whose third axis is the face normal, quaternion ->
matrix round trip, and the reference's own rot_to_quat_batch outputs
(tests/golden/rot_to_quat.npz).
"""
from __future__ import annotations

import numpy as np


def signed_pow(t, e):
    return np.sign(t) * np.abs(t) ** e


def parametric_sq(eta, omega, eps1, eps2):
    ce, se = signed_pow(np.cos(eta), eps1), signed_pow(np.sin(eta), eps1)
    co, so = signed_pow(np.cos(omega), eps2), signed_pow(np.sin(omega), eps2)
    return np.stack([ce * so, se, ce * co], axis=-1)


def quaternion_to_rotation_matrix(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        1 - 2 * y ** 2 - 2 * z ** 2, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w,
        2 * x * y + 2 * z * w, 1 - 2 * x ** 2 - 2 * z ** 2, 2 * y * z - 2 * x * w,
        2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x ** 2 - 2 * y ** 2,
    ], axis=-1).reshape(-1, 3, 3)


def normalize(v, eps=1e-12):
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / np.maximum(n, eps)


def rot_to_quat_batch(R):
    """pytorch3d-style best-conditioned branch, then w >= 0 (general_utils.py:34-87)."""
    m = R.reshape(-1, 9)
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = m.T
    pre = np.stack([1.0 + m00 + m11 + m22, 1.0 + m00 - m11 - m22, 1.0 - m00 + m11 - m22,
                    1.0 - m00 - m11 + m22], axis=-1)
    q_abs = np.where(pre > 0, np.sqrt(np.maximum(pre, 0)), 0.0)
    cand = np.stack([
        np.stack([q_abs[:, 0] ** 2, m21 - m12, m02 - m20, m10 - m01], -1),
        np.stack([m21 - m12, q_abs[:, 1] ** 2, m10 + m01, m02 + m20], -1),
        np.stack([m02 - m20, m10 + m01, q_abs[:, 2] ** 2, m12 + m21], -1),
        np.stack([m10 - m01, m20 + m02, m21 + m12, q_abs[:, 3] ** 2], -1),
    ], axis=-2)
    cand = cand / (2.0 * np.maximum(q_abs, 0.1))[..., None]
    k = np.argmax(q_abs, axis=-1)
    out = cand[np.arange(len(k)), k]
    return np.where(out[:, :1] < 0, -out, out)


def grid(n_eta: int = 64, n_omega: int = 64):
    """Vertex latitudes (n_eta + 1, cell-centred, poles excluded) and
    longitudes (n_omega, half-cell offset so no sin/cos is exactly 0)."""
    eta = -np.pi / 2 + (np.arange(n_eta + 1) + 0.5) * np.pi / (n_eta + 1)
    omega = -np.pi + (np.arange(n_omega) + 0.5) * 2 * np.pi / n_omega
    return eta, omega


def generate(params, occupancy, eta, omega, ratio=0.25, scale_min=0.2, surfel_scale=1.0, bary=None,
             dtype=np.float64):
    """Returns means3D, scales, rotations, opacities (the rasteriser inputs)."""
    p = np.asarray(params, dtype=dtype).reshape(-1, 12)
    occ = np.asarray(occupancy, dtype=dtype).reshape(-1)
    eta = np.asarray(eta, dtype=dtype)
    omega = np.asarray(omega, dtype=dtype)
    B = p.shape[0]
    ne, no = len(eta) - 1, len(omega)
    bary = np.full((1, 3), 1.0 / 3.0, dtype=dtype) if bary is None else np.asarray(bary, dtype=dtype)
    K = bary.shape[0]
    S = np.exp(p[:, 0:3]) + scale_min                             # get_mesh_scaling
    R = quaternion_to_rotation_matrix(normalize(p[:, 3:7]))       # get_mesh_rot
    t = p[:, 7:10]
    eps = 1.0 / (1.0 + np.exp(-p[:, 10:12])) * 1.8 + 0.1          # get_verts
    E, Om = np.meshgrid(eta, omega, indexing="ij")                # (ne+1, no)
    verts = parametric_sq(E[None], Om[None], eps[:, 0, None, None], eps[:, 1, None, None]) * ratio
    V = np.einsum("bijc,bcd->bijd", verts * S[:, None, None, :], R) + t[:, None, None, :]
    ii, jj = np.meshgrid(np.arange(ne), np.arange(no), indexing="ij")
    V0 = V[:, ii, jj]
    V1 = V[:, ii + 1, jj]
    V2 = V[:, ii, (jj + 1) % no]
    tri = np.stack([V0, V1, V2], axis=-2).reshape(B, ne * no, 3, 3)
    xyz = np.einsum("kv,bfvc->bfkc", bary, tri)
    e8 = 1e-8
    n = np.cross(tri[:, :, 1] - tri[:, :, 0], tri[:, :, 2] - tri[:, :, 0])
    v0 = n / (np.linalg.norm(n, axis=-1, keepdims=True) + e8)
    c = tri.mean(axis=2)
    v1 = tri[:, :, 1] - c
    v1n = np.linalg.norm(v1, axis=-1, keepdims=True) + e8
    v1 = v1 / v1n
    v2i = tri[:, :, 2] - c
    v2 = v2i - (v2i * v0).sum(-1, keepdims=True) * v0 - (v2i * v1).sum(-1, keepdims=True) * v1
    v2 = v2 / (np.linalg.norm(v2, axis=-1, keepdims=True) + e8)
    s = np.concatenate([v1n / 2.0, (v2i * v2).sum(-1, keepdims=True) / 2.0], axis=-1)
    scaling = np.log(np.maximum(surfel_scale * s, 0.0) + e8)      # _scaling
    M = np.stack([v1, v2, v0], axis=-1)                           # columns v1, v2, v0
    q = normalize(rot_to_quat_batch(M.reshape(-1, 3, 3)))         # get_rotation
    P = B * ne * no * K
    means = xyz.reshape(P, 3)
    scales = np.repeat(np.exp(scaling).reshape(B * ne * no, 2), K, axis=0)
    rots = np.repeat(q, K, axis=0)
    opac = np.repeat(1.0 / (1.0 + np.exp(-occ)), ne * no * K)[:, None]
    return means, scales, rots, opac


def random_blocks(rng, B, scale_range=(0.1, 0.5), eps_range=(0.2, 1.8)):
    """Block parameters: exp(sq_s) ~ U[scale_range] per axis, eps in
    eps_range, a random unit quaternion and a translation in the unit ball."""
    sq_s = np.log(rng.uniform(*scale_range, size=(B, 3)))
    q = rng.normal(size=(B, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    t = rng.normal(size=(B, 3))
    t = t / np.linalg.norm(t, axis=1, keepdims=True) * rng.uniform(0, 1, size=(B, 1)) ** (1 / 3)
    e = rng.uniform(*eps_range, size=(B, 2))
    sq_eps = np.log((e - 0.1) / 1.8 / (1 - (e - 0.1) / 1.8))
    occ = rng.normal(size=B)
    return np.concatenate([sq_s, q, t, sq_eps], axis=1), occ


# ---------------------------------------------------------------------------
# Superquadric signed distance and the block-stage losses
# (utils/superquadric.py:25-46, utils/pytorch.py:32-33,
# block_gaussian_model.py:522-582), FP64.  Pinned by tests/test_block_losses.py:
# eps1 = eps2 = 1 gives |p_block| - 1 (a sphere); the block layer's own
# surface vertices lie on the zero set at eps1 != eps2; a point on the
# block's y axis at height y gives (2^(e2/e1) 1e-6^(1/e1) + y^(2/e1))^(e1/2) - 1
# (the clamped x and z terms); the clamps' flat regions; the loss reductions
# on hand-made fields with known answers.
# ---------------------------------------------------------------------------
SQRT_EPS = 1e-6


def block_frame(params, scale_min=0.2):
    """S, R (row-major), T, eps1, eps2 of B x 12 block parameters."""
    p = np.asarray(params, dtype=np.float64).reshape(-1, 12)
    S = np.exp(p[:, 0:3]) + scale_min
    R = quaternion_to_rotation_matrix(normalize(p[:, 3:7]))
    eps = 1.0 / (1.0 + np.exp(-p[:, 10:12])) * 1.8 + 0.1
    return S, R, p[:, 7:10], eps[:, 0], eps[:, 1]


def implicit_sq_sdf(pinv, eps1, eps2):
    """implicit_sq(points, eps1, eps2, safe=True, as_sdf=2) on (..., 3)
    points with eps broadcast against the leading dims."""
    q = np.clip(pinv, -5.0, 5.0)
    e1, e2 = np.asarray(eps1, np.float64), np.asarray(eps2, np.float64)
    sp = lambda t, e: np.maximum(t, SQRT_EPS) ** e  # noqa: E731 (safe_pow)
    x = sp(q[..., 0] ** 2, 1.0 / e2)
    y = sp(q[..., 1] ** 2, 1.0 / e1)
    z = sp(q[..., 2] ** 2, 1.0 / e2)
    res = sp(x + z, e2 / e1) + y
    return sp(res, e1 / 2.0) - 1.0


def superquadric_sdf(params, points, ratio=0.25, scale_min=0.2):
    """B x N: implicit_sq of ((p - T) @ R^T) / (S ratio) for every block."""
    S, R, T, e1, e2 = block_frame(params, scale_min)
    pts = np.asarray(points, dtype=np.float64).reshape(1, -1, 3)
    pinv = np.einsum("bnj,bkj->bnk", pts - T[:, None, :], R) / (S[:, None, :] * ratio)
    return implicit_sq_sdf(pinv, e1[:, None], e2[:, None])


def overlap_value(sdf, alpha):
    """(sum_b sigmoid(-sdf / 0.005) alpha_b - 1.95).clamp(0).mean() (block_gaussian_model.py:534-543)."""
    occ = 1.0 / (1.0 + np.exp(sdf / 0.005)) * np.asarray(alpha, np.float64).reshape(-1, 1)
    return np.maximum(occ.sum(0) - 1.95, 0.0).mean()


def ray_terms(sdf, label_mask, occ_logits):
    """(ray_in_mask, ray_out_mask, primitive_opacity) of a B x rays x samples
    field (block_gaussian_model.py:553-582)."""
    m = np.asarray(label_mask, np.float64)
    rays = np.transpose(sdf, (1, 0, 2)).reshape(sdf.shape[1], -1)
    ray_in = np.maximum(rays.min(1)[m > 0.5] - 1e-3, 0.0).mean()
    out = np.maximum(np.max(-rays, axis=1) * (m < 0.5) - 1e-3, 0.0).sum() / (m < 0.5).sum()
    a = 1.0 / (1.0 + np.exp(-np.asarray(occ_logits, np.float64).reshape(-1)))
    alpha = np.where(sdf <= 1e-8, a[:, None, None], 0.0)
    alpha = np.transpose(alpha, (1, 0, 2)).reshape(sdf.shape[1], -1).max(1).clip(1e-3, 1 - 1e-3)
    bce = -(m * np.log(alpha) + (1 - m) * np.log(1 - alpha)).mean()
    return ray_in, out, bce
