"""Synthetic surfel scenes for tests and benchmarks (CPU tensors, seeded).

* ``surfel_cloud``  -- 16,384 surfels in [-1, 1]^3, log-uniform scales in
  [1e-3, 5e-2], random unit quaternions, opacity sigmoid(N(0, 1)), SH degree 0
  colours U[0, 1]; one camera at radius 3, fov 60 deg, white background.
* ``superquadric_surfels`` -- surfels sampled on the surfaces of random
  superquadrics (eps1, eps2 ~ U[0.1, 1.9], axes U[0.1, 0.5], random pose in
  the unit ball), oriented by the analytic surface normal, log-uniform
  scales in [1e-3, 1e-2], SH degree 3 with DC colour U[0, 1] and higher bands
  N(0, 0.05^2).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .camera import Camera, random_sphere_camera

SH_C0 = 0.28209479177387814


def rgb2sh(rgb):
    return (rgb - 0.5) / SH_C0


def quat_from_matrix(R: np.ndarray) -> np.ndarray:
    """(w, x, y, z) unit quaternions (w >= 0) of rotation matrices (N, 3, 3)."""
    R = np.asarray(R, dtype=np.float64)
    tr = R[:, 0, 0] + R[:, 1, 1] + R[:, 2, 2]
    cand = np.stack([
        1.0 + tr,
        1.0 + R[:, 0, 0] - R[:, 1, 1] - R[:, 2, 2],
        1.0 - R[:, 0, 0] + R[:, 1, 1] - R[:, 2, 2],
        1.0 - R[:, 0, 0] - R[:, 1, 1] + R[:, 2, 2],
    ], axis=1)
    k = np.argmax(cand, axis=1)
    q = np.zeros((R.shape[0], 4))
    s = np.sqrt(np.maximum(cand[np.arange(R.shape[0]), k], 1e-300)) * 2.0
    m = k == 0
    q[m, 0] = 0.25 * s[m]
    q[m, 1] = (R[m, 2, 1] - R[m, 1, 2]) / s[m]
    q[m, 2] = (R[m, 0, 2] - R[m, 2, 0]) / s[m]
    q[m, 3] = (R[m, 1, 0] - R[m, 0, 1]) / s[m]
    m = k == 1
    q[m, 0] = (R[m, 2, 1] - R[m, 1, 2]) / s[m]
    q[m, 1] = 0.25 * s[m]
    q[m, 2] = (R[m, 0, 1] + R[m, 1, 0]) / s[m]
    q[m, 3] = (R[m, 0, 2] + R[m, 2, 0]) / s[m]
    m = k == 2
    q[m, 0] = (R[m, 0, 2] - R[m, 2, 0]) / s[m]
    q[m, 1] = (R[m, 0, 1] + R[m, 1, 0]) / s[m]
    q[m, 2] = 0.25 * s[m]
    q[m, 3] = (R[m, 1, 2] + R[m, 2, 1]) / s[m]
    m = k == 3
    q[m, 0] = (R[m, 1, 0] - R[m, 0, 1]) / s[m]
    q[m, 1] = (R[m, 0, 2] + R[m, 2, 0]) / s[m]
    q[m, 2] = (R[m, 1, 2] + R[m, 2, 1]) / s[m]
    q[m, 3] = 0.25 * s[m]
    q *= np.where(q[:, :1] < 0, -1.0, 1.0)
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


@dataclass
class SurfelScene:
    means3D: torch.Tensor
    scales: torch.Tensor
    rotations: torch.Tensor
    opacities: torch.Tensor
    shs: torch.Tensor
    sh_degree: int
    camera: Camera
    bg: torch.Tensor

    @property
    def P(self) -> int:
        return int(self.means3D.shape[0])

    def to(self, device) -> "SurfelScene":
        return SurfelScene(self.means3D.to(device), self.scales.to(device), self.rotations.to(device),
                           self.opacities.to(device), self.shs.to(device), self.sh_degree,
                           self.camera.to(device), self.bg.to(device))


def surfel_cloud(seed: int = 0, P: int = 16384, width: int = 256, height: int = 256) -> SurfelScene:
    g = torch.Generator().manual_seed(seed)
    means = torch.rand(P, 3, generator=g) * 2 - 1
    scales = torch.exp(torch.rand(P, 2, generator=g) * (math.log(5e-2) - math.log(1e-3)) + math.log(1e-3))
    q = torch.randn(P, 4, generator=g)
    q = q / q.norm(dim=1, keepdim=True)
    opac = torch.sigmoid(torch.randn(P, 1, generator=g))
    rgb = torch.rand(P, 3, generator=g)
    shs = rgb2sh(rgb).reshape(P, 1, 3)
    cam = random_sphere_camera(seed, 3.0, width, height)
    bg = torch.ones(3)
    return SurfelScene(means.float(), scales.float(), q.float(), opac.float(), shs.float().contiguous(), 0, cam,
                       bg)


def _spow(t, e):
    return np.sign(t) * np.abs(t) ** e


def superquadric_surfels(seed: int = 0, P: int = 300_000, n_sq: int = 8, width: int = 800, height: int = 600,
                         sh_degree: int = 3) -> SurfelScene:
    rng = np.random.default_rng(seed)
    per = [P // n_sq + (1 if i < P % n_sq else 0) for i in range(n_sq)]
    pts, nrm = [], []
    for k in range(n_sq):
        e1, e2 = rng.uniform(0.1, 1.9, size=2)
        ax = rng.uniform(0.1, 0.5, size=3)
        Rk = random_rotations(rng, 1)[0]
        c = rng.normal(size=3)
        c = c / np.linalg.norm(c) * rng.uniform() ** (1 / 3)
        n = per[k]
        eta = rng.uniform(-math.pi / 2, math.pi / 2, size=n)
        om = rng.uniform(-math.pi, math.pi, size=n)
        ce, se, co, so = np.cos(eta), np.sin(eta), np.cos(om), np.sin(om)
        # surface point in the axis arrangement of superquadric.py:10-14
        x = ax[0] * _spow(ce, e1) * _spow(so, e2)
        y = ax[1] * _spow(se, e1)
        z = ax[2] * _spow(ce, e1) * _spow(co, e2)
        # analytic normal: gradient direction of the implicit surface
        nx = _spow(ce, 2 - e1) * _spow(so, 2 - e2) / ax[0]
        ny = _spow(se, 2 - e1) / ax[1]
        nz = _spow(ce, 2 - e1) * _spow(co, 2 - e2) / ax[2]
        p = np.stack([x, y, z], 1) @ Rk.T + c
        nn_ = np.stack([nx, ny, nz], 1) @ Rk.T
        nl = np.linalg.norm(nn_, axis=1, keepdims=True)
        nn_ = np.where(nl > 1e-12, nn_ / np.maximum(nl, 1e-12), np.array([0.0, 0.0, 1.0]))
        pts.append(p)
        nrm.append(nn_)
    pts = np.concatenate(pts)
    nrm = np.concatenate(nrm)
    helper = np.where(np.abs(nrm[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t1 = np.cross(helper, nrm)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    Rm = np.stack([t1, t2, nrm], axis=2)  # columns: tangent, bitangent, normal
    quats = quat_from_matrix(Rm)
    scales = np.exp(rng.uniform(math.log(1e-3), math.log(1e-2), size=(P, 2)))
    opac = 1.0 / (1.0 + np.exp(-rng.normal(size=(P, 1))))
    M = (sh_degree + 1) ** 2
    shs = rng.normal(0.0, 0.05, size=(P, M, 3))
    shs[:, 0, :] = rgb2sh(rng.uniform(0.0, 1.0, size=(P, 3)))
    cam = random_sphere_camera(seed, 3.0, width, height)
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a)).float()  # noqa: E731
    return SurfelScene(f32(pts), f32(scales), f32(quats), f32(opac), f32(shs), sh_degree, cam, torch.ones(3))


def upstream_grads(seed: int, H: int, W: int):
    """N(0, 1) upstream gradients for color (3,H,W) and others (7,H,W)."""
    g = torch.Generator().manual_seed(seed)
    return torch.randn(3, H, W, generator=g), torch.randn(7, H, W, generator=g)


def batched_views(seed: int = 0, n: int = 32, choices=((400, 300), (640, 480), (800, 600))):
    """n cameras on the radius-3 sphere looking at the origin (camera k from
    seed 1000 + k) with resolutions drawn from ``choices`` (seed ``seed``):
    the multi-view batch of 1,000,000-surfel scenes."""
    rng = np.random.default_rng(seed)
    res = [tuple(choices[int(k)]) for k in rng.integers(0, len(choices), size=n)]
    return [random_sphere_camera(1000 + k, 3.0, w, h) for k, (w, h) in enumerate(res)]
