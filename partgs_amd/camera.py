"""Camera conventions of the reference renderer.

* ``world_view_transform`` = (W2C)^T, built in FP64 numpy and cast to FP32
  (utils/graphics_utils.py:29-40, getWorld2View2; scene/cameras.py:58);
* ``projection_matrix`` = P^T with P from getProjectionMatrix
  (utils/graphics_utils.py:42-62): +z forward, P[3,2] = +1, depth in [0, 1];
* ``full_proj_transform`` = world_view_transform @ projection_matrix
  (scene/cameras.py:60), i.e. row-vector convention;
* ``camera_center`` = inverse(world_view_transform)[3, :3] (cameras.py:61).

The kernels read both matrices as flat 16-float arrays, column-major in the
column-vector sense (dsr/cuda_rasterizer/auxiliary.h:79-98).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


def get_world2view2(R: np.ndarray, t: np.ndarray, translate=np.array([0.0, 0.0, 0.0]), scale: float = 1.0):
    """W2C (4x4, float32) from the camera-to-world rotation R and translation t."""
    Rt = np.zeros((4, 4))
    Rt[:3, :3] = R.transpose()
    Rt[:3, 3] = t
    Rt[3, 3] = 1.0
    C2W = np.linalg.inv(Rt)
    cam_center = C2W[:3, 3]
    cam_center = (cam_center + translate) * scale
    C2W[:3, 3] = cam_center
    Rt = np.linalg.inv(C2W)
    return np.float32(Rt)


def get_projection_matrix(znear: float, zfar: float, fovX: float, fovY: float) -> torch.Tensor:
    tanHalfFovY = math.tan(fovY / 2)
    tanHalfFovX = math.tan(fovX / 2)
    top = tanHalfFovY * znear
    bottom = -top
    right = tanHalfFovX * znear
    left = -right
    P = torch.zeros(4, 4)
    z_sign = 1.0
    P[0, 0] = 2.0 * znear / (right - left)
    P[1, 1] = 2.0 * znear / (top - bottom)
    P[0, 2] = (right + left) / (right - left)
    P[1, 2] = (top + bottom) / (top - bottom)
    P[3, 2] = z_sign
    P[2, 2] = z_sign * zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    return P


@dataclass
class Camera:
    """The fields of the reference's Camera/MiniCam that the renderer reads."""

    image_width: int
    image_height: int
    FoVx: float
    FoVy: float
    znear: float
    zfar: float
    world_view_transform: torch.Tensor
    projection_matrix: torch.Tensor
    full_proj_transform: torch.Tensor
    camera_center: torch.Tensor

    def to(self, device) -> "Camera":
        return Camera(self.image_width, self.image_height, self.FoVx, self.FoVy, self.znear, self.zfar,
                      self.world_view_transform.to(device), self.projection_matrix.to(device),
                      self.full_proj_transform.to(device), self.camera_center.to(device))


def camera_from_Rt(R: np.ndarray, t: np.ndarray, fovx: float, fovy: float, width: int, height: int,
                   znear: float = 0.01, zfar: float = 100.0) -> Camera:
    """Camera from a camera-to-world rotation R and W2C translation t (CPU tensors)."""
    wvt = torch.tensor(get_world2view2(R, t)).transpose(0, 1).contiguous()
    proj = get_projection_matrix(znear, zfar, fovx, fovy).transpose(0, 1).contiguous()
    full = (wvt.unsqueeze(0).bmm(proj.unsqueeze(0))).squeeze(0).contiguous()
    center = wvt.inverse()[3, :3].contiguous()
    return Camera(width, height, fovx, fovy, znear, zfar, wvt, proj, full, center)


def look_at(position, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovx: float = math.radians(60.0),
            fovy: float = math.radians(60.0), width: int = 256, height: int = 256, znear: float = 0.01,
            zfar: float = 100.0) -> Camera:
    """Camera at ``position`` looking at ``target`` (x right, y down, z forward)."""
    pos = np.asarray(position, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - pos
    f /= np.linalg.norm(f)
    upv = np.asarray(up, dtype=np.float64)
    if abs(float(np.dot(f, upv))) > 0.999 * np.linalg.norm(upv):
        upv = np.array([0.0, 0.0, 1.0])
    r = np.cross(f, upv)
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    w2c_rot = np.stack([r, d, f], axis=0)  # rows: camera axes in world coordinates
    t = -w2c_rot @ pos
    return camera_from_Rt(w2c_rot.T, t, fovx, fovy, width, height, znear, zfar)


def random_sphere_camera(seed: int, radius: float = 3.0, width: int = 256, height: int = 256,
                         fov_deg: float = 60.0) -> Camera:
    """Camera at radius * u, u uniform on the unit sphere (seeded), looking at the origin."""
    g = np.random.default_rng(seed)
    u = g.normal(size=3)
    u /= np.linalg.norm(u)
    fov = math.radians(fov_deg)
    return look_at(radius * u, fovx=fov, fovy=fov, width=width, height=height)
