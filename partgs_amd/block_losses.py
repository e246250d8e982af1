"""Block-stage losses around the rasteriser, on the gfx950 HIP library.

BlockGaussianModel.summarize_losses and its terms
(games/block_mesh_splatting/scene/block_gaussian_model.py:510-648):

* ``superquadric_sdf``: implicit_sq(points_inv, eps1, eps2, as_sdf=2)
  (utils/superquadric.py:25-46) of world points in every block's frame, the
  B x N field every geometric term reduces; forward and adjoint are the
  fused kernels sb_sdf_fwd / sb_sdf_bwd (block parameters in the 12-float
  layout of the block layer: sq_s, sq_r, sq_t, sq_eps);
* ``overlap_loss`` (:522-545), ``parsimony_loss`` (:547-549),
  ``ray_coverage_loss`` (:575-582), ``ray_uncoverage_loss`` (:553-563),
  ``occupancy_opacity_loss`` (:565-573), ``primitive_scale_regularization``
  (:517-519) and ``summarize_losses`` (:584-648), reductions of that field in
  torch with the reference's operations;
* ``cast_rays`` (scene/cameras.py:79-96): stratified samples along random
  pixel rays.

The random draws (torch.rand / randn / randperm) take an optional
torch.Generator so a run can be repeated; without one they use torch's
global generator, as the reference does.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch
import torch.nn.functional as F

from . import _lib
from .losses import l1_loss, ssim

OVERLAP_N_POINTS = 2048     # block_gaussian_model.py:21
OVERLAP_N_BLOCKS = 1.95     # :22
OVERLAP_TEMPERATURE = 0.005  # :24
SQRT_EPS = 1e-6             # utils/pytorch.py:6


def _check(params, points):
    for name, t in (("params", params), ("points", points)):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float32:
            raise RuntimeError(f"{name} must be a float32 CUDA tensor")
    if params.dim() != 2 or params.shape[1] != 12:
        raise RuntimeError("params must be B x 12 (sq_s, sq_r, sq_t, sq_eps)")
    if points.dim() != 2 or points.shape[1] != 3:
        raise RuntimeError("points must be N x 3")


class _SuperquadricSdf(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, points, ratio, scale_min):
        lib = _lib.load()
        params = params.contiguous()
        points = points.detach().contiguous()
        _check(params, points)
        B, N = params.shape[0], points.shape[0]
        sdf = torch.empty((B, N), dtype=torch.float32, device=params.device)
        with torch.cuda.device(params.device):
            _lib.check(lib.sb_sdf_fwd(params.data_ptr(), B, float(ratio), float(scale_min), _lib.ptr(points), N,
                                      _lib.ptr(sdf), _lib.stream_handle(params.device)))
        ctx.save_for_backward(params, points)
        ctx.ratio, ctx.scale_min = float(ratio), float(scale_min)
        return sdf

    @staticmethod
    def backward(ctx, g):
        lib = _lib.load()
        params, points = ctx.saved_tensors
        B, N = params.shape[0], points.shape[0]
        g = g.contiguous()
        d = torch.empty_like(params)
        ws = torch.empty(int(lib.sb_sdf_bwd_workspace_bytes(B, N)), dtype=torch.uint8, device=params.device)
        with torch.cuda.device(params.device):
            _lib.check(lib.sb_sdf_bwd(params.data_ptr(), B, ctx.ratio, ctx.scale_min, _lib.ptr(points), N,
                                      _lib.ptr(g), d.data_ptr(), ws.data_ptr(), C.c_size_t(ws.numel()),
                                      _lib.stream_handle(params.device)))
        return d, None, None, None


def superquadric_sdf(params: torch.Tensor, points: torch.Tensor, ratio: float = 0.25,
                     scale_min: float = 0.2) -> torch.Tensor:
    """B x N signed distances of points (N x 3, world) to the B blocks,
    differentiable in params (B x 12)."""
    return _SuperquadricSdf.apply(params, points, ratio, scale_min)


def blocks_srt(params: torch.Tensor, scale_min: float = 0.2):
    """(S, R, T, eps1, eps2) of the blocks, as _blocks_SRT / _blocks_eps hold
    them (block_gaussian_model.py:112-121,190-192,285-288): S = exp(sq_s) +
    scale_min, R the row-major rotation of the normalised quaternion
    (superquadric.py:93-101), T = sq_t, eps = 1.8 sigmoid(sq_eps) + 0.1."""
    S = torch.exp(params[:, 0:3]) + scale_min
    q = F.normalize(params[:, 3:7], dim=-1)
    w, x, y, z = q.unbind(-1)
    R = torch.stack([1 - 2 * y ** 2 - 2 * z ** 2, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w,
                     2 * x * y + 2 * z * w, 1 - 2 * x ** 2 - 2 * z ** 2, 2 * y * z - 2 * x * w,
                     2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x ** 2 - 2 * y ** 2], -1).view(-1, 3, 3)
    eps = torch.sigmoid(params[:, 10:12]) * 1.8 + 0.1
    return S, R, params[:, 7:10], eps[:, 0:1], eps[:, 1:2]


def safe_pow(t, exponent, eps=SQRT_EPS):
    """utils/pytorch.py:32-33."""
    return t.clamp(eps).pow(exponent)


def overlap_loss(params, occ_logits, ratio: float = 0.25, scale_min: float = 0.2, n_points: int = OVERLAP_N_POINTS,
                 generator=None):
    """overlap_loss (block_gaussian_model.py:522-545): n_points uniform
    samples in each block's box (drawn without gradient), the occupancy
    sigmoid(-sdf / 0.005) of every sample in every block weighted by a noisy
    block opacity sigmoid(sq_occ + N(0, 1)), and the mean excess of the
    per-sample sum over 1.95 blocks."""
    B = params.shape[0]
    dev = params.device
    with torch.no_grad():
        S, R, T, _, _ = blocks_srt(params, scale_min)
        pts = torch.rand(B, n_points, 3, device=dev, generator=generator) * 2 - 1
        pts = (pts * ratio * S[:, None]) @ R + T[:, None]
        pts = pts.reshape(-1, 3)
    sdf = superquadric_sdf(params, pts, ratio, scale_min)
    occupancy = torch.sigmoid(-sdf / OVERLAP_TEMPERATURE)
    noise = torch.randn(occ_logits.shape, device=dev, generator=generator)
    alpha = torch.sigmoid(occ_logits.reshape(-1) + noise.reshape(-1))
    occupancy = occupancy * alpha[:, None]
    return {"overlap": (occupancy.sum(0) - OVERLAP_N_BLOCKS).clamp(0).mean()}


def parsimony_loss(occ_logits):
    """parsimony_loss (:547-549): mean of sqrt(sigmoid(sq_occ)) with the safe power."""
    return {"parsimony": safe_pow(torch.sigmoid(occ_logits), 0.5).mean()}


def primitive_scale_regularization(params, scale_min: float = 0.2, control_scale: float = 3.0):
    """primitive_scale_regularization (:517-519)."""
    S = torch.exp(params[:, 0:3]) + scale_min
    return {"scale_reg": (S - control_scale).clamp(min=0.0).sum()}


def ray_uncoverage_loss(sdf, label_mask):
    """ray_uncoverage_loss (:553-563); sdf is B x rays x samples."""
    sdf_rays = sdf.permute(1, 0, 2).flatten(start_dim=1, end_dim=2)
    max_values = torch.max(-sdf_rays, dim=1).values
    max_values = max_values * (label_mask < 0.5)
    ray_out_mask = torch.clamp(max_values - 1e-3, min=0)
    num = (label_mask < 0.5).sum()
    return {"ray_out_mask": ray_out_mask.sum() / num}


def occupancy_opacity_loss(sdf, label_mask, occ_logits):
    """occupancy_opacity_loss (:565-573)."""
    alpha = torch.sigmoid(occ_logits.reshape(-1))[:, None, None].expand(-1, sdf.shape[-2], sdf.shape[-1])
    alpha = torch.where(sdf <= 1e-8, alpha, 0)
    alpha = alpha.permute(1, 0, 2).flatten(start_dim=1, end_dim=2)
    alpha = torch.max(alpha, dim=1).values
    return {"primitive_opacity": F.binary_cross_entropy(alpha.clip(1e-3, 1.0 - 1e-3), label_mask).mean()}


def ray_coverage_loss(sdf, label_mask):
    """ray_coverage_loss (:575-582)."""
    sdf_rays = sdf.permute(1, 0, 2).flatten(start_dim=1, end_dim=2)
    min_values = torch.min(sdf_rays, dim=1).values
    return {"ray_in_mask": torch.clamp(min_values[label_mask > 0.5] - 1e-3, min=0).mean()}


def camera_intrinsics(camera) -> torch.Tensor:
    """K of scene/cameras.py:35-40."""
    import math

    K = torch.zeros((3, 3), dtype=torch.float32)
    K[0, 0] = (camera.image_width / 2.0) / math.tan(camera.FoVx / 2.0)
    K[1, 1] = (camera.image_height / 2.0) / math.tan(camera.FoVy / 2.0)
    K[0, 2] = camera.image_width / 2.0
    K[1, 2] = camera.image_height / 2.0
    K[2, 2] = 1.0
    return K


def _cam2world(X, world_view_transform):
    """Camera.cam2world with invert (scene/cameras.py:65-77)."""
    pose = world_view_transform.transpose(-1, -2)[:3, :]
    R, t = pose[..., :3], pose[..., 3:]
    R_inv = R.transpose(-1, -2)
    t_inv = (-R_inv @ t)[..., 0]
    pose_inv = torch.cat([R_inv, t_inv[..., None]], dim=-1)
    X_hom = torch.cat([X, torch.ones_like(X[..., :1])], dim=-1)
    return X_hom @ pose_inv.transpose(-1, -2)


def cast_rays(camera, depth_min: float = 0.0, depth_max: float = 10.0, num_samples: int = 100,
              num_rays: int = 2048, device="cuda", generator=None):
    """Camera.cast_rays (scene/cameras.py:79-96): num_rays random pixels
    (centres at +0.5), their rays from the camera centre and num_samples
    stratified depths in [depth_min, depth_max].  Returns (points
    (rays, samples, 3), ray_idx)."""
    H, W = int(camera.image_height), int(camera.image_width)
    ray_idx = torch.randperm(H * W, device=device, generator=generator)[:num_rays]
    xy = torch.stack([ray_idx % W, torch.div(ray_idx, W, rounding_mode="floor")], dim=-1).float().add_(0.5)
    xy_hom = torch.cat([xy, torch.ones_like(xy[..., :1])], dim=-1)
    K = camera_intrinsics(camera).to(device)
    wvt = camera.world_view_transform.to(device)
    points = xy_hom @ K.inverse().transpose(-1, -2)
    grid_3D = _cam2world(points, wvt)
    center_3D = _cam2world(torch.zeros_like(grid_3D), wvt)
    ray = grid_3D - center_3D
    rand = torch.rand(ray.shape[0], num_samples, 1, device=device, generator=generator)
    rand += torch.arange(num_samples, device=device)[None, :, None].float()
    depth = rand / num_samples * (depth_max - depth_min) + depth_min
    return center_3D.unsqueeze(1) + ray.unsqueeze(1) * depth, ray_idx


@dataclass
class BlockLossWeights:
    """The block stage's weights (arguments/__init__.py:111-122)."""

    lambda_dssim: float = 0.2
    rgb_weight: float = 1.0
    parsimony_weight: float = 0.002
    overlap_weight: float = 2.0
    occupancy_weight: float = 0.02
    ray_coverage_weigth: float = 10.0
    ray_uncoverage_weigth: float = 100.0
    primtive_scale_weight: float = 0.1
    data_type: str = "blender"


_NEAR_FAR = {"dtu": (0.5, 3.5), "blender": (2.0, 5.0), "blendermvs": (1.0, 3.0)}


def block_regularizers(params, occ_logits, camera, opt: BlockLossWeights, mask=None, ratio: float = 0.25,
                       scale_min: float = 0.2, use_reg: bool = True, generator=None, near_far=None,
                       device="cuda"):
    """The terms summarize_losses adds to the image loss
    (block_gaussian_model.py:598-646): primitive scale (dtu / blendermvs),
    parsimony, the three ray terms for one view's mask and overlap, each
    weighted.  Returns (weighted sum, loss_dict)."""
    loss_dict = {}
    loss = torch.zeros((), device=params.device)
    if opt.primtive_scale_weight and opt.data_type in ("dtu", "blendermvs"):
        d = primitive_scale_regularization(params, scale_min)
        loss_dict.update(d)
        loss = loss + d["scale_reg"] * opt.primtive_scale_weight
    if use_reg and opt.parsimony_weight > 0:
        d = parsimony_loss(occ_logits)
        loss_dict.update(d)
        loss = loss + d["parsimony"] * opt.parsimony_weight
    if use_reg and mask is not None:
        near, far = near_far if near_far is not None else _NEAR_FAR[opt.data_type]
        pts, ray_idx = cast_rays(camera, depth_min=near, depth_max=far, num_samples=128, device=params.device,
                                 generator=generator)
        label_mask = mask.flatten(start_dim=0, end_dim=1)[ray_idx]
        n_rays = label_mask.shape[0]
        sdf = superquadric_sdf(params, pts.reshape(-1, 3), ratio, scale_min)
        sdf = sdf.reshape(sdf.shape[0], n_rays, -1)
        for w, fn in ((opt.ray_coverage_weigth, lambda: ray_coverage_loss(sdf, label_mask)),
                      (opt.ray_uncoverage_weigth, lambda: ray_uncoverage_loss(sdf, label_mask)),
                      (opt.occupancy_weight, lambda: occupancy_opacity_loss(sdf, label_mask, occ_logits))):
            if w > 0:
                d = fn()
                loss_dict.update(d)
                for v in d.values():
                    loss = loss + v * w
    if opt.overlap_weight > 0:
        d = overlap_loss(params, occ_logits, ratio, scale_min, generator=generator)
        loss_dict.update(d)
        loss = loss + d["overlap"] * opt.overlap_weight
    return loss, loss_dict


def summarize_losses(image, gt_image, camera, opt: BlockLossWeights, params, occ_logits, mask=None,
                     ratio: float = 0.25, scale_min: float = 0.2, use_reg: bool = True, generator=None,
                     near_far=None):
    """summarize_losses (block_gaussian_model.py:584-648): the image term,
    then parsimony, the ray terms (with a mask) and overlap, each weighted as
    the reference does.  near_far overrides the data type's sampling range.
    Returns (Ll1, loss, loss_dict)."""
    Ll1 = l1_loss(image, gt_image) * opt.rgb_weight
    loss = (1.0 - opt.lambda_dssim) * Ll1 + opt.lambda_dssim * (1.0 - ssim(image, gt_image))
    reg, loss_dict = block_regularizers(params, occ_logits, camera, opt, mask=mask, ratio=ratio, scale_min=scale_min,
                                        use_reg=use_reg, generator=generator, near_far=near_far)
    return Ll1, loss + reg, loss_dict
