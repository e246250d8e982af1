"""Render-based block fitting (the plain optimiser of the block stage) and the
synthetic block scenes it is measured on.

A scene is K superquadric blocks (K in [4, 16]): exp(sq_s) ~ U[0.1, 0.5] per
axis, eps ~ U[0.2, 1.8], a uniform random rotation, a translation in the unit
ball and a colour per block.  Its surface points (for the initialisers) are
block-layer surfels with N(0, 0.005^2) noise; its target images come from
rasterising the same blocks from cameras on the radius-3 sphere.

fit_blocks optimises block parameters, occupancy and per-block colours
against the target images with the block stage's Adam groups
(games/block_mesh_splatting/scene/block_gaussian_model.py:429-442: rotation,
scale and translation 5e-3, eps and occupancy 1e-3, eps 1e-15) and the image
term of train.py ((1 - 0.2) L1 + 0.2 (1 - SSIM)), every view rasterised in one
batched call and the block layer run by the fused HIP kernel.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .batch import BatchConfig, BatchRasterizer
from .blocks import BlockGrid, generate_surfels, make_grid
from .camera import random_sphere_camera
from .losses import photometric


def logit(p):
    return math.log(p / (1.0 - p))


def random_block_params(rng: np.random.Generator, K: int):
    """K x 12 block parameters (sq_s, sq_r, sq_t, sq_eps)."""
    sq_s = np.log(rng.uniform(0.1, 0.5, size=(K, 3)))
    q = rng.normal(size=(K, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q *= np.where(q[:, :1] < 0, -1.0, 1.0)
    t = rng.normal(size=(K, 3))
    t = t / np.linalg.norm(t, axis=1, keepdims=True) * rng.uniform(0, 1, size=(K, 1)) ** (1 / 3)
    e = rng.uniform(0.2, 1.8, size=(K, 2))
    u = (e - 0.1) / 1.8
    sq_eps = np.log(u / (1 - u))
    return np.concatenate([sq_s, q, t, sq_eps], axis=1).astype(np.float32)


@dataclass
class BlockScene:
    params: torch.Tensor   # K x 12
    occ: torch.Tensor      # K
    colors: torch.Tensor   # K x 3
    points: torch.Tensor   # N x 3
    cameras: list
    targets: torch.Tensor  # flat batched colour images (the BatchRasterizer layout)
    grid: BlockGrid
    bg: torch.Tensor


def render_blocks(br: BatchRasterizer, params, occ, colors, grid: BlockGrid):
    """Flat colour buffer of all views (differentiable in all three inputs)."""
    means, scales, rots, opac = generate_surfels(params, occ, grid)
    per = grid.surfels_per_block()
    col = colors.repeat_interleave(per, dim=0)
    out_c, _, _ = br(means3D=means, opacities=opac, scales=scales, rotations=rots, colors_precomp=col)
    return torch.cat([c.reshape(-1) for c in out_c])


def random_block_scene(seed: int, n_min: int = 4, n_max: int = 16, n_points: int = 20000, views: int = 6,
                       width: int = 128, height: int = 96, grid_n: int = 24, device="cuda") -> BlockScene:
    rng = np.random.default_rng(seed)
    K = int(rng.integers(n_min, n_max + 1))
    dev = torch.device(device)
    params = torch.tensor(random_block_params(rng, K), device=dev)
    occ = torch.full((K,), logit(0.95), device=dev)
    colors = torch.tensor(rng.uniform(0.15, 0.95, size=(K, 3)).astype(np.float32), device=dev)
    grid = make_grid(grid_n, grid_n, device=dev)
    with torch.no_grad():
        means, _, _, _ = generate_surfels(params, occ, make_grid(64, 64, device=dev))
        pick = torch.tensor(rng.integers(0, means.shape[0], size=n_points), device=dev)
        noise = torch.tensor(rng.normal(0.0, 0.005, size=(n_points, 3)).astype(np.float32), device=dev)
        points = means[pick] + noise
    cams = [random_sphere_camera(seed * 97 + v, 3.0, width, height).to(dev) for v in range(views)]
    bg = torch.zeros(3, device=dev)
    br = BatchRasterizer(cams, bg, 0, BatchConfig(tile=16))
    with torch.no_grad():
        targets = render_blocks(br, params, occ, colors, grid).detach()
    return BlockScene(params, occ, colors, points, cams, targets, grid, bg)


@dataclass
class FitResult:
    params: torch.Tensor
    occ: torch.Tensor
    colors: torch.Tensor
    losses: list


def fit_blocks(scene: BlockScene, init_params: torch.Tensor, init_occ: torch.Tensor, iters: int = 200,
               br: BatchRasterizer | None = None) -> FitResult:
    """The plain optimiser: Adam on the block parameters through the renderer."""
    dev = init_params.device
    B = init_params.shape[0]
    sq_s = init_params[:, 0:3].clone().requires_grad_(True)
    sq_r = init_params[:, 3:7].clone().requires_grad_(True)
    sq_t = init_params[:, 7:10].clone().requires_grad_(True)
    sq_eps = init_params[:, 10:12].clone().requires_grad_(True)
    occ = init_occ.clone().requires_grad_(True)
    col_logit = torch.zeros((B, 3), device=dev, requires_grad=True)  # sigmoid(0) = 0.5 grey
    opt = torch.optim.Adam([
        {"params": [sq_r], "lr": 5e-3}, {"params": [sq_s], "lr": 5e-3}, {"params": [sq_t], "lr": 5e-3},
        {"params": [sq_eps], "lr": 1e-3}, {"params": [occ], "lr": 1e-3},
        {"params": [col_logit], "lr": 2.5e-2},
    ], lr=0.0, eps=1e-15)
    br = br or BatchRasterizer(scene.cameras, scene.bg, 0, BatchConfig(tile=16))
    losses = []
    for _ in range(iters):
        params = torch.cat([sq_s, sq_r, sq_t, sq_eps], dim=1)
        img = render_blocks(br, params, occ, torch.sigmoid(col_logit), scene.grid)
        loss = _image_loss(br, img, scene.targets)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
    params = torch.cat([sq_s, sq_r, sq_t, sq_eps], dim=1).detach()
    return FitResult(params, occ.detach(), torch.sigmoid(col_logit).detach(), losses)


def _image_loss(br: BatchRasterizer, img: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    total = 0.0
    for a, b in zip(br.split(img, 3), br.split(target, 3)):
        total = total + photometric(a, b)
    return total / br.V


def iterations_to_reach(losses, level: float) -> int:
    """First iteration (1-based) whose loss is <= level; len + 1 when none is."""
    for i, v in enumerate(losses):
        if v <= level:
            return i + 1
    return len(losses) + 1
