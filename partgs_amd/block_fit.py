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
scale and translation 5e-3, eps and occupancy 1e-3, eps 1e-15) and the
objective of summarize_losses (:584-648): the image term of train.py
((1 - 0.2) L1 + 0.2 (1 - SSIM)) over every view, rasterised in one batched
call, plus (with block_losses) parsimony, overlap and the three ray terms
against one view's object mask per iteration (views in turn, the mask being
the target's alpha > 0.5), the block layer and the superquadric distances run
by the fused HIP kernels.  The recorded curve is the image term, the loss
the comparison of initialisations is judged by.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .batch import BatchConfig, BatchRasterizer
from .blocks import BlockGrid, generate_surfels, make_grid
from .camera import random_sphere_camera
from .block_losses import BlockLossWeights, block_regularizers
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
    masks: list = None     # per view: H x W object mask (target alpha > 0.5)


def render_blocks(br: BatchRasterizer, params, occ, colors, grid: BlockGrid):
    """Flat colour buffer of all views (differentiable in all three inputs)."""
    means, scales, rots, opac = generate_surfels(params, occ, grid)
    per = grid.surfels_per_block()
    col = colors.repeat_interleave(per, dim=0)
    out_c, _, _ = br(means3D=means, opacities=opac, scales=scales, rotations=rots, colors_precomp=col)
    return torch.cat([c.reshape(-1) for c in out_c])


def block_scene_points(seed: int, n_min: int = 4, n_max: int = 16, n_points: int = 20000, device="cuda"):
    """The surface points of random_block_scene(seed) alone (the same draws,
    no cameras or target images)."""
    rng = np.random.default_rng(seed)
    K = int(rng.integers(n_min, n_max + 1))
    dev = torch.device(device)
    params = torch.tensor(random_block_params(rng, K), device=dev)
    occ = torch.full((K,), logit(0.95), device=dev)
    rng.uniform(0.15, 0.95, size=(K, 3))  # the colours' draws
    with torch.no_grad():
        means, _, _, _ = generate_surfels(params, occ, make_grid(64, 64, device=dev))
        pick = torch.tensor(rng.integers(0, means.shape[0], size=n_points), device=dev)
        noise = torch.tensor(rng.normal(0.0, 0.005, size=(n_points, 3)).astype(np.float32), device=dev)
        return means[pick] + noise


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
        means, scales, rots, opac = generate_surfels(params, occ, grid)
        col = colors.repeat_interleave(grid.surfels_per_block(), dim=0)
        out_c, out_o, _ = br(means3D=means, opacities=opac, scales=scales, rotations=rots, colors_precomp=col)
        targets = torch.cat([c.reshape(-1) for c in out_c]).detach()
        masks = [(o[1] > 0.5).float().contiguous() for o in out_o]
    return BlockScene(params, occ, colors, points, cams, targets, grid, bg, masks)


@dataclass
class FitResult:
    params: torch.Tensor
    occ: torch.Tensor
    colors: torch.Tensor
    losses: list


def fit_blocks(scene: BlockScene, init_params: torch.Tensor, init_occ: torch.Tensor, iters: int = 200,
               br: BatchRasterizer | None = None, block_losses: BlockLossWeights | None = None,
               near_far=(1.0, 5.0), seed: int = 0) -> FitResult:
    """The plain optimiser: Adam on the block parameters through the renderer.
    FitResult.losses holds the image term per iteration."""
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
    target = _stack_views(br, scene.targets)
    gen = torch.Generator(device=dev).manual_seed(seed) if block_losses is not None else None
    losses = []
    for it in range(iters):
        params = torch.cat([sq_s, sq_r, sq_t, sq_eps], dim=1)
        img = _stack_views(br, render_blocks(br, params, occ, torch.sigmoid(col_logit), scene.grid))
        image_loss = photometric(img, target)
        loss = image_loss
        if block_losses is not None:
            v = it % br.V
            mask = scene.masks[v] if scene.masks is not None else None
            reg, _ = block_regularizers(params, occ, scene.cameras[v], block_losses, mask=mask,
                                        generator=gen, near_far=near_far)
            loss = loss + reg
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        losses.append(image_loss.detach())
    params = torch.cat([sq_s, sq_r, sq_t, sq_eps], dim=1).detach()
    curve = torch.stack(losses).cpu().tolist() if losses else []  # one synchronisation per fit
    return FitResult(params, occ.detach(), torch.sigmoid(col_logit).detach(), curve)


def _stack_views(br: BatchRasterizer, flat: torch.Tensor) -> torch.Tensor:
    """V x 3 x H x W from the flat batched colour layout (views of one size)."""
    return torch.stack(list(br.split(flat, 3)))


def _image_loss(br: BatchRasterizer, img: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    return photometric(_stack_views(br, img), _stack_views(br, target))


def iterations_to_reach(losses, level: float) -> int:
    """First iteration (1-based) whose loss is <= level; len + 1 when none is."""
    for i, v in enumerate(losses):
        if v <= level:
            return i + 1
    return len(losses) + 1
