"""Image losses of the block stage: L1 and SSIM (utils/loss_utils.py:6-55),
restated in torch (a Gaussian 11 x 11 window with sigma 1.5, C1 = 0.01^2,
C2 = 0.03^2, zero padding)."""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F


def l1_loss(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    return (a - b).abs().mean()


_WINDOWS: dict = {}


def _window(channels: int, device, dtype, size: int = 11, sigma: float = 1.5):
    key = (channels, str(device), dtype, size)
    w = _WINDOWS.get(key)
    if w is None:
        g = torch.tensor([math.exp(-((x - size // 2) ** 2) / (2 * sigma ** 2)) for x in range(size)])
        g = g / g.sum()
        w2 = (g[:, None] @ g[None, :]).to(device=device, dtype=dtype)
        w = w2.expand(channels, 1, size, size).contiguous()
        _WINDOWS[key] = w
    return w


def ssim(img1: torch.Tensor, img2: torch.Tensor, window_size: int = 11) -> torch.Tensor:
    """Mean SSIM of two (C, H, W) or (B, C, H, W) images."""
    if img1.dim() == 3:
        img1, img2 = img1[None], img2[None]
    c = img1.shape[-3]
    w = _window(c, img1.device, img1.dtype, window_size)
    pad = window_size // 2
    mu1 = F.conv2d(img1, w, padding=pad, groups=c)
    mu2 = F.conv2d(img2, w, padding=pad, groups=c)
    mu1_sq, mu2_sq, mu12 = mu1 * mu1, mu2 * mu2, mu1 * mu2
    s1 = F.conv2d(img1 * img1, w, padding=pad, groups=c) - mu1_sq
    s2 = F.conv2d(img2 * img2, w, padding=pad, groups=c) - mu2_sq
    s12 = F.conv2d(img1 * img2, w, padding=pad, groups=c) - mu12
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu12 + C1) * (2 * s12 + C2)) / ((mu1_sq + mu2_sq + C1) * (s1 + s2 + C2))
    return m.mean()


def photometric(render: torch.Tensor, target: torch.Tensor, lambda_dssim: float = 0.2) -> torch.Tensor:
    """(1 - lambda) L1 + lambda (1 - SSIM), train.py's image term."""
    return (1.0 - lambda_dssim) * l1_loss(render, target) + lambda_dssim * (1.0 - ssim(render, target))
