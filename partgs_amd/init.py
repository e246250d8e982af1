"""The plain block initialisation (the fallback and yardstick of any learned one).

Restates readColmapBlockMeshInfo's block set-up
(games/block_mesh_splatting/scene/dataset_readers.py:487-503), cal_cluster_label
(games/block_mesh_splatting/utils/superquadric.py:68-90), fit_superquadric
(superquadric.py:112-120) and the parameter defaults of create_from_pcd
(games/block_mesh_splatting/scene/block_gaussian_model.py:146-166):

1. statistical outlier removal over 20 neighbours with std ratio 0.5 (open3d's
   ``remove_statistical_outlier``: drop points whose mean distance to their
   neighbours exceeds mean + ratio * std of that statistic);
2. k-means with k = number of blocks on the points min-max normalised to
   [-1, 1] per axis (scikit-learn's ``KMeans``: k-means++ seeding, Lloyd
   iterations to a centre-shift tolerance);
3. per cluster: sq_s = log(ratio) x 3, sq_r = (1, 0, 0, 0), sq_t = cluster
   mean + clamp(N(0,1)/4, -0.5, 0.5) * (scene extent per axis), sq_eps = 0,
   occupancy logit = logit(0.7).

Everything runs on the device of the input points; the k-nearest-neighbour
distances use chunked ``torch.cdist``.  scikit-learn and open3d are not
dependencies: their published algorithms are restated here, and their random
streams are not reproduced (the reference seeds neither).
"""
from __future__ import annotations

import math

import torch


def knn_mean_distance(points: torch.Tensor, k: int, chunk: int = 4096) -> torch.Tensor:
    """Mean Euclidean distance of every point to its k nearest other points."""
    n = points.shape[0]
    k = min(k, n - 1)
    if k <= 0:
        return torch.zeros(n, dtype=points.dtype, device=points.device)
    out = torch.empty(n, dtype=points.dtype, device=points.device)
    for s in range(0, n, chunk):
        d = torch.cdist(points[s:s + chunk], points)
        # k + 1 smallest include the point itself at distance 0
        out[s:s + chunk] = d.topk(k + 1, dim=1, largest=False).values[:, 1:].mean(dim=1)
    return out


def remove_statistical_outlier(points: torch.Tensor, nb_neighbors: int = 20, std_ratio: float = 0.5):
    """(inliers, mask): keep points whose mean neighbour distance <= mean + std_ratio * std."""
    if points.shape[0] <= 1:
        return points, torch.ones(points.shape[0], dtype=torch.bool, device=points.device)
    md = knn_mean_distance(points, nb_neighbors)
    mean = md.mean()
    std = md.std(unbiased=True) if md.numel() > 1 else torch.zeros((), device=md.device)
    keep = md <= mean + std_ratio * std
    return points[keep], keep


def _kmeanspp(x: torch.Tensor, k: int, gen: torch.Generator) -> torch.Tensor:
    """k-means++ seeding with 2 + floor(ln k) local trials per centre (scikit-learn's greedy variant)."""
    n = x.shape[0]
    trials = 2 + int(math.log(k))
    first = int(torch.randint(n, (1,), generator=gen, device="cpu"))
    centres = [x[first]]
    d2 = ((x - x[first]) ** 2).sum(1)
    for _ in range(1, k):
        pot = d2.sum()
        r = torch.rand(trials, generator=gen, device="cpu").to(x.device) * pot
        cand = torch.searchsorted(torch.cumsum(d2, 0), r).clamp_(max=n - 1)
        cd2 = torch.minimum(d2[None, :], ((x[None, :, :] - x[cand][:, None, :]) ** 2).sum(-1))
        best = int(torch.argmin(cd2.sum(1)))
        centres.append(x[cand[best]])
        d2 = cd2[best]
    return torch.stack(centres)


def kmeans(x: torch.Tensor, k: int, seed: int = 0, max_iter: int = 300, tol: float = 1e-4):
    """(labels, centres) of Lloyd's k-means with k-means++ seeding.

    tol is relative to the mean per-axis variance of x, as in scikit-learn.
    """
    gen = torch.Generator(device="cpu").manual_seed(seed)
    c = _kmeanspp(x, k, gen)
    tol_abs = tol * float(x.var(dim=0).mean())
    labels = torch.zeros(x.shape[0], dtype=torch.long, device=x.device)
    for _ in range(max_iter):
        labels = torch.cdist(x, c).argmin(dim=1)
        sums = torch.zeros_like(c).index_add_(0, labels, x)
        cnt = torch.bincount(labels, minlength=k).to(x.dtype)
        new_c = torch.where(cnt[:, None] > 0, sums / cnt.clamp(min=1)[:, None], c)
        shift = float(((new_c - c) ** 2).sum())
        c = new_c
        if shift <= tol_abs:
            break
    labels = torch.cdist(x, c).argmin(dim=1)
    return labels, c


def cluster_labels(points: torch.Tensor, num_k: int, seed: int = 0) -> torch.Tensor:
    """cal_cluster_label with num_k > 1: k-means on points normalised to [-1, 1] per axis."""
    lo = points.min(dim=0).values
    hi = points.max(dim=0).values
    x = 2 * (points - lo) / (hi - lo) - 1
    labels, _ = kmeans(x, num_k, seed=seed)
    return labels


def plain_block_init(points: torch.Tensor, n_blocks: int = 8, ratio_block_scene: float = 0.25, seed: int = 0,
                     remove_outliers: bool = True):
    """Block parameters (B x 12: sq_s, sq_r, sq_t, sq_eps) and occupancy logits (B)
    of the plain initialisation from a point cloud."""
    pts = points.float()
    if remove_outliers:
        pts, _ = remove_statistical_outlier(pts, nb_neighbors=20, std_ratio=0.5)
    labels = cluster_labels(pts, n_blocks, seed=seed)
    range_T = pts.max(dim=0).values - pts.min(dim=0).values
    gen = torch.Generator(device="cpu").manual_seed(seed + 1)
    params = torch.zeros((n_blocks, 12), dtype=torch.float32, device=pts.device)
    params[:, 0:3] = math.log(ratio_block_scene)
    params[:, 3] = 1.0
    for i in range(n_blocks):
        sel = pts[labels == i]
        mean = sel.mean(dim=0) if sel.shape[0] > 0 else torch.zeros(3, device=pts.device)
        jitter = (torch.randn(3, generator=gen) / 4).clamp(-0.5, 0.5).to(pts.device)
        params[i, 7:10] = mean + jitter * range_T
    occ = torch.full((n_blocks,), math.log(0.7 / 0.3), dtype=torch.float32, device=pts.device)
    return params, occ
